// host.cpp -- option handling, tau table and error state of libgbm3d.
//
// setBM3Dopts / checkBM3Dopts / the profile rule of GBM3D, restated:
//   utilities/setBM3Dopts.m:12-86, utilities/checkBM3Dopts.m:10-24,
//   source/GBM3D.m:15-19, utilities/getTau.m:9-51.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "internal.hpp"

namespace gbm3d {

namespace {
thread_local std::string g_last_error;

void put_spins(int32_t* dst, int32_t* n, std::initializer_list<int> v) {
  *n = 0;
  for (int x : v) dst[(*n)++] = x;
}
}  // namespace

void set_error(const std::string& msg) { g_last_error = msg; }
void clear_error() { g_last_error.clear(); }
int fail(int code, const std::string& msg) {
  set_error(msg);
  return code;
}

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(GBM3D_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
  return GBM3D_OK;
}

static void clear_opts(gbm3d_opts* o) {
  std::memset(o, 0, sizeof(*o));
  o->numMax = o->numMin = o->levels = o->cycleSpin = GBM3D_UNSET;
  o->nMatchSpins = o->nMatchSpinsWie = GBM3D_UNSET;
  o->tauMode = o->blockSize = o->blockSizeWie = o->matchSize = GBM3D_UNSET;
  o->Wiener = o->filtType = GBM3D_UNSET;
}

static void set_profile(const char* profile, gbm3d_opts* o) {
  std::string p = profile ? profile : "default";
  if (p != "default" && p != "fast" && p != "superFast" && p != "accuracy") p = "default";
  clear_opts(o);
  o->numMax = 16;
  o->numMin = 16;
  o->levels = 3;
  o->cycleSpin = 2;
  o->tauMode = 1;
  o->blockSize = p == "superFast" ? 32 : 16;
  o->blockSizeWie = 8;
  o->matchSize = 32;
  o->Wiener = p == "superFast" ? 0 : 1;
  o->filtType = GBM3D_FILT_HT;
  if (p == "default") {
    put_spins(o->matchSpins, &o->nMatchSpins, {0, 4});
    put_spins(o->matchSpinsWie, &o->nMatchSpinsWie, {0, 4});
  } else if (p == "accuracy") {
    put_spins(o->matchSpins, &o->nMatchSpins, {0, 4, 6});
    put_spins(o->matchSpinsWie, &o->nMatchSpinsWie, {0, 4, 6});
  } else {  // fast, superFast
    put_spins(o->matchSpins, &o->nMatchSpins, {0, 4, 6});
    put_spins(o->matchSpinsWie, &o->nMatchSpinsWie, {0});
  }
}

static void check_defaults(gbm3d_opts* o) {
  auto def = [](int32_t& f, int v) {
    if (f == GBM3D_UNSET) f = v;
  };
  def(o->numMax, 16);
  def(o->numMin, 16);
  def(o->levels, 3);
  def(o->tauMode, 1);
  def(o->filtType, GBM3D_FILT_HT);
  def(o->blockSize, 16);
  def(o->blockSizeWie, 8);
  def(o->matchSize, std::max(32, (int)o->blockSize));
  def(o->cycleSpin, 2);
  if (o->nMatchSpins == GBM3D_UNSET) {
    o->nMatchSpins = 2;
    o->matchSpins[0] = 0;
    o->matchSpins[1] = std::min(o->blockSize, o->blockSizeWie) / 2;
  }
  def(o->Wiener, 1);
}

int resolve(const gbm3d_opts* in, gbm3d_opts* out) {
  if (!out) return fail(GBM3D_ERR_ARG, "opts: NULL output");
  if (!in) {
    set_profile("default", out);
  } else if (in->profile[0]) {
    char name[17] = {0};
    std::memcpy(name, in->profile, 16);
    set_profile(name, out);
  } else {
    *out = *in;
  }
  check_defaults(out);
  return GBM3D_OK;
}

void tau_table(int levels, int mode, double sigma, double* tau) {
  for (int j = 0; j < 7; ++j)
    for (int i = 0; i < levels; ++i) {
      const double base = 3.3 - 0.3 * i;
      double t = 0.0;
      switch (mode) {
        case 1: t = base; break;
        case 2: t = j < 3 ? base : (j < 6 ? base + 0.5 : base + 1.0); break;
        case 3: t = j < 3 ? base : (j < 6 ? base + 0.25 : base + 1.5); break;
        case 4: t = j < 3 ? 2.9 : (j < 6 ? 2.9 + 0.5 : 2.9 + 1.0); break;
        default: t = 0.0; break;
      }
      tau[i + levels * j] = sigma * t;
    }
}

// getColoredTau (utilities/getColoredTau.m:9-57): for level i and band j,
// tau = 6 sqrt(sum(PSD .* |F|^2) / sum(|F|^2)), F the d1 x d2 DFT of the band's
// 2-D analysis filter (outer product of the 10-tap low / 2-tap high filters,
// nearest-neighbour upsampled by 2^(i-1)).  The filter is separable, so
// |F|^2 = |R(k1)|^2 |C(k2)|^2 and both sums factor per axis.
void colored_tau(int levels, const double* psd, int d1, int d2, double* tau) {
  const double Ca = kCa, Cb = kCb, Cc = kCc;
  const double lo[10] = {Ca, -Ca, -Cb, Cb, Cc, Cc, Cb, -Cb, -Ca, Ca};
  const double hi[10] = {0, 0, 0, 0, -Cc, Cc, 0, 0, 0, 0};
  // band j -> (rows filter, columns filter): Psi(:,:,j+1) = R'*C, 1 = low, 0 = high
  const int rlo[7] = {0, 1, 1, 0, 0, 1, 0}, clo[7] = {1, 0, 1, 0, 1, 0, 0};
  auto power = [](const double* f, int s, int n) {  // |DFT_n(kron(f, ones(s)))|^2 (fft2 pads or truncates)
    std::vector<double> p(n);
    const int L = std::min(10 * s, n);
    for (int k = 0; k < n; ++k) {
      double re = 0.0, im = 0.0;
      for (int t = 0; t < L; ++t) {
        const double a = -2.0 * M_PI * (double)(((long long)k * t) % n) / n;
        re += f[t / s] * std::cos(a);
        im += f[t / s] * std::sin(a);
      }
      p[k] = re * re + im * im;
    }
    return p;
  };
  for (int i = 0; i < levels; ++i) {
    const int s = 1 << i;
    const std::vector<double> Rl = power(lo, s, d1), Rh = power(hi, s, d1), Cl = power(lo, s, d2),
                              Ch = power(hi, s, d2);
    for (int j = 0; j < 7; ++j) {
      const std::vector<double>& R = rlo[j] ? Rl : Rh;
      const std::vector<double>& C = clo[j] ? Cl : Ch;
      double num = 0.0, sr = 0.0, sc = 0.0;
      for (int k2 = 0; k2 < d2; ++k2) {
        double col = 0.0;
        for (int k1 = 0; k1 < d1; ++k1) col += psd[k1 + (size_t)d1 * k2] * R[k1];
        num += col * C[k2];
      }
      for (double v : R) sr += v;
      for (double v : C) sc += v;
      tau[i + levels * j] = 6.0 * std::sqrt(num / (sr * sc));
    }
  }
}

// imresize(A, [out_len ...], 'bilinear') along one axis, as MATLAB's imresize
// builds it (the documented `contributions` construction): output sample i
// (1-based) sits at u = i/s + (1 - 1/s)/2 of the input, s = out/in; the
// triangle kernel, widened to 2/s and scaled by s when shrinking (antialiasing,
// on by default), weighs input samples floor(u - w/2) + (0 .. ceil(w) + 1),
// weights normalised to sum 1, indices past either end mirrored
// ([1:n, n:-1:1]).  Returns, per output sample, its (index, weight) pairs.
static std::vector<std::vector<std::pair<int, double>>> resize_weights(int in, int out) {
  const double s = (double)out / in;
  const bool shrink = s < 1.0;
  const double kw = shrink ? 2.0 / s : 2.0;
  auto tri = [](double x) { return (-1.0 <= x && x < 0.0) ? x + 1.0 : ((0.0 <= x && x <= 1.0) ? 1.0 - x : 0.0); };
  auto h = [&](double x) { return shrink ? s * tri(s * x) : tri(x); };
  std::vector<std::vector<std::pair<int, double>>> w(out);
  const int P = (int)std::ceil(kw) + 2;
  for (int i = 1; i <= out; ++i) {
    const double u = i / s + 0.5 * (1.0 - 1.0 / s);
    const long left = (long)std::floor(u - kw / 2.0);
    double sum = 0.0;
    std::vector<std::pair<int, double>> c;
    for (int p = 0; p < P; ++p) {
      const long idx = left + p;
      const double wt = h(u - (double)idx);
      // mirror into 1..in through the period-2n sequence [1:n, n:-1:1]
      long m = ((idx - 1) % (2L * in) + 2L * in) % (2L * in);
      const int src = (int)(m < in ? m : 2L * in - 1 - m);  // 0-based
      c.emplace_back(src, wt);
      sum += wt;
    }
    for (auto& e : c) e.second /= sum;
    w[i - 1] = std::move(c);
  }
  return w;
}

// getColoredWie (utilities/getColoredWie.m:6-19): the PSD smoothed by a
// Gaussian (makeGausPSF([d1 d2], 3): not part of the reference; here a
// circular Gaussian of standard deviation 3 px centred at the origin with
// unit sum, so that ifft2(ghat .* fft2(PSD)) is that circular convolution),
// resized to d x d as imresize(., [d d], 'bilinear') does, then
// sigmaWie(j, i) = sqrt(sum(PSD_d .* |fft2(idct2(e_ji))|^2)) / d.  idct2 of
// the unit impulse at (j, i) is the outer product of rows j and i of the
// orthonormal DCT-II matrix T, so |fft2(.)|^2 = A_j(u) A_i(v), A_k = |fft(T(k,:))|^2.
void colored_wie(const double* psd, int d1, int d2, int d, double* sigmaWie) {
  const double omega = 3.0;
  // 1-D circular Gaussians (the 2-D one is their product; unit sum each)
  auto gauss = [&](int n) {
    std::vector<double> g(n);
    double sum = 0.0;
    for (int x = 0; x < n; ++x) {
      const double dx = std::min(x, n - x);
      g[x] = std::exp(-dx * dx / (2.0 * omega * omega));
      sum += g[x];
    }
    for (double& v : g) v /= sum;
    return g;
  };
  const std::vector<double> g1 = gauss(d1), g2 = gauss(d2);
  // offsets -48..48 on long axes (taps past 16 standard deviations are below
  // 1e-55 of the centre one), every offset of the circle on the others
  auto taps = [](const std::vector<double>& g, int n) {
    std::vector<std::pair<int, double>> t;
    if (n <= 2 * 48 + 1) {
      for (int x = 0; x < n; ++x) t.emplace_back(x, g[x]);
    } else {
      for (int x = -48; x <= 48; ++x) t.emplace_back(x, g[(x + n) % n]);
    }
    return t;
  };
  const auto t1 = taps(g1, d1), t2 = taps(g2, d2);
  std::vector<double> a((size_t)d1 * d2), b((size_t)d1 * d2);
  for (int c = 0; c < d2; ++c)  // along dim 1
    for (int rr = 0; rr < d1; ++rr) {
      double acc = 0.0;
      for (const auto& [x, w] : t1) acc += w * psd[(size_t)(((rr - x) % d1 + d1) % d1) + (size_t)d1 * c];
      a[rr + (size_t)d1 * c] = acc;
    }
  for (int c = 0; c < d2; ++c)  // along dim 2
    for (int rr = 0; rr < d1; ++rr) {
      double acc = 0.0;
      for (const auto& [y, w] : t2) acc += w * a[rr + (size_t)d1 * (((c - y) % d2 + d2) % d2)];
      b[rr + (size_t)d1 * c] = acc;
    }
  // imresize to d x d: along dim 1 first (equal scale factors, or the smaller
  // one first, as imresize orders the passes), then dim 2
  const bool rows_first = (double)d / d1 <= (double)d / d2;
  std::vector<double> r((size_t)d * d);
  if (rows_first) {
    const auto w1 = resize_weights(d1, d), w2 = resize_weights(d2, d);
    std::vector<double> m1((size_t)d * d2);
    for (int c = 0; c < d2; ++c)
      for (int i = 0; i < d; ++i) {
        double acc = 0.0;
        for (const auto& [k, w] : w1[i]) acc += w * b[k + (size_t)d1 * c];
        m1[i + (size_t)d * c] = acc;
      }
    for (int j = 0; j < d; ++j)
      for (int i = 0; i < d; ++i) {
        double acc = 0.0;
        for (const auto& [k, w] : w2[j]) acc += w * m1[i + (size_t)d * k];
        r[i + (size_t)d * j] = acc;
      }
  } else {
    const auto w1 = resize_weights(d1, d), w2 = resize_weights(d2, d);
    std::vector<double> m2((size_t)d1 * d);
    for (int j = 0; j < d; ++j)
      for (int i = 0; i < d1; ++i) {
        double acc = 0.0;
        for (const auto& [k, w] : w2[j]) acc += w * b[i + (size_t)d1 * k];
        m2[i + (size_t)d1 * j] = acc;
      }
    for (int j = 0; j < d; ++j)
      for (int i = 0; i < d; ++i) {
        double acc = 0.0;
        for (const auto& [k, w] : w1[i]) acc += w * m2[k + (size_t)d1 * j];
        r[i + (size_t)d * j] = acc;
      }
  }
  // A_k(u) = |sum_p T(k, p) e^{-2 pi i u p / d}|^2, T(k, p) = c_k cos(pi (2p + 1) k / (2d))
  std::vector<double> A((size_t)d * d);
  for (int k = 0; k < d; ++k) {
    const double ck = k == 0 ? std::sqrt(1.0 / d) : std::sqrt(2.0 / d);
    for (int u = 0; u < d; ++u) {
      double re = 0.0, im = 0.0;
      for (int p = 0; p < d; ++p) {
        const double t = ck * std::cos(M_PI * (2 * p + 1) * k / (2.0 * d));
        const double ang = -2.0 * M_PI * (double)((u * p) % d) / d;
        re += t * std::cos(ang);
        im += t * std::sin(ang);
      }
      A[k + (size_t)d * u] = re * re + im * im;
    }
  }
  for (int i = 0; i < d; ++i)
    for (int j = 0; j < d; ++j) {
      double sum = 0.0;
      for (int v = 0; v < d; ++v) {
        double col = 0.0;
        for (int u = 0; u < d; ++u) col += r[u + (size_t)d * v] * A[j + (size_t)d * u];
        sum += col * A[i + (size_t)d * v];
      }
      sigmaWie[j + (size_t)d * i] = std::sqrt(sum) / d;
    }
}

}  // namespace gbm3d

extern "C" {

int gbm3d_opts_clear(gbm3d_opts* opts) {
  if (!opts) return gbm3d::fail(GBM3D_ERR_ARG, "opts: NULL");
  gbm3d::clear_opts(opts);
  return GBM3D_OK;
}

int gbm3d_set_opts(const char* profile, gbm3d_opts* opts) {
  if (!opts) return gbm3d::fail(GBM3D_ERR_ARG, "opts: NULL");
  gbm3d::set_profile(profile, opts);
  return GBM3D_OK;
}

int gbm3d_check_opts(gbm3d_opts* opts) {
  if (!opts) return gbm3d::fail(GBM3D_ERR_ARG, "opts: NULL");
  gbm3d::check_defaults(opts);
  return GBM3D_OK;
}

int gbm3d_resolve_opts(const gbm3d_opts* in, gbm3d_opts* out) { return gbm3d::resolve(in, out); }

int gbm3d_get_tau(int levels, int mode, double sigma, double* tau) {
  if (levels < 1 || !tau) return gbm3d::fail(GBM3D_ERR_ARG, "get_tau: levels < 1 or NULL");
  gbm3d::tau_table(levels, mode, sigma, tau);
  return GBM3D_OK;
}

int gbm3d_get_colored_tau(int levels, const double* sigmaPSD, int d1, int d2, double* tau) {
  if (levels < 1 || !sigmaPSD || !tau || d1 < 1 || d2 < 1)
    return gbm3d::fail(GBM3D_ERR_ARG, "get_colored_tau: bad argument");
  gbm3d::colored_tau(levels, sigmaPSD, d1, d2, tau);
  return GBM3D_OK;
}

int gbm3d_get_colored_wie(const double* sigmaPSD, int d1, int d2, int d, double* sigmaWie) {
  if (!sigmaPSD || !sigmaWie || d1 < 1 || d2 < 1 || d < 1 || d > 64)
    return gbm3d::fail(GBM3D_ERR_ARG, "get_colored_wie: bad argument");
  gbm3d::colored_wie(sigmaPSD, d1, d2, d, sigmaWie);
  return GBM3D_OK;
}

const char* gbm3d_last_error(void) { return gbm3d::g_last_error.c_str(); }

}  // extern "C"
