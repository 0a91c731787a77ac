// host.cpp -- option handling, tau table and error state of libgbm3d.
//
// setBM3Dopts / checkBM3Dopts / the profile rule of GBM3D, restated:
//   utilities/setBM3Dopts.m:12-86, utilities/checkBM3Dopts.m:10-24,
//   source/GBM3D.m:15-19, utilities/getTau.m:9-51.
#include <algorithm>
#include <cstring>
#include <string>

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

const char* gbm3d_last_error(void) { return gbm3d::g_last_error.c_str(); }

}  // extern "C"
