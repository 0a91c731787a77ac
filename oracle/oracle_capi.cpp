// oracle_capi.cpp -- extern "C" entry points of the oracle (liboracle_gbm3d.so)
// for the Python test harness.  TEST INFRASTRUCTURE ONLY: never linked into the
// product library.  Every function returns 0 on success, a gbm3d_status code
// otherwise.  Layouts are the reference's (column-major), see gbm3d_oracle.hpp.
#include "gbm3d_oracle.hpp"

using namespace gbm3d_oracle;

namespace {
bool ok_sizes(int m, int n, int d, int d2) {
  return m > 0 && n > 0 && d > 0 && d2 > 0 && (d2 % 2) == 0 && m >= d2 && n >= d2 && m % d == 0 && n % d == 0;
}
template <class F>
int guarded(F f) {
  try {
    return f();
  } catch (...) {
    return GBM3D_ERR_NOMEM;
  }
}
void threads(int nt) {
#ifdef _OPENMP
  if (nt > 0) omp_set_num_threads(nt);
#else
  (void)nt;
#endif
}
}  // namespace

extern "C" {

int orc_set_threads(int nt) {
  threads(nt);
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

int orc_set_opts(const char* profile, gbm3d_opts* out) {
  set_opts(profile, out);
  return 0;
}
int orc_check_opts(gbm3d_opts* o) {
  check_opts(o);
  return 0;
}
int orc_resolve_opts(const gbm3d_opts* in, gbm3d_opts* out) {
  resolve_opts(in, out);
  return 0;
}
int orc_get_tau(int levels, int mode, double sigma, double* tau) {
  get_tau(levels, mode, sigma, tau);
  return 0;
}

int orc_mirror_pad_f64(const double* I, int m, int n, int M, int N, double* out) {
  if (M < m || N < n || M > 2 * m || N > 2 * n) return GBM3D_ERR_ARG;
  auto v = mirror_pad<double>(I, m, n, M, N);
  std::copy(v.begin(), v.end(), out);
  return 0;
}

int orc_search_pad_f64(const double* X, int m, int n, int d2, double* out) {
  if (m < d2 || n < d2) return GBM3D_ERR_ARG;
  auto v = search_pad<double>(X, m, n, d2);
  std::copy(v.begin(), v.end(), out);
  return 0;
}

#define MATCH1(SUF, T)                                                                               \
  int orc_match_##SUF(const T* I, int m, int n, int d, int d2, int K, uint32_t* S, T* Volume,      \
                      T* dist) {                                                                     \
    if (!ok_sizes(m, n, d, d2) || K < 1 || K > d2 * d2) return GBM3D_ERR_ARG;                     \
    return guarded([&] {                                                                             \
      match_step1<T>(I, m, n, d, d2, K, S, Volume, dist);                                            \
      return 0;                                                                                      \
    });                                                                                              \
  }                                                                                                  \
  int orc_match_tol_##SUF(const T* I, int m, int n, int d, int d2, int K, int Kmin, double sigma,    \
                          uint32_t* S, T* Volume) {                                                  \
    if (!ok_sizes(m, n, d, d2) || K < 1 || K > d2 * d2 || Kmin < 1 || Kmin > K) return GBM3D_ERR_ARG; \
    int st = 0;                                                                                      \
    const int g = guarded([&] {                                                                      \
      st = match_step1<T>(I, m, n, d, d2, K, S, Volume, nullptr, Kmin, sigma);                       \
      return 0;                                                                                      \
    });                                                                                              \
    return g != 0 ? g : st;                                                                          \
  }                                                                                                  \
  int orc_match_wie_##SUF(const T* I, const T* In, int m, int n, int d, int d2, int K, uint32_t* S, \
                          T* V1, T* V2, T* dist) {                                                   \
    if (!ok_sizes(m, n, d, d2) || K < 1 || K > d2 * d2) return GBM3D_ERR_ARG;                     \
    return guarded([&] {                                                                             \
      match_step2<T>(I, In, m, n, d, d2, K, S, V1, V2, dist);                                        \
      return 0;                                                                                      \
    });                                                                                              \
  }
MATCH1(f64, double)
MATCH1(f32, float)

#define HT(SUF, T)                                                                                   \
  int orc_ht_filter_##SUF(const T* V, int m, int n, int K, int levels, int cycleSpin,               \
                          const double* tau, T* Vavg, double* min_margin) {                          \
    if (levels < 1 || (m % (1 << levels)) || (n % (1 << levels)) || (K % (1 << levels)) ||          \
        cycleSpin < 1)                                                                               \
      return GBM3D_ERR_ARG;                                                                          \
    return guarded([&] {                                                                             \
      if (min_margin) *min_margin = 1e300;                                                           \
      ht_filter<T>(V, m, n, K, levels, cycleSpin, tau, Vavg, min_margin);                            \
      return 0;                                                                                      \
    });                                                                                              \
  }
HT(f64, double)
HT(f32, float)

#define MED(SUF, T)                                                                                  \
  int orc_median3_##SUF(const T* V, int m, int n, int K, T* Vavg) {                                 \
    if (m < 1 || n < 1 || K < 1) return GBM3D_ERR_ARG;                                              \
    return guarded([&] {                                                                             \
      median3<T>(V, m, n, K, Vavg);                                                                  \
      return 0;                                                                                      \
    });                                                                                              \
  }
MED(f64, double)
MED(f32, float)

#define AGG(SUF, T)                                                                                  \
  int orc_agg_ht_##SUF(const T* Vavg, const uint32_t* S, int m, int n, int d, int d2, int K, T* U,  \
                       T* W, int gather) {                                                           \
    if (!ok_sizes(m, n, d, d2)) return GBM3D_ERR_ARG;                                                \
    return guarded([&] {                                                                             \
      aggregate_ht<T>(Vavg, S, m, n, d, d2, K, U, W, gather != 0);                                   \
      return 0;                                                                                      \
    });                                                                                              \
  }                                                                                                  \
  int orc_agg_wie_##SUF(const T* V2f, const T* gain2, const uint32_t* S, int m, int n, int d, int d2, \
                        int K, T* U, T* W, int gather) {                                             \
    if (!ok_sizes(m, n, d, d2)) return GBM3D_ERR_ARG;                                                \
    return guarded([&] {                                                                             \
      aggregate_wie<T>(V2f, gain2, S, m, n, d, d2, K, U, W, gather != 0);                            \
      return 0;                                                                                      \
    });                                                                                              \
  }                                                                                                  \
  int orc_agg_weights_##SUF(const T* Vavg, int m, int n, int d, int K, T* w) {                       \
    if (m % d || n % d) return GBM3D_ERR_ARG;                                                        \
    auto v = agg1_weights<T>(Vavg, m, n, d, K);                                                      \
    std::copy(v.begin(), v.end(), w);                                                                \
    return 0;                                                                                        \
  }                                                                                                  \
  int orc_wiener_##SUF(const T* V1, const T* V2, int d, int K, long G, double sigma, T* V2f,        \
                       T* gain2) {                                                                   \
    if (d < 1 || K < 4 || (K % 4)) return GBM3D_ERR_ARG;                                             \
    return guarded([&] {                                                                             \
      wiener<T>(V1, V2, d, K, G, sigma, V2f, gain2);                                                 \
      return 0;                                                                                      \
    });                                                                                              \
  }                                                                                                  \
  int orc_wiener01_##SUF(const T* V1, const T* V2, int d, int K, long G, double sigma, T* V2f,      \
                         T* gain2) {                                                                 \
    if (d < 1 || K < 4 || (K % 4)) return GBM3D_ERR_ARG;                                             \
    return guarded([&] {                                                                             \
      wiener<T>(V1, V2, d, K, G, sigma, V2f, gain2, nullptr, true);                                  \
      return 0;                                                                                      \
    });                                                                                              \
  }                                                                                                  \
  int orc_wiener_tab_##SUF(const T* V1, const T* V2, int d, int K, long G, const double* sigWie,     \
                           T* V2f, T* gain2) {                                                       \
    if (d < 1 || K < 4 || (K % 4) || !sigWie) return GBM3D_ERR_ARG;                                  \
    return guarded([&] {                                                                             \
      wiener<T>(V1, V2, d, K, G, 1.0, V2f, gain2, sigWie);                                           \
      return 0;                                                                                      \
    });                                                                                              \
  }
AGG(f64, double)
AGG(f32, float)

int orc_haar_dec_f64(double* x, int R, int levels) {
  haar_dec<double>(x, R, levels);
  return 0;
}
int orc_haar_rec_f64(double* x, int R, int levels) {
  haar_rec<double>(x, R, levels);
  return 0;
}

// One level of the 3-D analysis (8 subbands, b = dim1_hi | dim2_hi<<1 | z_hi<<2),
// each (p/2) x (q/2) x (r/2), written consecutively to C.
int orc_wav_dec_level_f64(const double* X, int p, int q, int r, double* C) {
  if (p % 2 || q % 2 || r % 2) return GBM3D_ERR_ARG;
  Vol<double> V(p, q, r);
  std::copy(X, X + (size_t)p * q * r, V.v.begin());
  Vol<double> B[8];
  analysis_level<double>(V, B, Bank<double>());
  size_t sz = (size_t)p * q * r / 8;
  for (int b = 0; b < 8; ++b) std::copy(B[b].v.begin(), B[b].v.end(), C + sz * b);
  return 0;
}
int orc_wav_rec_level_f64(const double* C, int p, int q, int r, double* X) {
  Vol<double> B[8];
  size_t sz = (size_t)p * q * r / 8;
  for (int b = 0; b < 8; ++b) {
    B[b] = Vol<double>(p / 2, q / 2, r / 2);
    std::copy(C + sz * b, C + sz * (b + 1), B[b].v.begin());
  }
  Vol<double> Y = synthesis_level<double>(B, Bank<double>());
  std::copy(Y.v.begin(), Y.v.end(), X);
  return 0;
}

int orc_gbm3d_f64(const double* I, int m, int n, double sigma, const gbm3d_opts* resolved, double* U,
                  double* U1) {
  if (!I || !U || !resolved || m < 1 || n < 1) return GBM3D_ERR_ARG;
  return guarded([&] { return gbm3d<double>(I, m, n, sigma, *resolved, U, U1, false); });
}
int orc_gbm3d_f32(const float* I, int m, int n, float sigma, const gbm3d_opts* resolved, float* U,
                  float* U1, int nthreads) {
  if (!I || !U || !resolved || m < 1 || n < 1) return GBM3D_ERR_ARG;
  threads(nthreads);
  return guarded([&] { return gbm3d<float>(I, m, n, (double)sigma, *resolved, U, U1, true); });
}


void orc_sym4_taps(double* lo, double* hi) {
  for (int k = 0; k < 8; ++k) {
    lo[k] = kSym4Lo[k];
    hi[k] = kSym4Hi[k];
  }
}

#define PREFILT(SUF, T)                                                                              \
  int orc_wav2_prefilter_##SUF(const T* I, int m, int n, double tau, T* out) {                       \
    if (!I || !out || m % 8 || n % 8) return GBM3D_ERR_ARG;                                          \
    return guarded([&] {                                                                             \
      const std::vector<T> r = wav2_prefilter<T>(I, m, n, tau);                                      \
      std::copy(r.begin(), r.end(), out);                                                            \
      return 0;                                                                                      \
    });                                                                                              \
  }
PREFILT(f64, double)
PREFILT(f32, float)

#define GMATCH(SUF, T)                                                                               \
  int orc_match_global_##SUF(const T* I, int m, int n, double sigma, int d, int numMax, int numMin,     \
                             uint32_t* S, uint32_t* counts, T* filtered) {                             \
    if (!I || !S) return GBM3D_ERR_ARG;                                                              \
    int st = GBM3D_OK;                                                                               \
    const int g = guarded([&] {                                                                      \
      st = match_global<T>(I, m, n, sigma, d, numMax, numMin, S, counts, filtered);                  \
      return 0;                                                                                      \
    });                                                                                              \
    return g != 0 ? g : st;                                                                          \
  }
GMATCH(f64, double)
GMATCH(f32, float)

long orc_ht_near_mask_f32(const float* V, int m, int n, int K, int levels, int cycleSpin, const double* tau,
                          double tol, uint8_t* vmask) {
  if (!V || !tau || !vmask || levels < 1 || (m % (1 << levels)) || (n % (1 << levels)) || (K % (1 << levels)) ||
      cycleSpin < 1)
    return -1;
  long r = -1;
  const int st = guarded([&] {
    r = ht_near_mask<float>(V, m, n, K, levels, cycleSpin, tau, tol, vmask);
    return 0;
  });
  return st == 0 ? r : -1;
}

long orc_u1_near_mask_f32(const float* I, int m, int n, float sigma, const gbm3d_opts* resolved, double tol,
                          uint8_t* mask) {
  if (!I || !resolved || !mask || m < 1 || n < 1) return -1;
  long r = -1;
  const int st = guarded([&] {
    r = u1_near_mask<float>(I, m, n, (double)sigma, *resolved, tol, mask);
    return 0;
  });
  return st == 0 ? r : -1;
}

int orc_colored_tau(int levels, const double* psd, int d1, int d2, double* tau) {
  if (levels < 1 || !psd || !tau || d1 < 1 || d2 < 1) return GBM3D_ERR_ARG;
  colored_tau(levels, psd, d1, d2, tau);
  return 0;
}

int orc_gbm3d_colored_f64(const double* I, int m, int n, const double* psd, int d1, int d2, const double* sigWie,
                          const gbm3d_opts* resolved, double* U, double* U1) {
  if (!I || !U || !resolved || !psd) return GBM3D_ERR_ARG;
  return guarded([&] {
    std::vector<double> tau((size_t)resolved->levels * 7);
    colored_tau(resolved->levels, psd, d1, d2, tau.data());
    // sigma enters only the tolerance branch there: mean(sqrt(sigma(:))) (matchBlocksCPU.m:21)
    double ss = 0.0;
    for (int e = 0; e < d1 * d2; ++e) ss += std::sqrt(psd[e]);
    return gbm3d<double>(I, m, n, ss / (d1 * d2), *resolved, U, U1, true, tau.data(), sigWie);
  });
}
int orc_gbm3d_colored_f32(const float* I, int m, int n, const double* psd, int d1, int d2, const double* sigWie,
                          const gbm3d_opts* resolved, float* U, float* U1) {
  if (!I || !U || !resolved || !psd) return GBM3D_ERR_ARG;
  return guarded([&] {
    std::vector<double> tau((size_t)resolved->levels * 7);
    colored_tau(resolved->levels, psd, d1, d2, tau.data());
    // sigma enters only the tolerance branch there: mean(sqrt(sigma(:))) (matchBlocksCPU.m:21)
    double ss = 0.0;
    for (int e = 0; e < d1 * d2; ++e) ss += std::sqrt(psd[e]);
    return gbm3d<float>(I, m, n, ss / (d1 * d2), *resolved, U, U1, true, tau.data(), sigWie);
  });
}

int orc_distribution(int m, int n, const int32_t* spins, int nspins, int k, int32_t* rows, int max_rows) {
  std::vector<TileCase> R = distribution(m, n, spins, nspins, k);
  if ((int)R.size() > max_rows) return -(int)R.size();
  for (size_t i = 0; i < R.size(); ++i) {
    rows[5 * i + 0] = R[i].y1;
    rows[5 * i + 1] = R[i].y2;
    rows[5 * i + 2] = R[i].x1;
    rows[5 * i + 3] = R[i].x2;
    rows[5 * i + 4] = R[i].spin;
  }
  return (int)R.size();
}

int orc_interp_window(int a, int b, int k, double* w) {
  if (a < 1 || b < 1 || k < 2) return GBM3D_ERR_ARG;
  std::vector<double> v = interp_window(a, b, k);
  std::copy(v.begin(), v.end(), w);
  return 0;
}

int orc_gbm3d_distributed_f64(const double* I, int m, int n, double sigma, const gbm3d_opts* resolved, double* U,
                              double* U2, const double* pilot) {
  if (!I || !U || !resolved || m < 1 || n < 1) return GBM3D_ERR_ARG;
  return guarded([&] { return gbm3d_distributed<double>(I, m, n, sigma, *resolved, U, U2, pilot); });
}
int orc_gbm3d_distributed_f32(const float* I, int m, int n, float sigma, const gbm3d_opts* resolved, float* U,
                              float* U2, const float* pilot) {
  if (!I || !U || !resolved || m < 1 || n < 1) return GBM3D_ERR_ARG;
  return guarded([&] { return gbm3d_distributed<float>(I, m, n, sigma, *resolved, U, U2, pilot); });
}

}  // extern "C"
