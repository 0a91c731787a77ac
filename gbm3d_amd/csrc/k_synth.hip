// k_synth.hip -- seeded synthetic clean/noisy image pairs on the device.
//
// The construction (values on [0, 255]): background 0; 12 axis-aligned
// rectangles, then 8 disks, each painted with an intensity U[0, 255]; an
// oriented sinusoid of amplitude 20 and period U[6, 16] px added on a
// rectangle; clip to [0, 255].  Noise N(0, sigma^2) is added unclipped (the
// reference demo, demoDenoise.m:15, does not clip).  Random numbers come from a
// counter-based SplitMix64 stream: draw c of stream key is
// mix(mix(key) + c * golden) with mix(x) the SplitMix64 output function of
// x + golden, mapped to [0, 1) by its top 53 bits.  Shape draws use key
// 2*seed + 1 (counters 0..98), the noise of pixel i = r + m*c uses key
// 2*seed + 2, counters 2i and 2i + 1, through Box-Muller.  Pixel values are
// formed in double and rounded once to float, so the host construction in the
// Python package gives the same images.
#include <cmath>

#include "internal.hpp"

namespace gbm3d {

namespace {

constexpr uint64_t kGolden = 0x9E3779B97F4A7C15ull;

__host__ __device__ inline uint64_t splitmix64(uint64_t x) {
  uint64_t z = x + kGolden;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__host__ __device__ inline double uniform01(uint64_t key_mixed, uint64_t ctr) {
  const uint64_t z = splitmix64(key_mixed + ctr * kGolden);
  return (double)(z >> 11) * (1.0 / 9007199254740992.0);
}

}  // namespace

void synth_params(uint64_t seed, int m, int n, SynthParams* p) {
  const uint64_t key = splitmix64(seed * 2 + 1);
  uint64_t ctr = 0;
  auto draw = [&](double lo, double hi) { return lo + (hi - lo) * uniform01(key, ctr++); };
  for (int i = 0; i < 12; ++i) {
    const int h = (int)draw(m / 16.0, m / 2.0);
    const int w = (int)draw(n / 16.0, n / 2.0);
    const int r0 = (int)draw(0.0, (double)std::max(m - h, 1));
    const int c0 = (int)draw(0.0, (double)std::max(n - w, 1));
    p->rect[i][0] = r0;
    p->rect[i][1] = r0 + h;
    p->rect[i][2] = c0;
    p->rect[i][3] = c0 + w;
    p->rect_val[i] = draw(0.0, 255.0);
  }
  const int s = std::min(m, n);
  for (int i = 0; i < 8; ++i) {
    const double rad = draw(s / 32.0, s / 6.0);
    p->disk[i][0] = draw(0.0, (double)m);
    p->disk[i][1] = draw(0.0, (double)n);
    p->disk[i][2] = rad * rad;
    p->disk[i][3] = draw(0.0, 255.0);
  }
  const int h = (int)draw(m / 8.0, m / 3.0);
  const int w = (int)draw(n / 8.0, n / 3.0);
  const int r0 = (int)draw(0.0, (double)std::max(m - h, 1));
  const int c0 = (int)draw(0.0, (double)std::max(n - w, 1));
  p->tex[0] = r0;
  p->tex[1] = r0 + h;
  p->tex[2] = c0;
  p->tex[3] = c0 + w;
  p->tex_theta = draw(0.0, M_PI);
  p->tex_period = draw(6.0, 16.0);
  p->tex_phase = draw(0.0, 2.0 * M_PI);
}

__global__ __launch_bounds__(256) void synth_pair_kernel(SynthParams p, uint64_t noise_key, int m, int n,
                                                         double sigma, float* __restrict__ clean,
                                                         float* __restrict__ noisy) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)m * n) return;
  const int r = (int)(i % m), c = (int)(i / m);
  double v = 0.0;
  for (int k = 0; k < 12; ++k)
    if (r >= p.rect[k][0] && r < p.rect[k][1] && c >= p.rect[k][2] && c < p.rect[k][3]) v = p.rect_val[k];
  const double rr = (double)r, cc = (double)c;
  for (int k = 0; k < 8; ++k) {
    const double dr = rr - p.disk[k][0], dc = cc - p.disk[k][1];
    if (dr * dr + dc * dc <= p.disk[k][2]) v = p.disk[k][3];
  }
  if (r >= p.tex[0] && r < p.tex[1] && c >= p.tex[2] && c < p.tex[3]) {
    const double t = rr * cos(p.tex_theta) + cc * sin(p.tex_theta);
    v = v + 20.0 * sin(2.0 * M_PI * t / p.tex_period + p.tex_phase);
  }
  v = v < 0.0 ? 0.0 : (v > 255.0 ? 255.0 : v);
  if (clean) clean[i] = (float)v;
  if (noisy) {
    const double u1 = uniform01(noise_key, 2 * (uint64_t)i);
    const double u2 = uniform01(noise_key, 2 * (uint64_t)i + 1);
    const double z = sqrt(-2.0 * log1p(-u1)) * cos(2.0 * M_PI * u2);
    noisy[i] = (float)(v + sigma * z);
  }
}

int launch_synth_pair(const SynthParams& p, uint64_t seed, int m, int n, float sigma, float* clean, float* noisy,
                      hipStream_t st) {
  if (m <= 0 || n <= 0) return fail(GBM3D_ERR_ARG, "synth: empty image");
  const size_t total = (size_t)m * n;
  hipLaunchKernelGGL(synth_pair_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, p,
                     splitmix64(seed * 2 + 2), m, n, (double)sigma, clean, noisy);
  return check_launch("synth_pair");
}

}  // namespace gbm3d
