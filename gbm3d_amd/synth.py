"""Seeded synthetic test images (clean + AWGN) for the denoising workloads.

Construction per image (m rows x n columns, values on [0, 255]):
  background 0; 12 axis-aligned rectangles then 8 disks, each with an intensity
  drawn U[0, 255], painted in that order; one oriented sinusoidal texture patch
  of amplitude 20 and period U[6, 16] px added on a rectangle; clip to [0, 255].
Noise: additive N(0, sigma^2), not clipped (as demoDenoise.m:15 does).

The random stream is counter based (SplitMix64 of (key, counter)), so the same
seed gives the same image on every machine and in every language binding.
Noise sample of pixel (r, c) uses counters 2i, 2i+1 with i = r + m*c
(column-major) and Box-Muller.
"""
from __future__ import annotations

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)


def _splitmix64(x: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = (x + _GOLDEN).astype(np.uint64)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform(key: int, ctr) -> np.ndarray:
    """U[0,1) doubles for counters ctr under stream key."""
    k = _splitmix64(np.array([key & 0xFFFFFFFFFFFFFFFF], dtype=np.uint64))[0]
    c = np.asarray(ctr, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = _splitmix64(k + c * _GOLDEN)
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


class _Stream:
    def __init__(self, key: int):
        self.key = key
        self.ctr = 0

    def u(self, lo: float = 0.0, hi: float = 1.0) -> float:
        v = float(uniform(self.key, self.ctr))
        self.ctr += 1
        return lo + (hi - lo) * v


# Uniform draws per image, in stream order: 12 rectangles x (h, w, r0, c0,
# value), 8 disks x (radius, row, col, value), texture (h, w, r0, c0, theta,
# period, phase).
N_DRAWS = 12 * 5 + 8 * 4 + 7


def shape_params(seed: int, m: int, n: int) -> dict:
    """Shape parameters of the clean image (the first N_DRAWS stream values)."""
    u = uniform(seed * 2 + 1, np.arange(N_DRAWS, dtype=np.uint64))
    it = iter(u.tolist())

    def draw(lo, hi):
        return lo + (hi - lo) * next(it)

    rects, disks = [], []
    for _ in range(12):
        h = int(draw(m / 16, m / 2))
        w = int(draw(n / 16, n / 2))
        r0 = int(draw(0, max(m - h, 1)))
        c0 = int(draw(0, max(n - w, 1)))
        rects.append((r0, r0 + h, c0, c0 + w, draw(0, 255)))
    s = min(m, n)
    for _ in range(8):
        rad = draw(s / 32, s / 6)
        cr = draw(0, m)
        ccn = draw(0, n)
        disks.append((cr, ccn, rad * rad, draw(0, 255)))
    h = int(draw(m / 8, m / 3))
    w = int(draw(n / 8, n / 3))
    r0 = int(draw(0, max(m - h, 1)))
    c0 = int(draw(0, max(n - w, 1)))
    theta = draw(0, np.pi)
    period = draw(6, 16)
    phase = draw(0, 2 * np.pi)
    tex = (r0, r0 + h, c0, c0 + w, theta, period, phase)
    return {"rects": rects, "disks": disks, "texture": tex}


def clean_image(seed: int, m: int, n: int) -> np.ndarray:
    """The clean synthetic image (float64, shape (m, n))."""
    prm = shape_params(seed, m, n)
    img = np.zeros((m, n), dtype=np.float64)
    rr, cc = np.meshgrid(np.arange(m, dtype=np.float64), np.arange(n, dtype=np.float64), indexing="ij")
    for r0, r1, c0, c1, val in prm["rects"]:
        img[r0:r1, c0:c1] = val
    for cr, ccn, rad2, val in prm["disks"]:
        img[(rr - cr) ** 2 + (cc - ccn) ** 2 <= rad2] = val
    r0, r1, c0, c1, theta, period, phase = prm["texture"]
    box = (slice(r0, r1), slice(c0, c1))
    img[box] += 20.0 * np.sin(2 * np.pi * (rr[box] * np.cos(theta) + cc[box] * np.sin(theta)) / period + phase)
    return np.clip(img, 0.0, 255.0)


def noise(seed: int, m: int, n: int) -> np.ndarray:
    """Standard normal field (float64, shape (m, n)), Box-Muller."""
    i = np.arange(m * n, dtype=np.uint64)
    key = seed * 2 + 2
    u1 = uniform(key, 2 * i)
    u2 = uniform(key, 2 * i + 1)
    z = np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2 * np.pi * u2)
    return z.reshape(n, m).T  # counter i = r + m*c


def make_pair(seed: int, m: int, n: int, sigma: float):
    """(clean, noisy) float32 images of shape (m, n)."""
    clean = clean_image(seed, m, n)
    noisy = clean + sigma * noise(seed, m, n)
    return clean.astype(np.float32), noisy.astype(np.float32)


def psnr(ref: np.ndarray, rec: np.ndarray, max_i: float = 255.0) -> float:
    """myPSNR (utilities/myPSNR.m:11-12): 10*log10(maxI^2/MSE)."""
    ref = np.asarray(ref, dtype=np.float64)
    rec = np.asarray(rec, dtype=np.float64)
    mse = np.sum((ref - rec) ** 2) / ref.size
    return float(10.0 * np.log10(max_i ** 2 / mse))
