"""The FP32 matcher ranks candidates by num = |x|^2 + |y|^2 - 2<x,y> and divides
only those with num <= theta + |theta| 2^-20 + |x|^2 2^-126 (theta = the K-th
smallest of the 64 lane minima of num), then keys them by (rounded quotient,
index).  Property checked here in float32 on adversarial inputs: that set always
contains the exact top K by (num / |x|^2 rounded, index), including quotient
ties at the cut, negative numerators and quotients below the normal range.
(The GPU kernel itself is checked bit-exact against the oracle's FP32
restatement, which divides every candidate, in test_gpu_parity.py.)"""
import numpy as np


def _ord(x):
    b = x.view(np.uint32).astype(np.uint64)
    b = np.where(b == 0x80000000, 0, b)
    o = np.where(b & 0x80000000, (~b) & 0xFFFFFFFF, b | 0x80000000)
    return np.where(np.isnan(x), 0xFFFFFFFF, o).astype(np.uint64)


def _from_ord(o):
    o = np.uint32(o)
    b = (o & np.uint32(0x7FFFFFFF)) if (o & np.uint32(0x80000000)) else ~o
    return np.array([b], dtype=np.uint32).view(np.float32)[0]


def test_numerator_bound_keeps_exact_top_k():
    rng = np.random.default_rng(7)
    K = 16
    idx = np.arange(1024, dtype=np.uint64)
    for trial in range(4000):
        nx = np.float32(10 ** rng.uniform(-3, 7))
        base = np.float32(rng.uniform(-1e-3, 5) * nx)
        mode = trial % 4
        if mode == 0:
            num = rng.standard_normal(1024) * nx * 0.5 + base
        elif mode == 1:  # values a few ulps apart: rounded quotients tie across the cut
            num = base + rng.integers(-40, 40, 1024) * np.spacing(base)
        elif mode == 2:  # quotients below the normal range
            num = rng.standard_normal(1024) * 1e-38 * nx
        else:
            num = np.abs(rng.standard_normal(1024)) * rng.choice([1e-30, 1.0, 1e20])
        num = num.astype(np.float32)
        key = (_ord((num / nx).astype(np.float32)) << np.uint64(32)) | idx
        th = np.sort(_ord(num.reshape(64, 16)).min(axis=1))[K - 1]
        tnum = _from_ord(th)
        bound = np.float32(np.float32(np.abs(tnum) * np.float32(2.0 ** -20)) + tnum)
        bound = np.float32(bound + np.float32(nx * np.float32(2.0 ** -126)))
        sel = num <= bound
        assert sel.sum() >= K
        np.testing.assert_array_equal(np.sort(key[sel])[:K], np.sort(key)[:K])
