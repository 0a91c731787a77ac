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


def test_gpu_match_images_cover_every_selection_size():
    """The matcher sorts the compacted set over 32 lanes (at most 32 members),
    over 64 lanes (33-64), by a per-lane sort plus K wave-minimum rounds
    (65-256) or by K rounds over every candidate (more).  The synthetic images of
    test_gpu_parity.py::test_match_bit_exact_vs_fp32_restatement hold tiles of
    the first three sizes (the degenerate images there the last), counted here
    with the kernel's lane layout (lane l: window column l % 32, rows
    16 (l // 32) .. +15) on the FP32 restatement's distances (the quotient is
    monotone in the numerator for |x|^2 > 0)."""
    import oracle as O
    from gbm3d_amd import synth

    sizes = []
    for m, n, d, seed in ((256, 192, 8, 4), (256, 256, 16, 3)):
        _, noisy = synth.make_pair(seed, m, n, 25.0)
        _, _, dist = O.match(noisy, d, 32, 16, precision=32, want_dist=True)
        dist = np.asarray(dist)
        for i in range(dist.shape[1]):
            for j in range(dist.shape[2]):
                dd = dist[:, i, j].reshape(32, 32, order="F")  # [a, b], window index a + 32 b
                lanes = np.stack([dd[16 * (l // 32):16 * (l // 32) + 16, l % 32] for l in range(64)])
                theta = np.sort(lanes.min(axis=1))[15]
                sizes.append(int((lanes <= theta * (1 + 2.0**-20)).sum()))
    sizes = np.array(sizes)
    assert (sizes <= 32).sum() > 0
    assert ((sizes > 32) & (sizes <= 64)).sum() > 0
    assert ((sizes > 64) & (sizes <= 256)).sum() > 10
