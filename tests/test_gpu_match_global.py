"""matchBlocksGPU (source/matchBlocksGPU.m:18-91) on the GPU through the C ABI
(gbm3d_match_global) against the FP32 restatement of the oracle: the index
lists, their lengths and the prefiltered image bit for bit (integer and index
work; the prefilter and the distances are the same FP32 fused multiply-add
chains in the same order)."""
import numpy as np
import pytest

import gbm3d_amd
import oracle as O
from gbm3d_amd import _abi, synth
from gbm3d_amd._abi import GBM3DError

pytestmark = pytest.mark.gpu

CASES = [  # m, n, d, sigma, numMax, numMin, scale
    (32, 48, 4, 25.0, 16, 16, 255.0),     # prefilter (sigma > 30/255), 4 x 4 blocks
    (48, 40, 8, 0.05, 8, 8, 1.0),         # [0, 1] image, no prefilter
    (40, 37, 8, 0.1, 16, 16, 1.0),        # not a block / window multiple
    (96, 96, 8, 0.4, 16, 4, 255.0),       # tolerance branch
    (64, 80, 4, 0.3, 16, 4, 255.0),
    (64, 64, 16, 25.0, 8, 8, 255.0),
    (64, 96, 32, 25.0, 8, 2, 255.0),
    (256, 256, 8, 25.0, 16, 16, 255.0),
    (384, 320, 16, 50.0, 32, 32, 255.0),
    (8, 24, 8, 25.0, 64, 64, 255.0),      # smallest prefilter size (filters cut), K up to 64
]


def _image(m, n, d, scale):
    _, x = synth.make_pair(700 + m + n + d, m, n, 10.0)
    return (x * (scale / 255.0)).astype(np.float32)


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}x{c[1]}_d{c[2]}_s{c[3]}_{c[4]}_{c[5]}" for c in CASES])
def test_match_global_bit_exact_vs_fp32_restatement(ctx, case):
    m, n, d, sigma, nmax, nmin, scale = case
    I = _image(m, n, d, scale)
    print(f"match_global {m}x{n} d={d} sigma={sigma} numMax={nmax} numMin={nmin}", flush=True)
    S, cnt, F = ctx.match_global(I, sigma, d, nmax, nmin, want_filtered=True)
    So, cnto, Fo = O.match_global(I, sigma, d, nmax, nmin, precision=32, want_filtered=True)
    np.testing.assert_array_equal(F, Fo)
    np.testing.assert_array_equal(cnt, cnto)
    np.testing.assert_array_equal(S, So)
    if nmin < nmax and sigma < 1:
        assert ((cnt > nmin) & (cnt < nmax)).any()


def test_match_global_python_mirror_is_the_cell_array(ctx):
    I = _image(96, 96, 8, 255.0)
    opts = {"blockSize": 8, "numMax": 16, "numMin": 4}
    cells = gbm3d_amd.matchBlocksGPU(I, 0.4, opts)
    So, cnto = O.match_global(I, 0.4, 8, 16, 4, precision=32)
    assert cells.shape == (12, 12)
    for i in range(12):
        for j in range(12):
            assert cells[i, j].dtype == np.uint32
            np.testing.assert_array_equal(cells[i, j], So[:cnto[i, j], i, j])


def test_match_global_repeat_calls_and_arguments(ctx):
    """A smaller call after a larger one on the same context (the arena is
    reused), and the argument errors, raised before anything is launched."""
    a = _image(128, 128, 8, 255.0)
    b = _image(40, 48, 4, 255.0)
    S1, _ = ctx.match_global(a, 25.0, 8, 16)
    S2, c2 = ctx.match_global(b, 25.0, 4, 8, 8)
    np.testing.assert_array_equal(S2, O.match_global(b, 25.0, 4, 8, precision=32)[0])
    np.testing.assert_array_equal(S1, O.match_global(a, 25.0, 8, 16, precision=32)[0])
    for args, status in [((_image(36, 40, 4, 255.0), 25.0, 4, 8), _abi.GBM3D_ERR_ARG),   # prefilter: m % 8
                         ((a, 25.0, 8, 8, 9), _abi.GBM3D_ERR_ARG),                        # numMin > numMax
                         ((a, 25.0, 8, 65), _abi.GBM3D_ERR_ARG),
                         ((a, 25.0, 12, 8), _abi.GBM3D_ERR_UNSUPPORTED)]:
        with pytest.raises(GBM3DError) as e:
            ctx.match_global(*args)
        assert e.value.status == status
    S3, _ = ctx.match_global(b, 25.0, 4, 8)
    np.testing.assert_array_equal(S3, S2)


def _degenerate(kind, m, n):
    if kind == "zeros":
        return np.zeros((m, n), np.float32)
    if kind == "constant":
        return np.full((m, n), 100.0, np.float32)
    if kind == "ramp":
        r, c = np.meshgrid(np.arange(m), np.arange(n), indexing="ij")
        return (r + 3.0 * c).astype(np.float32)
    I = np.zeros((m, n), np.float32)              # sparse: a few bright pixels
    I[::13, ::7] = 200.0
    return I


@pytest.mark.parametrize("kind", ["zeros", "constant", "ramp", "sparse"])
@pytest.mark.parametrize("nmin", [16, 4])
def test_match_global_degenerate_images(ctx, kind, nmin):
    """Exact ties, 0/0 and Inf distances (zero-norm reference blocks), all-tied
    windows: the running bound never excludes a candidate that can still tie,
    and ties resolve to the lower index as mink does."""
    I = _degenerate(kind, 64, 48)
    S, cnt = ctx.match_global(I, 0.05, 8, 16, nmin)
    So, cnto = O.match_global(I, 0.05, 8, 16, nmin, precision=32)
    np.testing.assert_array_equal(cnt, cnto)
    np.testing.assert_array_equal(S, So)
