"""The BF16-MFMA matcher (precision form) against the FP32 restatement.

Its approximate scores carry rigorous error bounds and every candidate the
bounds cannot exclude from the top K is rescored with the FP32 chain, so the
index lists and gathered blocks must equal the FP32 restatement's bit for
bit -- exact distance ties included -- whatever the first rescored set's size.
"""
import numpy as np
import pytest

import gbm3d_amd
import oracle as O
from gbm3d_amd import synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def cbf():
    c = gbm3d_amd.Context(0)
    yield c
    c.close()


def _tiles(m, n, d):
    return (m // d) * (n // d)


@pytest.mark.parametrize("m,n,d,d2,seed", [(256, 256, 8, 32, 91), (256, 192, 16, 32, 92), (192, 256, 8, 48, 93),
                                           (128, 128, 16, 48, 94), (256, 200, 8, 32, 96), (96, 80, 16, 48, 97),
                                           (64, 48, 16, 32, 98)])
@pytest.mark.parametrize("rescore", [0, 16, 64])
def test_match_bf16_equals_fp32_restatement(cbf, m, n, d, d2, seed, rescore):
    clean, noisy = synth.make_pair(seed, m, n, 25.0)
    cbf.set_variant(match_bf16=True, rescore=rescore, match_stats=True)
    cbf.bf16_match_stats(reset=True)
    S, V = cbf.match(noisy, d, d2, 16)
    st = cbf.bf16_match_stats(reset=True)
    So, Vo, _ = O.match(noisy, d, d2, 16, precision=32)
    np.testing.assert_array_equal(S, So)
    np.testing.assert_array_equal(V, Vo)
    assert st["first_set"] + st["widened"] + st["exact_all"] == _tiles(m, n, d)


@pytest.mark.parametrize("d,K,spin", [(8, 16, 0), (8, 16, 4), (16, 16, 0), (8, 8, 4), (8, 32, 0)])
def test_match_wie_bf16_on_a_pilot_equals_fp32_restatement(cbf, d, K, spin):
    """Step-2 matching runs on the (smooth) pilot: the bounds must hold there too."""
    m, n = 192, 160
    clean, noisy = synth.make_pair(300 + d + K + spin, m, n, 25.0)
    pilot = np.roll(np.maximum(clean + 1.5 * synth.noise(301, m, n), 0).astype(np.float32), (spin, spin), (0, 1))
    In = np.roll(noisy, (spin, spin), (0, 1))
    cbf.set_variant(match_bf16=True)
    S, V1, V2 = cbf.match_wie(pilot, In, d, 32, K)
    So, V1o, V2o, _ = O.match_wie(pilot, In, d, 32, K, precision=32)
    np.testing.assert_array_equal(S, So)
    np.testing.assert_array_equal(V1, V1o)
    np.testing.assert_array_equal(V2, V2o)


@pytest.mark.parametrize("kind", ["zeros", "half", "constant", "steps", "large"])
def test_match_bf16_degenerate_images(cbf, kind):
    """Zero-norm reference blocks (every exact distance Inf or NaN), long runs
    of exactly tied candidates, a constant image, and values far from [0, 255]:
    all equal the FP32 restatement, ties broken by the lower index as there."""
    clean, noisy = synth.make_pair(99, 128, 144, 25.0)
    if kind == "zeros":
        I = np.zeros_like(noisy)
    elif kind == "half":
        I = noisy.copy()
        I[:, :64] = 0.0
    elif kind == "constant":
        I = np.full_like(noisy, 117.0)
    elif kind == "steps":
        I = (np.floor(clean / 64.0) * 64.0).astype(np.float32)
    else:
        I = (noisy * 1e4 + 3e6).astype(np.float32)
    cbf.set_variant(match_bf16=True, rescore=32)
    for d in (8, 16):
        S, V = cbf.match(I, d, 32, 16)
        So, Vo, _ = O.match(I, d, 32, 16, precision=32)
        np.testing.assert_array_equal(S, So)
        np.testing.assert_array_equal(V, Vo)


def test_match_bf16_bounds_widen_the_set_when_needed(cbf):
    """With the first set at its minimum (R = K) some tiles need the widened
    set; the result is still the FP32 restatement's."""
    clean, noisy = synth.make_pair(310, 256, 256, 10.0)
    cbf.set_variant(match_bf16=True, rescore=16, match_stats=True)
    cbf.bf16_match_stats(reset=True)
    S, _ = cbf.match(noisy, 8, 32, 16, want_volume=False)
    st = cbf.bf16_match_stats(reset=True)
    So, _, _ = O.match(noisy, 8, 32, 16, precision=32)
    np.testing.assert_array_equal(S, So)
    assert st["widened"] > 0, st
    assert st["first_set"] + st["widened"] + st["exact_all"] == _tiles(256, 256, 8)


@pytest.mark.parametrize("opts", [None, {"blockSize": 8, "blockSizeWie": 8, "matchSpinsWie": [0, 4]},
                                  {"profile": "accuracy"}], ids=["default", "p8", "accuracy"])
def test_bf16_form_end_to_end_bitwise_equal_to_plain(ctx, cbf, opts):
    clean, noisy = synth.make_pair(95, 256, 256, 25.0)
    cbf.set_variant(match_bf16=True)
    U32, o32 = ctx.denoise(noisy, 25.0, opts)
    Ubf, obf = cbf.denoise(noisy, 25.0, opts)
    np.testing.assert_array_equal(obf["U1"], o32["U1"])
    np.testing.assert_array_equal(Ubf, U32)
