"""Option combinations beyond the profiles: every block / match / group size
the kernels are built for runs against the oracle, and every other one is
rejected before anything is launched (the context keeps working).

Bars as in test_gpu_parity.py: index sets and gathered blocks bit-exact
against the FP32 restatement; float stages on identical inputs within 1e-3
(max abs, [0, 255] scale) of the FP64 oracle; end to end chained within 1e-3
of the FP32 restatement, PSNR within 0.01 dB of the FP64 oracle.
"""
import ctypes as C

import numpy as np
import pytest

import gbm3d_amd
import oracle as O
from gbm3d_amd import _abi, synth
from gbm3d_amd._abi import GBM3DError

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("m,n,d,d2,K,seed", [(96, 96, 4, 48, 16, 601), (128, 96, 32, 48, 16, 602),
                                             (96, 128, 32, 48, 8, 603)])
def test_match_size_48_small_and_large_blocks_bit_exact(ctx, m, n, d, d2, K, seed):
    clean, noisy = synth.make_pair(seed, m, n, 25.0)
    S, V = ctx.match(noisy, d, d2, K)
    So, Vo, _ = O.match(noisy, d, d2, K, precision=32)
    np.testing.assert_array_equal(S, So)
    np.testing.assert_array_equal(V, Vo)


@pytest.mark.parametrize("d,K,seed", [(4, 8, 611), (4, 32, 612), (16, 8, 613), (16, 32, 614), (8, 8, 615),
                                      (8, 32, 616)])
def test_wiener_group_sizes_vs_fp64_oracle(ctx, d, K, seed):
    """denoiseAndAggregateWieFast.m:19-66 at blockSizeWie 4/8/16 and numMax
    8/16/32 (levels = 2 below 16 slices, :24; DC protection of slices 15:end)."""
    m, n = 96, 128
    clean, noisy = synth.make_pair(seed, m, n, 25.0)
    pilot = np.maximum(clean + 3.0 * synth.noise(seed + 7, m, n), 0)
    S, V1, V2, _ = O.match_wie(pilot, noisy, d, 32, K, precision=32)
    V2f, g2 = ctx.wiener(V1, V2, 25.0)
    V2o, g2o = O.wiener(V1.astype(np.float64), V2.astype(np.float64), 25.0, precision=64)
    assert np.abs(V2f - V2o).max() <= 1e-3
    assert np.abs(g2 - g2o).max() <= 1e-5 * g2o.max()
    S2, V1g, V2g = ctx.match_wie(pilot.astype(np.float32), noisy, d, 32, K)
    np.testing.assert_array_equal(S2, S)
    np.testing.assert_array_equal(V1g, V1)


@pytest.mark.parametrize("opts", [
    {"blockSize": 8, "blockSizeWie": 4, "numMax": 32, "numMin": 32, "matchSpinsWie": [0, 4]},
    {"blockSize": 16, "blockSizeWie": 16, "numMax": 8, "numMin": 8, "matchSpinsWie": [0, 4]},
    {"blockSize": 32, "blockSizeWie": 8, "matchSize": 48, "matchSpinsWie": [0, 4]},
], ids=["wie4_K32", "wie16_K8", "bs32_d2_48"])
def test_end_to_end_other_sizes_vs_oracle(ctx, opts):
    m, n = 128, 96
    clean, noisy = synth.make_pair(620 + opts["blockSizeWie"], m, n, 25.0)
    U, out = ctx.denoise(noisy, 25.0, opts)
    _, U132 = O.gbm3d(noisy, 25.0, opts, precision=32)
    assert np.max(np.abs(out["U1"] - U132)) <= 1e-3
    Uc = O.step2_from_pilot(out["U1"], noisy, 25.0, opts)
    assert np.max(np.abs(U - Uc)) <= 1e-3
    U64, _ = O.gbm3d(noisy, 25.0, opts, precision=64)
    assert abs(synth.psnr(clean, U) - synth.psnr(clean, U64)) <= 0.01


BAD = {"numMax": 24, "numMin": 24, "matchSpinsWie": [0]}   # a Wiener group of 24 slices is not built


def test_unbuilt_options_rejected_before_any_launch(ctx):
    """Each entry point checks the whole support table first, so nothing is
    issued (no half-captured graph, no forked lane left unjoined) and the
    context denoises normally afterwards."""
    opts = {"blockSize": 8, "blockSizeWie": 8, "matchSpinsWie": [0, 4]}
    _, big = synth.make_pair(630, 600, 520, 25.0)
    U_ref, o_ref = ctx.denoise_distributed(big, 25.0, opts)
    for call in (lambda: ctx.denoise_distributed(big, 25.0, BAD), lambda: ctx.denoise(big[:96, :96], 25.0, BAD),
                 lambda: ctx.denoise_batch([big[:96, :96], big[:64, :128]], [25.0, 20.0], BAD)):
        with pytest.raises(GBM3DError, match="Wiener step") as e:
            call()
        assert e.value.status == _abi.GBM3D_ERR_UNSUPPORTED
    U, o = ctx.denoise_distributed(big, 25.0, opts)
    assert np.array_equal(U, U_ref) and np.array_equal(o["U2"], o_ref["U2"])
    fresh = gbm3d_amd.Context(0)
    try:
        V, _ = fresh.denoise(big[:96, :96], 25.0, opts)
        W, _ = ctx.denoise(big[:96, :96], 25.0, opts)
        assert np.array_equal(V, W)
    finally:
        fresh.close()


def test_bf16_form_on_unbuilt_block_sizes_is_the_fp32_search():
    """The BF16 matcher is built for 8 x 8 and 16 x 16 blocks; 'superFast'
    (32 x 32 blocks) with the BF16 form runs the FP32 search, bit for bit."""
    _, noisy = synth.make_pair(640, 128, 128, 25.0)
    plain, bf = gbm3d_amd.Context(0), gbm3d_amd.Context(0)
    try:
        bf.set_variant(match_bf16=True)
        U0, o0 = plain.denoise(noisy, 25.0, {"profile": "superFast"})
        U1, o1 = bf.denoise(noisy, 25.0, {"profile": "superFast"})
        assert np.array_equal(U0, U1) and np.array_equal(o0["U1"], o1["U1"])
    finally:
        plain.close()
        bf.close()


def test_colored_sigma_wie_size_checked_by_the_c_abi(ctx):
    """A sigmaWie that is not blockSizeWie x blockSizeWie is rejected by the
    library itself (no read past a short array), not only by the Python mirror."""
    lib = gbm3d_amd.load_library()
    _, noisy = synth.make_pair(650, 64, 64, 25.0)
    I = np.asfortranarray(noisy, dtype=np.float32)
    psd = np.full((16, 16), 625.0, order="F")
    sw = np.full((4, 4), 25.0, order="F")                 # blockSizeWie is 8
    o = _abi.opts_to_struct({"blockSize": 8, "blockSizeWie": 8, "matchSpinsWie": [0]})
    U = np.zeros_like(I)
    st = lib.gbm3d_denoise_colored(ctx.handle, _abi.ptr(I), 64, 64, _abi.ptr(psd), 16, 16, _abi.ptr(sw), 4, 4,
                                   C.byref(o), _abi.ptr(U), None)
    assert st == _abi.GBM3D_ERR_ARG
    assert b"blockSizeWie x blockSizeWie" in lib.gbm3d_last_error()


@pytest.mark.parametrize("d,K", [(8, 16), (4, 8), (16, 32)])
def test_denoise_and_aggregate_wie_dct_twin_vs_oracle(ctx, d, K):
    """denoiseAndAggregateWie (source/denoiseAndAggregateWie.m): filter with
    the [0, 1] clip, then aggregation, against the FP64 oracle on a
    [0, 1]-scale image; clip01 = 0 is WieFast."""
    m, n = 96, 128
    clean, noisy = synth.make_pair(660 + d, m, n, 25.0)
    clean, noisy = clean / 255.0, (noisy / 255.0).astype(np.float32)
    pilot = np.maximum(clean + 0.02 * synth.noise(661, m, n), 0)
    S, V1, V2, _ = O.match_wie(pilot, noisy, d, 32, K, precision=32)
    s = 25.0 / 255.0
    for clip in (True, False):
        U, W = ctx.denoise_aggregate_wie(V1, V2, S, s, 32, clip01=clip)
        V2o, g2o = O.wiener(V1.astype(np.float64), V2.astype(np.float64), s, precision=64, clip01=clip)
        Uo, Wo = O.aggregate_wie(V2o, g2o, S, m, n, 32, precision=64)
        assert np.abs(U / W - Uo / Wo).max() <= 1e-3 / 255.0
    V2c, _ = O.wiener(V1.astype(np.float64), V2.astype(np.float64), s, precision=64, clip01=True)
    assert np.mean(V2c == 1.0) > 1e-4        # the clip is active on these inputs


def _periodic(seed=3, m=64, n=64):
    rng = np.random.default_rng(seed)
    tile = rng.uniform(20, 235, (8, 8))
    return (np.tile(tile, (m // 8, n // 8)) + 0.5 * rng.standard_normal((m, n))).astype(np.float32)


@pytest.mark.parametrize("nmax,nmin,sigma", [(16, 8, 0.177), (12, 4, 0.177), (16, 8, 25.0), (16, 4, 0.05)])
def test_match_tolerance_branch_bit_exact(ctx, nmax, nmin, sigma):
    """matchBlocksCPU's tolerance branch (numMax ~= numMin, matchBlocksCPU.m:
    89-111) against the FP32 restatement: find order when exactly numMax
    distances are below tol, mink otherwise; where the reference fails (fewer
    than numMax) both report it."""
    I = _periodic()
    try:
        So, Vo = O.match_tol(I, 8, 32, nmax, nmin, sigma, precision=32)
    except RuntimeError:
        with pytest.raises(GBM3DError) as e:
            ctx.match_tol(I, 8, 32, nmax, nmin, sigma)
        assert e.value.status == _abi.GBM3D_ERR_ARG
        return
    S, V = ctx.match_tol(I, 8, 32, nmax, nmin, sigma)
    np.testing.assert_array_equal(S, So)
    np.testing.assert_array_equal(V, Vo)


def test_tolerance_branch_end_to_end_vs_oracle(ctx):
    opts = {"blockSize": 8, "blockSizeWie": 8, "numMax": 16, "numMin": 8, "matchSpinsWie": [0, 4]}
    I = _periodic(5, 96, 64)
    sigma = 0.177
    U, out = ctx.denoise(I, sigma, opts)
    U32, U132 = O.gbm3d(I, sigma, opts, precision=32)
    assert np.max(np.abs(out["U1"] - U132)) <= 1e-3
    assert np.max(np.abs(U - O.step2_from_pilot(out["U1"], I, sigma, opts))) <= 1e-3
    Um, _ = ctx.denoise(I, sigma, dict(opts, numMin=16))
    assert not np.array_equal(U, Um)                        # the find order changes the groups
    with pytest.raises(GBM3DError) as e:                    # numMax 32: every tile is short
        ctx.denoise(I, sigma, dict(opts, numMax=32, numMin=8))
    assert e.value.status == _abi.GBM3D_ERR_ARG
    V, _ = ctx.denoise(I, sigma, opts)                      # the context still works
    assert np.array_equal(U, V)


def test_tolerance_branch_coloured_noise_uses_mean_sqrt_psd(ctx):
    """With a matrix sigma the tolerance branch takes sigma = mean(sqrt(PSD(:)))
    (matchBlocksCPU.m:21): a flat PSD of 0.177^2 puts exactly numMax copies
    below tol on the period-8 image, so the lists come in find order."""
    I = _periodic(7, 64, 64)
    psd = np.full((16, 16), 0.177 ** 2)
    sw = np.full((8, 8), 0.177)
    opts = {"blockSize": 8, "blockSizeWie": 8, "numMax": 16, "numMin": 8, "matchSpinsWie": [0, 4]}
    U, out = ctx.denoise_colored(I, psd, sw, opts)
    U32, U132 = O.gbm3d_colored(I, psd, sw, opts, precision=32)
    assert np.max(np.abs(out["U1"] - U132)) <= 1e-3
    Um, outm = ctx.denoise_colored(I, psd, sw, dict(opts, numMin=16))
    assert not np.array_equal(out["U1"], outm["U1"])
