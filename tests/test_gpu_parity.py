"""Parity of the HIP path (called through the C ABI) with the oracle.

Bars (written in each test):
  * index sets S and gathered blocks: bit-exact against the FP32 restatement,
    whose distances use the same fused-multiply-add order; against the FP64
    oracle they may differ only at near ties (normalised distance gap < 1e-5);
  * floating point stages, given identical inputs: max abs error <= 1e-3 on the
    [0, 255] scale against the FP64 oracle;
  * end to end: PSNR within 0.01 dB of the FP64 oracle, bitwise reproducible.
"""
import glob
import json
import os

import numpy as np
import pytest

import gbm3d_amd
import oracle as O
from gbm3d_amd import synth
from gbm3d_amd._abi import GBM3DError

pytestmark = pytest.mark.gpu
FIX = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))


def _near_tie_only(S_gpu, S64, dist64, tol=1e-5):
    """Every tile whose index list differs from the FP64 oracle's holds a near
    tie: two of its K+1 best FP64 distances closer than tol.  Distances are
    normalised by |x|^2, and the FP32 expansion |x|^2+|y|^2-2<x,y> loses about
    1e-6 of that scale to cancellation, so tol is absolute in those units."""
    bad = np.where(np.any(S_gpu != S64, axis=0))
    for i, j in zip(*bad):
        srt = np.sort(dist64[:, i, j])[: S64.shape[0] + 1]
        if np.diff(srt).min() >= tol:
            return False
    return True


@pytest.mark.parametrize("m,n,d,d2,seed", [(64, 64, 16, 32, 1), (64, 64, 8, 32, 2), (256, 256, 16, 32, 3),
                                           (256, 192, 8, 32, 4), (96, 128, 8, 48, 5), (128, 96, 16, 48, 6),
                                           (64, 96, 4, 32, 7), (96, 64, 32, 32, 8)])
def test_match_bit_exact_vs_fp32_restatement(ctx, m, n, d, d2, seed):
    clean, noisy = synth.make_pair(seed, m, n, 25.0)
    S, V = ctx.match(noisy, d, d2, 16)
    So, Vo, _ = O.match(noisy, d, d2, 16, precision=32)
    np.testing.assert_array_equal(S, So)
    np.testing.assert_array_equal(V, Vo)
    S64, _, dist = O.match(noisy, d, d2, 16, precision=64, want_dist=True)
    assert _near_tie_only(S, S64, dist)


@pytest.mark.parametrize("m,n,d,d2,K,seed", [(64, 64, 8, 32, 8, 11), (128, 96, 8, 32, 32, 12), (96, 64, 16, 32, 32, 13),
                                             (64, 96, 4, 32, 8, 14), (96, 128, 8, 48, 32, 15)])
def test_match_bit_exact_other_group_sizes(ctx, m, n, d, d2, K, seed):
    """K = 8 and 32 (numMax): the K-th lane minimum, the survivor bound and the
    32/64-lane sorts at other group sizes, bit for bit against the restatement."""
    clean, noisy = synth.make_pair(seed, m, n, 25.0)
    S, V = ctx.match(noisy, d, d2, K)
    So, Vo, _ = O.match(noisy, d, d2, K, precision=32)
    np.testing.assert_array_equal(S, So)
    np.testing.assert_array_equal(V, Vo)


@pytest.mark.parametrize("kind", ["zeros", "constant", "ramp", "sparse"])
def test_match_degenerate_images_vs_fp32_restatement(ctx, kind):
    """Exact ties and 0/0 distances (all-zero reference blocks) resolve as mink does."""
    m = n = 64
    if kind == "zeros":
        I = np.zeros((m, n), np.float32)
    elif kind == "constant":
        I = np.full((m, n), 17.0, np.float32)
    elif kind == "ramp":
        I = np.add.outer(np.arange(m), np.arange(n)).astype(np.float32) % 5
    else:
        I = np.zeros((m, n), np.float32)
        I[::7, ::5] = 100.0
    for d in (8, 16):
        S, V = ctx.match(I, d, 32, 16)
        So, Vo, _ = O.match(I, d, 32, 16, precision=32)
        np.testing.assert_array_equal(S, So)
        np.testing.assert_array_equal(V, Vo)


@pytest.mark.parametrize("m,n,seed", [(64, 64, 11), (256, 200, 12)])
def test_match_wie_bit_exact(ctx, m, n, seed):
    clean, noisy = synth.make_pair(seed, m, n, 25.0)
    pilot = np.maximum(clean + 2.0 * synth.noise(seed + 1, m, n), 0).astype(np.float32)
    S, V1, V2 = ctx.match_wie(pilot, noisy, 8, 32, 16)
    So, V1o, V2o, _ = O.match_wie(pilot, noisy, 8, 32, 16, precision=32)
    np.testing.assert_array_equal(S, So)
    np.testing.assert_array_equal(V1, V1o)
    np.testing.assert_array_equal(V2, V2o)


def _volume(seed, m, n, d=16):
    clean, noisy = synth.make_pair(seed, m, n, 25.0)
    S, V, _ = O.match(noisy, d, 32, 16, precision=64)
    return S, V


@pytest.mark.parametrize("m,n,cs,seed", [(64, 64, 2, 21), (256, 256, 2, 22), (128, 96, 1, 23), (96, 160, 3, 24)])
def test_ht_filter_vs_fp64_oracle(ctx, m, n, cs, seed):
    S, V = _volume(seed, m, n)
    tau = O.get_tau(3, 1, 25.0)
    Vg = ctx.ht_filter(V, 3, cs, tau)
    Vo, margin = O.ht_filter(V, 3, cs, tau, precision=64, want_margin=True)
    # FP32 coefficients carry ~1e-5 absolute error (~1e-7 of tau ~ 70); a margin of
    # 1e-6*tau keeps every hard-threshold decision identical to the oracle's
    assert margin > 1e-6, "input has a coefficient within 1e-6*tau of a threshold; pick another seed"
    assert np.abs(Vg - Vo).max() <= 1e-3


@pytest.mark.parametrize("m,n,K,levels,seed", [(64, 64, 16, 3, 51), (96, 160, 16, 3, 52), (48, 80, 8, 2, 53),
                                                (16, 16, 2, 1, 54), (256, 200, 16, 3, 55), (18, 22, 2, 1, 56),
                                                (20, 28, 2, 1, 58)])
def test_ht_fused_cycle_spins_equal_per_spin_kernels(ctx, monkeypatch, m, n, K, levels, seed):
    """The finest level of both cycle spins in one pass (anchors of parity s in
    original coordinates) is the per-spin computation with its operations in
    the same order: bitwise equal, in the mosaic layout and, end to end, in the
    group-contiguous layout the aggregation reads, with the step-1 weights the
    fused kernel writes at 8x8 blocks (the per-spin path runs the weights
    kernel); end to end on 136x104 images too, whose last output tiles are
    partial (64x32 tiles).  18x22 has row quads that straddle the volume's
    edge (32-bit stores), 20x28 a last run of 4 columns (16-byte stores)."""
    rng = np.random.default_rng(seed)
    V = (rng.standard_normal((m, n, K)) * 40 + 100).astype(np.float32)
    tau = O.get_tau(levels, 1, 25.0)
    a = ctx.ht_filter(V, levels, 2, tau)
    monkeypatch.setenv("GBM3D_HT_UNFUSED", "1")
    b = ctx.ht_filter(V, levels, 2, tau)
    np.testing.assert_array_equal(a, b)
    clean, noisy = synth.make_pair(seed, *((136, 104) if seed % 2 else (128, 96)), 25.0)
    Ub, ob = ctx.denoise(noisy, 25.0, {"blockSize": 8, "matchSpinsWie": [0, 4]})
    monkeypatch.delenv("GBM3D_HT_UNFUSED")
    Ua, oa = ctx.denoise(noisy, 25.0, {"blockSize": 8, "matchSpinsWie": [0, 4]})
    np.testing.assert_array_equal(oa["U1"], ob["U1"])
    np.testing.assert_array_equal(Ua, Ub)


@pytest.mark.parametrize("d", [4, 8, 16])
def test_aggregate_buffer_and_pointer_loads_agree(ctx, monkeypatch, d):
    """Hit blocks read through a buffer resource (arrays below 2 GB) and through
    64-bit pointers (GBM3D_AGG_POINTER=1, the path of larger arrays): bitwise
    equal, both steps and end to end."""
    clean, noisy = synth.make_pair(60 + d, 128, 96, 25.0)
    S, V, _ = O.match(noisy, d, 32, 16, precision=32)
    Vavg = np.maximum(V, 0).astype(np.float32)
    a = ctx.aggregate_ht(Vavg, S, d, 32)
    monkeypatch.setenv("GBM3D_AGG_POINTER", "1")
    b = ctx.aggregate_ht(Vavg, S, d, 32)
    Ub, ob = ctx.denoise(noisy, 25.0, {"blockSize": d, "matchSpinsWie": [0, 4]})
    monkeypatch.delenv("GBM3D_AGG_POINTER")
    Ua, oa = ctx.denoise(noisy, 25.0, {"blockSize": d, "matchSpinsWie": [0, 4]})
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)
    np.testing.assert_array_equal(oa["U1"], ob["U1"])
    np.testing.assert_array_equal(Ua, Ub)


def test_ht_filter_perfect_reconstruction(ctx):
    """tau = 0 keeps every coefficient: the filter is max(V, 0) (closed form)."""
    rng = np.random.default_rng(3)
    V = (rng.standard_normal((128, 96, 16)) * 60).astype(np.float32)
    Vg = ctx.ht_filter(V, 3, 2, np.zeros((3, 7)))
    assert np.abs(Vg - np.maximum(V, 0)).max() <= 1e-3


def test_ht_filter_tau_modes(ctx):
    S, V = _volume(25, 128, 128)
    for mode in (2, 3, 4):
        tau = O.get_tau(3, mode, 25.0)
        Vo, margin = O.ht_filter(V, 3, 2, tau, precision=64, want_margin=True)
        if margin < 1e-6:
            continue
        assert np.abs(ctx.ht_filter(V, 3, 2, tau) - Vo).max() <= 1e-3


@pytest.mark.parametrize("d,seed", [(16, 31), (8, 32)])
def test_aggregate_ht_vs_fp64_oracle(ctx, d, seed):
    clean, noisy = synth.make_pair(seed, 128, 96, 25.0)
    S, V, _ = O.match(noisy, d, 32, 16, precision=64)
    Vavg = O.ht_filter(V, 3, 2, O.get_tau(3, 1, 25.0)).astype(np.float32)
    U, W = ctx.aggregate_ht(Vavg, S, d, 32)
    Uo, Wo = O.aggregate_ht(Vavg.astype(np.float64), S, d, 32, precision=64)
    assert np.abs(U - Uo).max() <= 1e-5 * np.abs(Uo).max()
    assert np.abs(W - Wo).max() <= 1e-5 * np.abs(Wo).max()
    assert np.abs(U / W - Uo / Wo).max() <= 1e-3


@pytest.mark.parametrize("seed", [41, 42])
def test_wiener_vs_fp64_oracle(ctx, seed):
    clean, noisy = synth.make_pair(seed, 96, 128, 25.0)
    pilot = np.maximum(clean + 3.0 * synth.noise(seed + 7, 96, 128), 0)
    S, V1, V2, _ = O.match_wie(pilot, noisy, 8, 32, 16, precision=32)
    V2f, g2 = ctx.wiener(V1, V2, 25.0)
    V2o, g2o = O.wiener(V1.astype(np.float64), V2.astype(np.float64), 25.0, precision=64)
    assert np.abs(V2f - V2o).max() <= 1e-3
    assert np.abs(g2 - g2o).max() <= 1e-5 * g2o.max()
    U, W = ctx.aggregate_wie(V2o.astype(np.float32), g2o.astype(np.float32), S, 96, 128, 32)
    Uo, Wo = O.aggregate_wie(V2o, g2o, S, 96, 128, 32, precision=64)
    assert np.abs(U / W - Uo / Wo).max() <= 1e-3


def _chain_step2(ctx, noisy, sigma, opts, U):
    """U against the FP32 restatement's step 2 run on the GPU's own pilot; an
    image that is not a block multiple is chained on its mirror-padded form
    (the pipeline's first step), whose crop must be the plain call's U."""
    m, n = noisy.shape
    bs = O.resolve_opts(opts)["blockSize"]
    M, N = -(-m // bs) * bs, -(-n // bs) * bs
    Ip = O.mirror_pad(noisy, M, N).astype(np.float32) if (M, N) != (m, n) else np.asarray(noisy, np.float32)
    Up, outp = ctx.denoise(Ip, sigma, opts)
    assert np.array_equal(Up[:m, :n], U)
    U2 = O.step2_from_pilot(outp["U1"], Ip, sigma, opts, precision=32)
    assert np.abs(Up - U2).max() <= 1e-3


def _step2_chainable(opts):
    # 'fast'-like option sets (fewer Wiener spins than step-1 spins) keep
    # step-1 weights in the final quotient (source/GBM3D.m:26-27,82)
    ro = O.resolve_opts(opts)
    return ro["Wiener"] and len(ro["matchSpinsWie"]) >= len(ro["matchSpins"])


@pytest.mark.parametrize("path", FIX, ids=[os.path.basename(p) for p in FIX])
def test_end_to_end_golden(ctx, path):
    """PSNR of U and of the pilot U1 within 0.01 dB of the FP64 golden run;
    chained: U1 within 1e-3 of the FP32 restatement, U within 1e-3 of the
    restatement's step 2 on the GPU's own pilot."""
    f = np.load(path, allow_pickle=False)
    opts = json.loads(str(f["opts"]))
    sigma = float(f["sigma"])
    U, out = ctx.denoise(f["noisy"], sigma, opts)
    clean = f["clean"]
    assert abs(synth.psnr(clean, U) - synth.psnr(clean, f["U"])) < 0.01
    assert abs(synth.psnr(clean, out["U1"]) - synth.psnr(clean, f["U1"])) < 0.01
    _, U132 = O.gbm3d(f["noisy"], sigma, opts, precision=32)
    assert np.abs(out["U1"] - U132).max() <= 1e-3
    if _step2_chainable(opts):
        _chain_step2(ctx, f["noisy"], sigma, opts, U)


@pytest.mark.parametrize("opts,m,n", [(None, 256, 256), ({"blockSize": 8, "matchSpinsWie": [0, 4]}, 256, 192),
                                      ({"profile": "accuracy"}, 128, 160), ({"profile": "superFast"}, 256, 256),
                                      ({"profile": "fast"}, 200, 232),
                                      ({"blockSize": 8, "matchSpinsWie": [0, 4]}, 70, 66)])  # crop not a multiple of 4
def test_end_to_end_vs_oracle(ctx, opts, m, n):
    """Both steps, chained: the GPU pilot U1 against the FP32 restatement run on
    the same input (bit-exact indices, so only transform round-off remains),
    and the GPU output U against the restatement's step 2 run on the GPU's own
    pilot; max abs error <= 1e-3.  PSNR within 0.01 dB of the FP64 oracle."""
    clean, noisy = synth.make_pair(77 + m, m, n, 25.0)
    U, out = ctx.denoise(noisy, 25.0, opts)
    Uf, U1f = O.gbm3d(noisy, 25.0, opts, precision=32)
    assert np.abs(out["U1"] - U1f).max() <= 1e-3
    if _step2_chainable(opts):
        _chain_step2(ctx, noisy, 25.0, opts, U)
    Uo, U1o = O.gbm3d(noisy, 25.0, opts, precision=64)
    assert abs(synth.psnr(clean, U) - synth.psnr(clean, Uo)) < 0.01
    assert abs(synth.psnr(clean, out["U1"]) - synth.psnr(clean, U1o)) < 0.01


def test_end_to_end_is_bitwise_reproducible(ctx):
    clean, noisy = synth.make_pair(78, 192, 160, 25.0)
    a, oa = ctx.denoise(noisy, 25.0, None)
    b, ob = ctx.denoise(noisy, 25.0, None)
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(oa["U1"], ob["U1"])


@pytest.mark.parametrize("sigma", [10.0, 25.0, 50.0])
def test_full_size_2048_properties(ctx, sigma):
    """2048^2 (the benchmark size): finite, a large PSNR gain, reproducible,
    and scale equivariant (U(I/255, sigma/255) = U(I, sigma)/255)."""
    clean, noisy = synth.make_pair(2048, 2048, 2048, sigma)
    U, out = ctx.denoise(noisy, sigma, None)
    assert np.isfinite(U).all()
    gain = synth.psnr(clean, U) - synth.psnr(clean, noisy)
    assert gain > 6.0, gain
    U2, _ = ctx.denoise(noisy, sigma, None)
    np.testing.assert_array_equal(U, U2)
    Us, _ = ctx.denoise(noisy / 255.0, sigma / 255.0, None)
    assert abs(synth.psnr(clean, Us * 255.0) - synth.psnr(clean, U)) < 0.01


def test_error_paths(ctx):
    clean, noisy = synth.make_pair(1, 64, 64, 25.0)
    with pytest.raises(GBM3DError, match="unrecognized filtering"):
        ctx.denoise(noisy, 25.0, {"filtType": 7, "matchSpinsWie": [0]})
    with pytest.raises(GBM3DError, match="matchSpinsWie"):
        ctx.denoise(noisy, 25.0, {"blockSize": 8})
    with pytest.raises(GBM3DError, match="numMin"):
        ctx.denoise(noisy, 25.0, {"numMax": 8, "numMin": 16, "matchSpinsWie": [0]})
    # tolerance branch at sigma 25 on the [0, 255] scale: tol = 0.005 sigma^2 d^2
    # = 800 (16 x 16 blocks): here more than numMax candidates lie below it, so it
    # keeps the numMax nearest
    Ut, _ = ctx.denoise(noisy, 25.0, {"numMax": 16, "numMin": 8, "matchSpinsWie": [0]})
    Um, _ = ctx.denoise(noisy, 25.0, {"numMax": 16, "numMin": 16, "matchSpinsWie": [0]})
    assert np.array_equal(Ut, Um)
    with pytest.raises(GBM3DError):
        ctx.denoise(noisy[:12, :12], 25.0, None)  # pads to 16x16, smaller than matchSize
    U, _ = ctx.denoise(noisy, 25.0, None)  # the context still works afterwards
    assert np.isfinite(U).all()


def _batch_inputs():
    sizes = [(96, 128), (200, 150), (64, 64), (130, 260), (256, 192), (100, 100)]
    sigmas = [10.0, 25.0, 40.0, 5.0, 50.0, 20.0]
    imgs = [synth.make_pair(300 + i, m, n, s)[1] for i, ((m, n), s) in enumerate(zip(sizes, sigmas))]
    return imgs, sigmas


@pytest.mark.parametrize("graph", [False, True])
def test_batch_equals_single_image_runs(ctx, graph):
    """A batch over parallel streams (and its hipGraph) gives, bit for bit,
    the per-image results; graph replays interleaved with eager work too."""
    imgs, sigmas = _batch_inputs()
    ref = [ctx.denoise(I, s, None)[0] for I, s in zip(imgs, sigmas)]
    import gbm3d_amd

    bctx = gbm3d_amd.Context(0)
    out = bctx.denoise_batch(imgs, sigmas, None, graph=graph)
    for a, b in zip(out, ref):
        np.testing.assert_array_equal(a, b)
    if graph:
        from gbm3d_amd import device

        dI = [device.DeviceArray.from_host(I) for I in imgs]
        dU = [device.DeviceArray(I.shape) for I in imgs]
        for rep in range(3):
            bctx.denoise_batch_device([d.ptr for d in dI], [I.shape[0] for I in imgs], [I.shape[1] for I in imgs],
                                      sigmas, None, [d.ptr for d in dU], graph=True)
            bctx.synchronize()
            for d, r in zip(dU, ref):
                np.testing.assert_array_equal(d.to_host(), r)
            bctx.denoise(imgs[rep], sigmas[rep], None)  # eager work between replays
    bctx.close()


def test_batch_autotune_picks_a_valid_shape_and_keeps_results(ctx):
    imgs, sigmas = _batch_inputs()
    ref = [ctx.denoise(I, s, None)[0] for I, s in zip(imgs, sigmas)]
    import gbm3d_amd

    bctx = gbm3d_amd.Context(0)
    out = bctx.denoise_batch(imgs, sigmas, None, graph=True, tune=True)
    info = bctx.batch_info()
    assert info["streams"] in (1, 2, 4, 8) and info["match_waves"] in (1, 2, 4) and info["tuned_ms"] > 0
    for a, b in zip(out, ref):
        np.testing.assert_array_equal(a, b)
    bctx.close()


def test_parallel_spin_lanes_match_serial_run(ctx):
    """Match spins of one image run on parallel streams; under stage timing they
    run serially on the context stream.  Both give the same bits."""
    _, noisy = synth.make_pair(301, 192, 160, 25.0)
    opts = {"blockSize": 8, "blockSizeWie": 8, "matchSpinsWie": [0, 4]}
    U_par, o_par = ctx.denoise(noisy, 25.0, opts)
    ctx.set_timing(True)
    try:
        U_ser, o_ser = ctx.denoise(noisy, 25.0, opts)
    finally:
        ctx.set_timing(False)
    assert np.array_equal(U_par, U_ser) and np.array_equal(o_par["U1"], o_ser["U1"])


@pytest.mark.parametrize("shape,seed", [((40, 24, 16), 5), ((64, 64, 8), 6)])
def test_median_filter_bit_exact(ctx, shape, seed):
    """medfilt3(Volume, [3 3 3]) is a selection: bit-exact vs the oracle,
    including heavy ties (integer values) and the volume edges."""
    rng = np.random.default_rng(seed)
    V = rng.normal(100, 40, shape).astype(np.float32)
    V[: shape[0] // 2] = np.round(V[: shape[0] // 2] / 20) * 20
    out = ctx.median_filter(V)
    ref = O.median3(V, precision=32)
    assert np.array_equal(out, ref)


@pytest.mark.parametrize("m,n,sigma,seed", [(128, 96, 25.0, 7), (96, 136, 15.0, 8)])
def test_median_branch_end_to_end_vs_oracle(ctx, m, n, sigma, seed):
    """filtType 'median' (denoiseAndAggregateGPU.m:49-52) through the whole path."""
    opts = {"blockSize": 8, "blockSizeWie": 8, "matchSpinsWie": [0, 4], "filtType": "median"}
    clean, noisy = synth.make_pair(seed, m, n, sigma)
    U, out = ctx.denoise(noisy, sigma, opts)
    U32, U132 = O.gbm3d(noisy, sigma, opts, precision=32)
    assert np.max(np.abs(out["U1"] - U132)) <= 1e-3
    U2 = O.step2_from_pilot(out["U1"], noisy, sigma, opts)
    assert np.max(np.abs(U - U2)) <= 1e-3
    U64, _ = O.gbm3d(noisy, sigma, opts, precision=64)
    assert abs(synth.psnr(clean, U) - synth.psnr(clean, U64)) <= 0.01


@pytest.mark.parametrize("m,n,opts", [
    (600, 520, {"blockSize": 8, "blockSizeWie": 8, "matchSpinsWie": [0, 4]}),
    (560, 300, None),
])
def test_distributed_vs_oracle(ctx, m, n, opts):
    """GBM3D_distributed (source/GBM3D_distributed.m): tiles of 256 + blockSize,
    per-case step 1 / step 2, ramp-window stitching.  Chained like the
    single-image test: U2 vs the FP32 restatement, U vs the restatement's step 2
    on the GPU's own U2, PSNR vs the FP64 restatement."""
    clean, noisy = synth.make_pair(400 + m, m, n, 25.0)
    U, out = ctx.denoise_distributed(noisy, 25.0, opts)
    U32, U232 = O.gbm3d_distributed(noisy, 25.0, opts, precision=32)
    assert np.max(np.abs(out["U2"] - U232)) <= 1e-3
    Uc, _ = O.gbm3d_distributed(noisy, 25.0, opts, precision=32, pilot=out["U2"])
    assert np.max(np.abs(U - Uc)) <= 1e-3
    U64, _ = O.gbm3d_distributed(noisy, 25.0, opts, precision=64)
    assert abs(synth.psnr(clean, U) - synth.psnr(clean, U64)) <= 0.01


def test_distributed_single_tile_and_small_images_are_gbm3d(ctx):
    opts = {"blockSize": 8, "blockSizeWie": 8, "matchSpinsWie": [0, 4]}
    _, noisy = synth.make_pair(90, 264, 264, 25.0)      # exactly one tile
    U, out = ctx.denoise_distributed(noisy, 25.0, opts)
    V, o1 = ctx.denoise(noisy, 25.0, opts)
    assert np.max(np.abs(U - V)) <= 1e-4 and np.max(np.abs(out["U2"] - o1["U1"])) <= 1e-4
    _, small = synth.make_pair(91, 200, 300, 25.0)      # below 256 + blockSize: GBM3D itself
    U, out = ctx.denoise_distributed(small, 25.0, opts)
    V, o1 = ctx.denoise(small, 25.0, opts)
    assert np.array_equal(U, V) and np.array_equal(out["U2"], o1["U1"])


def test_distributed_graph_replays_match_eager_run(ctx):
    """The distributed driver's cases are captured into a graph and replayed;
    replays interleaved with other work all equal the eager (stage-timed, serial) run."""
    opts = {"blockSize": 8, "blockSizeWie": 8, "matchSpinsWie": [0, 4]}
    _, noisy = synth.make_pair(92, 544, 288, 20.0)
    _, other = synth.make_pair(93, 96, 96, 20.0)
    ctx.set_timing(True)
    try:
        U_eager, o_eager = ctx.denoise_distributed(noisy, 20.0, opts)
    finally:
        ctx.set_timing(False)
    for _ in range(3):
        U, out = ctx.denoise_distributed(noisy, 20.0, opts)
        assert np.array_equal(U, U_eager) and np.array_equal(out["U2"], o_eager["U2"])
        ctx.denoise(other, 20.0, opts)


def test_colored_noise_vs_oracle(ctx):
    """Matrix sigma (source/GBM3D.m:29-36): getColoredTau thresholds and a
    per-coefficient Wiener sigma (the caller's getColoredWie output)."""
    m, n = 96, 80
    clean, _ = synth.make_pair(501, m, n, 0.0)
    rng = np.random.default_rng(502)
    fy, fx = np.fft.fftfreq(m)[:, None], np.fft.fftfreq(n)[None, :]
    ghat = np.exp(-(fy ** 2 + fx ** 2) / (2 * 0.15 ** 2))            # low-pass noise colouring
    white = rng.normal(0, 30.0, (m, n))
    noisy = (clean + np.real(np.fft.ifft2(np.fft.fft2(white) * ghat))).astype(np.float32)
    psd = 30.0 ** 2 * np.abs(ghat) ** 2
    sigma_wie = np.asfortranarray(np.linspace(4, 12, 64).reshape(8, 8))
    opts = {"blockSize": 8, "blockSizeWie": 8, "matchSpinsWie": [0, 4]}
    U, out = ctx.denoise_colored(noisy, psd, sigma_wie, opts)
    U32, U132 = O.gbm3d_colored(noisy, psd, sigma_wie, opts, precision=32)
    assert np.max(np.abs(out["U1"] - U132)) <= 1e-3
    U64, _ = O.gbm3d_colored(noisy, psd, sigma_wie, opts, precision=64)
    # the end-to-end PSNR bar of the other chained tests: within 0.01 dB of the
    # FP32 restatement, and no further from the FP64 oracle than it is + 0.01 dB
    # (a U1 within round-off of the restatement's can still flip near-tied
    # step-2 matches on a 96 x 80 image, where 0.01 dB is 0.2 % of the MSE)
    p_gpu, p32, p64 = synth.psnr(clean, U), synth.psnr(clean, U32), synth.psnr(clean, U64)
    assert abs(p_gpu - p32) <= 0.01
    assert abs(p_gpu - p64) <= abs(p32 - p64) + 0.01
    U2 = O.step2_from_pilot(out["U1"], noisy, 1.0, opts, sigma_wie=sigma_wie)   # chained step 2
    assert np.max(np.abs(U - U2)) <= 1e-3
    # no sigmaWie: the library computes getColoredWie(sigmaPSD, blockSizeWie)
    # itself (source/GBM3D.m:35), the same as passing its result
    sw = gbm3d_amd.getColoredWie(psd, 8)
    Ua, _ = ctx.denoise_colored(noisy, psd, None, opts)
    Ub, _ = ctx.denoise_colored(noisy, psd, sw, opts)
    assert np.array_equal(Ua, Ub)
