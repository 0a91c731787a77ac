"""The benchmark workloads at their full sizes against the oracle.

  * 2048 x 2048, 8 x 8 blocks in both steps, sigma 10 / 25 / 50 (the timed
    workload): U1 against the FP32 restatement outside near-threshold pixels,
    U against the restatement's step 2 on the GPU's own U1, PSNR against the
    FP32 restatement and the FP64 oracle.
  * 64 images of uneven sizes and sigmas in one graph-captured batch: equal to
    the per-image runs bit for bit; the smallest, the largest and the most
    elongated image against the restatement as above.
  * 4096 x 4096, matchSize 48: the FP32 and the BF16 matcher's index lists
    and blocks against the FP32 restatement on the noisy image (step 1) and on
    the pilot (step 2), match spins 0 and 4; the BF16 form end to end equal to
    the FP32 path bit for bit.
  * 64 images of 512 x 512 with the learned shrinkage in one graph-captured
    batch: two pilots against the learned-pilot restatement, the device
    judge's flags, and its fallback equal to the plain batch bit for bit.
Bars: 1e-3 max abs on the [0, 255] scale, 0.01 dB (see _chained_vs_oracle).
"""
import numpy as np
import pytest

import gbm3d_amd
import oracle as O
from gbm3d_amd import synth

pytestmark = pytest.mark.gpu

P8 = {"blockSize": 8, "blockSizeWie": 8, "matchSpinsWie": [0, 4]}


from chain import chained_vs_oracle as _chained_vs_oracle  # noqa: E402


@pytest.mark.parametrize("sigma", [10.0, 25.0, 50.0])
def test_benchmark_workload_2048_vs_oracle(ctx, sigma):
    clean, noisy = synth.make_pair(2048, 2048, 2048, sigma)
    _chained_vs_oracle(ctx, clean, noisy, sigma, P8)


@pytest.fixture(scope="module")
def batch():
    import bench

    return bench.batch_workload()


def test_batched_64_images_equal_per_image_runs(ctx, batch):
    imgs, cleans, sig = batch
    assert len(imgs) == 64
    single = [ctx.denoise(x, s, P8)[0] for x, s in zip(imgs, sig)]
    for graph, tune in ((True, True), (True, False), (False, False)):
        out = ctx.denoise_batch(imgs, sig, P8, graph=graph, tune=tune)
        for i, (a, b) in enumerate(zip(out, single)):
            assert np.array_equal(a, b), (graph, tune, i)


@pytest.mark.parametrize("which", ["smallest", "largest", "elongated"])
def test_batched_images_vs_oracle(ctx, batch, which):
    imgs, cleans, sig = batch
    area = [x.size for x in imgs]
    aspect = [max(x.shape) / min(x.shape) for x in imgs]
    i = {"smallest": int(np.argmin(area)), "largest": int(np.argmax(area)),
         "elongated": int(np.argmax(aspect))}[which]
    out = ctx.denoise_batch([imgs[i]], [sig[i]], P8)[0]
    U = _chained_vs_oracle(ctx, cleans[i], imgs[i], sig[i], P8)
    assert np.array_equal(out, U)


P48 = {"blockSize": 8, "blockSizeWie": 8, "matchSize": 48, "matchSpinsWie": [0, 4]}


@pytest.fixture(scope="module")
def precision_case(ctx):
    clean, noisy = synth.make_pair(4096, 4096, 4096, 25.0)
    _, out = ctx.denoise(noisy, 25.0, P48)
    return noisy, out["U1"]


@pytest.fixture(scope="module")
def cbf():
    c = gbm3d_amd.Context(0)
    c.set_variant(match_bf16=True)
    yield c
    c.close()


@pytest.mark.parametrize("step,spin", [(1, 0), (1, 4), (2, 0), (2, 4)])
def test_precision_4096_index_lists_vs_fp32_restatement(ctx, cbf, precision_case, step, spin):
    noisy, U1 = precision_case
    In = np.roll(noisy, (spin, spin), (0, 1))
    if step == 1:
        So, Vo, _ = O.match(In, 8, 48, 16, precision=32)
        for c in (ctx, cbf):
            S, V = c.match(In, 8, 48, 16)
            np.testing.assert_array_equal(S, So)
            np.testing.assert_array_equal(V, Vo)
    else:
        P = np.roll(U1, (spin, spin), (0, 1)).astype(np.float32)
        So, _, _, _ = O.match_wie(P, In, 8, 48, 16, precision=32)
        for c in (ctx, cbf):
            S, _, _ = c.match_wie(P, In, 8, 48, 16, want_groups=False)
            np.testing.assert_array_equal(S, So)


def test_precision_4096_bf16_form_equals_fp32_path(ctx, cbf, precision_case):
    noisy, U1 = precision_case
    Ubf, obf = cbf.denoise(noisy, 25.0, P48)
    U32, o32 = ctx.denoise(noisy, 25.0, P48)
    np.testing.assert_array_equal(obf["U1"], U1)
    np.testing.assert_array_equal(Ubf, U32)


@pytest.fixture(scope="module")
def learned_case():
    import bench
    from gbm3d_amd import shrink

    imgs, cleans, sig = bench.learned_workload()
    c = gbm3d_amd.Context(0)
    c.set_variant(learned=True)
    U, U1 = _batch_with_pilots(c, imgs, sig)
    stats = [c.shrink_stats(i) for i in range(len(imgs))]
    yield imgs, cleans, sig, U, U1, stats, c, shrink.load_weights()
    c.close()


def _batch_with_pilots(c, imgs, sig):
    from gbm3d_amd import device

    dI = [device.DeviceArray.from_host(x) for x in imgs]
    dU = [device.DeviceArray(x.shape) for x in imgs]
    dU1 = [device.DeviceArray(x.shape) for x in imgs]
    c.denoise_batch_device([d.ptr for d in dI], [x.shape[0] for x in imgs], [x.shape[1] for x in imgs], sig, P8,
                           [d.ptr for d in dU], [d.ptr for d in dU1], graph=True, tune=False)
    c.synchronize()
    return [d.to_host() for d in dU], [d.to_host() for d in dU1]


@pytest.mark.parametrize("which", ["lowest_sigma", "highest_sigma"])
def test_learned_64x512_batch_vs_restatement(learned_case, which):
    """64 x 512^2 learned inference in one graph-captured batch: the pilot of
    the images with the lowest and the highest sigma against the learned-pilot
    restatement (oracle.shrink_oracle.learned_pilot, FP32 stages, BF16
    network), PSNR within 0.01 dB and mean abs within 1e-3; U against the
    restatement's step 2 on the GPU's pilot; no spin fell back at the
    default bound (removed energy between 0.5 and 1 noise energies)."""
    from oracle import shrink_oracle as so

    imgs, cleans, sig, U, U1, stats, c, params = learned_case
    i = int(np.argmin(sig)) if which == "lowest_sigma" else int(np.argmax(sig))
    ref = so.learned_pilot(imgs[i].astype(np.float64), sig[i], params, P8, precision=32)
    err = np.abs(U1[i] - ref)
    print(f"image {i} sigma {sig[i]:.2f}: pilot PSNR gpu {synth.psnr(cleans[i], U1[i]):.4f} "
          f"restatement {synth.psnr(cleans[i], ref):.4f}, mean abs {err.mean():.2e}")
    assert abs(synth.psnr(cleans[i], U1[i]) - synth.psnr(cleans[i], ref)) <= 0.01
    assert err.mean() <= 1e-3
    assert np.max(np.abs(U[i] - O.step2_from_pilot(U1[i], imgs[i], sig[i], P8))) <= 1e-3
    assert all(f == 0 and 0.5 < el < 1.0 for el, _, f in stats)


def test_learned_64x512_fallback_equals_plain_batch(ctx, learned_case):
    """With a bound below every spin's removed energy the device judge makes
    every spin fall back: the batch equals the plain batch bit for bit."""
    imgs, cleans, sig, U, U1, stats, c, params = learned_case
    Up, U1p = _batch_with_pilots(ctx, imgs, sig)
    try:
        c.set_variant(learned=True, learned_bound=1e-3, weights=params)
        Uf, U1f = _batch_with_pilots(c, imgs, sig)
        fell = [c.shrink_stats(i)[2] for i in range(len(imgs))]
    finally:
        c.set_variant(learned=True, weights=params)
    assert fell == [len(gbm3d_amd.resolve_opts(P8)["matchSpins"])] * len(imgs)
    for i in range(len(imgs)):
        assert np.array_equal(Uf[i], Up[i]) and np.array_equal(U1f[i], U1p[i]), i
