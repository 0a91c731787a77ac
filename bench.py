"""Benchmark: GBM3D two-step denoise of a 2048x2048 FP32 image, sigma = 25.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

A step is one full two-step GBM3D denoise (source/GBM3D.m) of one synthetic
2048^2 image whose input is already resident in HBM.  Prints one JSON line.
The per-kernel roofline figure uses the library's own HIP-event stage timers
(recorded on the stream the kernels run on).  The CPU baseline is the
repository's FP32 C++ restatement (the oracle, OpenMP) on a 512^2 sample of
the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "Mpix/s two-step denoise, sigma=25, 2048^2: plain vs chosen form; dPSNR (dB)"
SIZE = 2048
SIGMA = 25.0
SEED = 2048
# 8x8 blocks in both steps, K = 16, matchSize 32 (radius 16), spins [0 4]/[0 4],
# cycleSpin 2, 3 levels, tau = sigma*[3.3 3.0 2.7]
OPTS = {"blockSize": 8, "blockSizeWie": 8, "matchSpinsWie": [0, 4]}
PEAK_HBM_GBS = 8000.0
PEAK_FP32_TFLOPS = 157.3          # MI355X FP32 with packed (dual) FP32 instructions
PEAK_FP32_VALU_TFLOPS = 78.6      # FP32 FMA without packed instructions (the build does not emit them)
PEAK_BF16_TFLOPS = 2500.0
# learned shrinkage: 4 conv layers x 2 FLOP x 16 positions x 64 out x 192 in, per group
SHRINK_FLOP_PER_GROUP = 4 * 2 * 16 * 64 * 192
DPSNR_BOUND = -0.05   # a form is eligible as the chosen form when PSNR >= plain - 0.05 dB


def kernel_rooflines(stage_ms, stage_n, M, N, K, d1, d2, n1, n2, bw=8):
    """Algorithmic work per launch of each single-kernel stage (d1: step-1
    block size, bw: step-2 block size, d2: match size)."""
    MN = M * N
    vol_bytes = MN * K * 4
    pad_bytes = (M + d2) * (N + d2) * 4
    G2 = MN // (bw * bw)
    out = {}
    flops_match = 2.0 * MN * d2 * d2  # |x|^2+|y|^2-2<x,y>: d^2 FMAs per candidate, d2^2 candidates per d^2 pixels
    # the FP32 matcher runs plain FP32 FMAs on the VALU: bound by their issue rate
    if stage_n["match1"]:
        out["match1"] = dict(bound="valu", unit="TFLOP/s", work=flops_match, launches=stage_n["match1"],
                             ms=stage_ms["match1"] / stage_n["match1"], peak=PEAK_FP32_VALU_TFLOPS)
    if stage_n["match2"]:
        out["match2"] = dict(bound="valu", unit="TFLOP/s", work=flops_match, launches=stage_n["match2"],
                             ms=stage_ms["match2"] / stage_n["match2"], peak=PEAK_FP32_VALU_TFLOPS)
    if stage_n["wiener"]:
        # groups gathered straight from the padded pilot and noisy search images
        # through S (V1, V2 never materialised): two pads + S in, V2f + gains out
        out["wiener"] = dict(bound="hbm", unit="GB/s", work=float(2 * pad_bytes + G2 * K * 4 + vol_bytes + G2 * 4),
                             launches=stage_n["wiener"], ms=stage_ms["wiener"] / stage_n["wiener"], peak=PEAK_HBM_GBS)
    if stage_n["agg2"]:
        out["agg2"] = dict(bound="hbm", unit="GB/s", work=float(vol_bytes + G2 * K * 4 + G2 * 4 + 2 * MN * 4),
                           launches=stage_n["agg2"], ms=stage_ms["agg2"] / stage_n["agg2"], peak=PEAK_HBM_GBS)
    if stage_n["ht_syn0"]:
        # finest-level threshold + synthesis of both cycle spins (at 8x8 blocks
        # with the step-1 weights, which read nothing more): read the mosaic
        # volume, write Vavg (the coarse level read is 1/8 of it)
        out["ht_syn0"] = dict(bound="hbm", unit="GB/s", work=2.0 * vol_bytes, launches=stage_n["ht_syn0"],
                              ms=stage_ms["ht_syn0"] / stage_n["ht_syn0"], peak=PEAK_HBM_GBS)
    if stage_n["agg1_prep"]:
        # read the group-contiguous Vavg once, write one weight per group
        G1 = MN // (d1 * d1)
        out["agg1_prep"] = dict(bound="hbm", unit="GB/s", work=float(vol_bytes + G1 * 4), launches=stage_n["agg1_prep"],
                                ms=stage_ms["agg1_prep"] / stage_n["agg1_prep"], peak=PEAK_HBM_GBS)
    if stage_n.get("shrink"):
        # fused learned shrinkage of one spin: the network's BF16 MFMA work
        G1 = MN // (d1 * d1)
        out["shrink"] = dict(bound="mfma", unit="TFLOP/s", work=float(G1 * SHRINK_FLOP_PER_GROUP),
                             launches=stage_n["shrink"], ms=stage_ms["shrink"] / stage_n["shrink"],
                             peak=PEAK_BF16_TFLOPS)
    if stage_n["agg1"]:
        G1 = MN // (d1 * d1)
        out["agg1"] = dict(bound="hbm", unit="GB/s", work=float(vol_bytes + G1 * K * 4 + G1 * 4 + 2 * MN * 4),
                           launches=stage_n["agg1"], ms=stage_ms["agg1"] / stage_n["agg1"], peak=PEAK_HBM_GBS)
    for k, v in out.items():
        s = v["ms"] * 1e-3
        v["achieved"] = v["work"] / s / (1e12 if v["unit"] == "TFLOP/s" else 1e9)
        v["frac"] = v["achieved"] / v["peak"]
    return out


def host_cores():
    """The host cores this process may run on (its CPU affinity set) and the
    CPU model."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return n, model


def cpu_baseline():
    import oracle
    from gbm3d_amd import synth

    threads, model = host_cores()
    sz = SIZE  # the timed workload itself: about 10 s on 16 host threads
    clean, noisy = synth.make_pair(SEED, sz, sz, SIGMA)
    oracle.set_threads(threads)
    t0 = time.perf_counter()
    U, _ = oracle.gbm3d(noisy, SIGMA, OPTS, precision=32, nthreads=threads)
    dt = time.perf_counter() - t0
    return {"value": sz * sz / dt / 1e6, "unit": "Mpix/s", "cores": threads, "kind": "port",
            "cpu_model": model, "os_cpu_count": os.cpu_count(),
            "sample": f"FP32 C++ restatement (OpenMP, one thread per core of this process's affinity set) of the "
                      f"same two-step GBM3D on the same {sz}x{sz} synthetic image (seed {SEED}), sigma=25, same "
                      f"options; {dt:.2f} s",
            "psnr_db": synth.psnr(clean, U)}


def precision_report(args):
    """4096x4096, sigma 25, K=16, matchSize 48 (radius 24): BF16-MFMA distances
    with FP32 rescoring vs the FP32 path; errors against an FP64 oracle run."""
    import gbm3d_amd
    import oracle
    from gbm3d_amd import synth

    size, opts = args.size or 4096, {"blockSize": 8, "blockSizeWie": 8, "matchSize": 48, "matchSpinsWie": [0, 4]}
    clean, noisy = synth.make_pair(4096, size, size, SIGMA)
    c32, cbf = gbm3d_amd.Context(0), gbm3d_amd.Context(0)
    cbf.set_variant(match_bf16=True)
    out = {"workload": f"precision: {size}x{size}, sigma 25, K 16, matchSize 48, 8x8 blocks", "size": size}
    _, o32 = c32.denoise(noisy, SIGMA, opts)
    # index lists of every matching call (steps 1 and 2, both match spins)
    out["index_tiles_per_call"] = (size // 8) ** 2
    out["index_identical_frac"] = bf16_index_agreement(cbf, noisy, o32["U1"], opts)
    cbf.set_variant(match_bf16=True, match_stats=True)
    cbf.bf16_match_stats(reset=True)
    cbf.denoise(noisy, SIGMA, opts)
    out["bf16_match_paths"] = cbf.bf16_match_stats(reset=True)
    cbf.set_variant(match_bf16=True)
    times = {}
    for name, c in (("fp32", c32), ("bf16_rescored", cbf)):
        c.denoise(noisy, SIGMA, opts)
        c.set_timing(True)
        t0 = time.perf_counter()
        U, o1 = c.denoise(noisy, SIGMA, opts)
        dt = time.perf_counter() - t0
        st = c.stage_times()
        times[name] = (U, dt, st)
    t0 = time.perf_counter()
    nth, _ = host_cores()
    oracle.set_threads(nth)
    U64, _ = oracle.gbm3d(noisy, SIGMA, opts, precision=64, nthreads=nth)
    out["fp64_oracle_s"] = round(time.perf_counter() - t0, 2)
    out["bf16_max_abs_vs_fp32"] = float(np.abs(times["bf16_rescored"][0] - times["fp32"][0]).max())
    for name, (U, dt, st) in times.items():
        dev_ms = sum(v[0] for v in st.values())
        mlaunch = st["match1"][1] + st["match2"][1]
        mms = st["match1"][0] + st["match2"][0]
        # distance work 2*m*n*d2^2 per launch (the useful FMAs; the band GEMM issues more)
        tflops = 2.0 * size * size * 48 * 48 * mlaunch / (mms * 1e-3) / 1e12
        peak = PEAK_BF16_TFLOPS if name == "bf16_rescored" else PEAK_FP32_VALU_TFLOPS
        out[name] = {"psnr_db": round(synth.psnr(clean, U), 4), "psnr_vs_fp64_oracle_db": round(synth.psnr(U64, U), 3),
                     "max_abs_vs_fp64_oracle": float(np.abs(U - U64).max()),
                     "match_ms": round(mms, 3), "device_ms": round(dev_ms, 3),
                     "mpix_s_device": round(size * size / (dev_ms * 1e-3) / 1e6, 2),
                     "match_kernel": {"launches": mlaunch, "avg_launch_ms": round(mms / mlaunch, 4),
                                      "achieved_tflops": round(tflops, 2), "peak_tflops": peak,
                                      "frac": round(tflops / peak, 4)}}
    out["fp64_oracle_psnr_db"] = round(synth.psnr(clean, U64), 4)
    print(json.dumps(out))


def bf16_index_agreement(ctx, noisy, pilot, opts=None):
    """Index lists of every matching call of the timed workload (step 1 on the
    noisy image, step 2 on the pilot, match spins 0 and 4), BF16 form against
    the FP32 matcher: the fraction of reference tiles whose lists are equal."""
    import gbm3d_amd

    ro = gbm3d_amd.resolve_opts(opts or OPTS)
    c32 = gbm3d_amd.Context(0)
    out = {}
    try:
        for step, spins in ((1, ro["matchSpins"]), (2, ro["matchSpinsWie"])):
            for sp in spins:
                I = np.roll(noisy, (sp, sp), (0, 1)).astype(np.float32)
                if step == 1:
                    a, _ = c32.match(I, ro["blockSize"], ro["matchSize"], ro["numMax"], want_volume=False)
                    b, _ = ctx.match(I, ro["blockSize"], ro["matchSize"], ro["numMax"], want_volume=False)
                else:
                    P = np.roll(pilot, (sp, sp), (0, 1)).astype(np.float32)
                    a, _, _ = c32.match_wie(P, I, ro["blockSizeWie"], ro["matchSize"], ro["numMax"], False)
                    b, _, _ = ctx.match_wie(P, I, ro["blockSizeWie"], ro["matchSize"], ro["numMax"], False)
                out[f"step{step}_spin{sp}"] = float(np.mean(np.all(a == b, axis=0)))
    finally:
        c32.close()
    return out


def batch_workload(seed=64, count=64):
    """64 synthetic images: side m ~ U[256,1536], aspect n/m ~ U[0.5,2] (n kept in
    [256,1536]), sigma ~ U[5,50]; the same construction as the single image."""
    from gbm3d_amd import synth

    imgs, cleans, sig = [], [], []
    key = 0xBA7C4 + seed
    for i in range(count):
        u = synth.uniform(key, np.arange(3 * i, 3 * i + 3))
        m = int(256 + u[0] * (1536 - 256))
        n = int(np.clip(round(m * (0.5 + 1.5 * u[1])), 256, 1536))
        s = float(5 + 45 * u[2])
        c, x = synth.make_pair(seed * 1000 + i, m, n, s)
        imgs.append(x)
        cleans.append(c)
        sig.append(s)
    return imgs, cleans, sig


def batched_report(args):
    """One graph launch for 64 uneven images vs the plain per-image runs."""
    import gbm3d_amd
    from gbm3d_amd import device, synth

    imgs, cleans, sig = batch_workload()
    mp = sum(x.size for x in imgs) / 1e6
    ctx = gbm3d_amd.Context(0)
    stream = device.Stream()
    ctx.set_stream(stream.handle)
    dI = [device.DeviceArray.from_host(x) for x in imgs]
    dU = [device.DeviceArray(x.shape) for x in imgs]
    ms_, ns_ = [x.shape[0] for x in imgs], [x.shape[1] for x in imgs]
    e0, e1 = device.Event(), device.Event()

    def per_image():
        for i in range(len(imgs)):
            ctx.denoise_device(dI[i].ptr, ms_[i], ns_[i], sig[i], OPTS, dU[i].ptr)

    per_image()
    stream.synchronize()
    e0.record(stream)
    for _ in range(args.steps):
        per_image()
    e1.record(stream)
    t_plain = e0.elapsed_ms(e1) / args.steps
    U_plain = [d.to_host() for d in dU]
    t0 = time.perf_counter()
    ctx.denoise_batch_device([d.ptr for d in dI], ms_, ns_, sig, OPTS, [d.ptr for d in dU], graph=True, tune=True)
    stream.synchronize()
    tune_s = time.perf_counter() - t0
    info = ctx.batch_info()
    e0.record(stream)
    for _ in range(args.steps):
        ctx.denoise_batch_device([d.ptr for d in dI], ms_, ns_, sig, OPTS, [d.ptr for d in dU], graph=True,
                                 tune=True)
    e1.record(stream)
    t_batch = e0.elapsed_ms(e1) / args.steps
    U_b = [d.to_host() for d in dU]
    p_plain = [synth.psnr(c, u) for c, u in zip(cleans, U_plain)]
    p_b = [synth.psnr(c, u) for c, u in zip(cleans, U_b)]
    print(json.dumps({
        "workload": "batched: 64 synthetic images, sides U[256,1536], aspect U[0.5,2], sigma U[5,50], 8x8 blocks",
        "images": len(imgs), "mpix": round(mp, 3),
        "plain_per_image": {"ms": round(t_plain, 3), "mpix_s": round(mp / (t_plain * 1e-3), 2)},
        "batched_graph": {"ms": round(t_batch, 3), "mpix_s": round(mp / (t_batch * 1e-3), 2), "tuned": info,
                          "tune_and_capture_s": round(tune_s, 3)},
        "speedup": round(t_plain / t_batch, 3),
        "psnr_mean_db": round(float(np.mean(p_b)), 4),
        "psnr_max_abs_diff_vs_plain_db": float(np.max(np.abs(np.array(p_b) - np.array(p_plain)))),
        "bitwise_equal_to_plain": all(np.array_equal(a, b) for a, b in zip(U_b, U_plain))}))


def learned_workload():
    """64 synthetic 512x512 images, sigma ~ U[5, 50] -> (noisy, clean, sigmas)."""
    from gbm3d_amd import synth

    count, size = 64, 512
    u = synth.uniform(0x5EED5, np.arange(count))
    sig = [float(5 + 45 * x) for x in u]
    pairs = [synth.make_pair(5000 + i, size, size, sig[i]) for i in range(count)]
    return [p[1] for p in pairs], [p[0] for p in pairs], sig


def learned_report(args):
    """64 synthetic 512x512 images, sigma ~ U[5,50]: one graph-captured batch
    launch with the learned shrinkage replacing the step-1 hard threshold, vs
    the same batch on the plain path (the judge of its PSNR); errors of both
    against FP64 oracle runs of the same pipelines on image 0."""
    import gbm3d_amd
    import oracle
    from gbm3d_amd import device, shrink, synth
    from oracle import shrink_oracle as so

    imgs, cleans, sig = learned_workload()
    count, size = len(imgs), imgs[0].shape[0]
    mp = count * size * size / 1e6
    ctx = gbm3d_amd.Context(0)
    stream = device.Stream()
    ctx.set_stream(stream.handle)
    dI = [device.DeviceArray.from_host(x) for x in imgs]
    dU = [device.DeviceArray(x.shape) for x in imgs]
    dU1 = [device.DeviceArray(x.shape) for x in imgs]
    e0, e1 = device.Event(), device.Event()
    res = {}
    for name, kw in (("plain", {}), ("learned", {"learned": True})):
        ctx.set_variant(**kw)

        def run():
            ctx.denoise_batch_device([d.ptr for d in dI], [size] * count, [size] * count, sig, OPTS,
                                     [d.ptr for d in dU], [d.ptr for d in dU1], graph=True, tune=True)

        t0 = time.perf_counter()
        run()
        stream.synchronize()
        tune_s = time.perf_counter() - t0
        for _ in range(args.warmup):
            run()
        e0.record(stream)
        for _ in range(args.steps):
            run()
        e1.record(stream)
        stream.synchronize()
        ms = e0.elapsed_ms(e1) / args.steps
        U = [d.to_host() for d in dU]
        U1 = [d.to_host() for d in dU1]
        r = {"ms": round(ms, 3), "mpix_s": round(mp / (ms * 1e-3), 2), "tune_and_capture_s": round(tune_s, 3),
             "tuned": ctx.batch_info(), "psnr": [synth.psnr(c, x) for c, x in zip(cleans, U)],
             "psnr_pilot": [synth.psnr(c, x) for c, x in zip(cleans, U1)], "U0": U[0], "U10": U1[0]}
        if kw.get("learned"):
            st = [ctx.shrink_stats(i) for i in range(count)]
            r["fell_back_spins"] = int(sum(x[2] for x in st))
            r["removed_energy_vs_noise"] = [round(x[0], 4) for x in st]
        res[name] = r
    ctx.set_variant()
    pp, pl = np.array(res["plain"]["psnr"]), np.array(res["learned"]["psnr"])
    qp, ql = np.array(res["plain"]["psnr_pilot"]), np.array(res["learned"]["psnr_pilot"])
    # FP64 oracle runs of the same pipelines on image 0
    params = shrink.load_weights()
    nth, _ = host_cores()
    oracle.set_threads(nth)
    U64, U164 = oracle.gbm3d(imgs[0], sig[0], OPTS, precision=64, nthreads=nth)
    L164 = so.learned_pilot(imgs[0].astype(np.float64), sig[0], params, OPTS, precision=64, exact=True)
    L64 = oracle.step2_from_pilot(L164, imgs[0], sig[0], OPTS, precision=64)
    err = {"plain": {"max_abs_U": float(np.abs(res["plain"]["U0"] - U64).max()),
                     "max_abs_U1": float(np.abs(res["plain"]["U10"] - U164).max()),
                     "dpsnr_vs_fp64_db": round(synth.psnr(cleans[0], res["plain"]["U0"]) - synth.psnr(cleans[0], U64),
                                               4)},
           "learned": {"max_abs_U": float(np.abs(res["learned"]["U0"] - L64).max()),
                       "max_abs_U1": float(np.abs(res["learned"]["U10"] - L164).max()),
                       "mean_abs_U1": float(np.abs(res["learned"]["U10"] - L164).mean()),
                       "dpsnr_vs_fp64_db": round(synth.psnr(cleans[0], res["learned"]["U0"])
                                                 - synth.psnr(cleans[0], L64), 4),
                       "note": "FP64 oracle network without BF16 rounding"}}
    meta = shrink.weights_meta()
    print(json.dumps({
        "workload": "learned shrinkage: 64 synthetic 512x512 images, sigma U[5,50], 8x8 blocks, one graph launch",
        "images": count, "mpix": round(mp, 3),
        "plain_batched": {k: res["plain"][k] for k in ("ms", "mpix_s", "tuned")},
        "learned_batched": {k: res["learned"][k] for k in ("ms", "mpix_s", "tuned", "fell_back_spins")},
        "speedup": round(res["plain"]["ms"] / res["learned"]["ms"], 3),
        "psnr_plain_mean_db": round(float(pp.mean()), 4), "psnr_learned_mean_db": round(float(pl.mean()), 4),
        "dpsnr_mean_db": round(float((pl - pp).mean()), 4), "dpsnr_min_db": round(float((pl - pp).min()), 4),
        "dpsnr_max_db": round(float((pl - pp).max()), 4),
        "pilot_psnr_plain_mean_db": round(float(qp.mean()), 4), "pilot_psnr_learned_mean_db": round(float(ql.mean()), 4),
        "images_within_0.05db": int(np.sum(pl - pp >= DPSNR_BOUND)),
        "vs_fp64_oracle_image0": err,
        "training": {k: meta.get(k) for k in ("images", "groups_per_image", "size", "sigma", "epochs", "batch",
                                              "baseline_val_loss")} | {"val_loss": meta["history"][-1]["val_loss"]},
    }))


def distributed_report(args):
    """GBM3D_distributed (source/GBM3D_distributed.m) on the 2048^2 sigma 25
    image: tiles of 256 + blockSize on parallel streams, ramp stitching; vs the
    single-image path, and vs the FP64 restatement on a 600x520 crop."""
    import gbm3d_amd
    import oracle
    from gbm3d_amd import device, synth

    clean, noisy = synth.make_pair(SEED, SIZE, SIZE, SIGMA)
    ctx = gbm3d_amd.Context(0)
    stream = device.Stream()
    ctx.set_stream(stream.handle)
    dI = device.DeviceArray.from_host(noisy)
    dU = device.DeviceArray((SIZE, SIZE))
    e0, e1 = device.Event(), device.Event()
    res = {}
    for name, fn in (("single_image", ctx.denoise_device), ("distributed", ctx.denoise_distributed_device)):
        def run():
            fn(dI.ptr, SIZE, SIZE, SIGMA, OPTS, dU.ptr)
        for _ in range(args.warmup):
            run()
        e0.record(stream)
        for _ in range(args.steps):
            run()
        e1.record(stream)
        stream.synchronize()
        ms = e0.elapsed_ms(e1) / args.steps
        res[name] = {"ms": round(ms, 3), "mpix_s": round(SIZE * SIZE / (ms * 1e-3) / 1e6, 2),
                     "psnr_db": round(synth.psnr(clean, dU.to_host()), 4)}
    c2, n2 = synth.make_pair(600, 600, 520, SIGMA)
    U, _ = ctx.denoise_distributed(n2, SIGMA, OPTS)
    U64, _ = oracle.gbm3d_distributed(n2, SIGMA, OPTS, precision=64)
    print(json.dumps({
        "workload": "GBM3D_distributed, 2048x2048 synthetic image, sigma 25, 8x8 blocks, tiles of 264",
        "cases_per_step": len(oracle.distribution(SIZE, SIZE, [0, 4], 8)),
        **res,
        "crop_600x520_vs_fp64_oracle": {"max_abs": float(np.abs(U - U64).max()),
                                        "dpsnr_db": round(synth.psnr(c2, U) - synth.psnr(c2, U64), 4)}}))


def global_match_report(args):
    """matchBlocksGPU (source/matchBlocksGPU.m:18-91) on a size x size synthetic
    image (default 512), sigma 25 (the sym4 prefilter runs), 8 x 8 blocks,
    numMax = numMin = 16, through the host entry point (upload, prefilter,
    norms, search, download).  The search's algorithmic work is 2 d^2 flops
    per (tile, position) pair, G = (size / 8)^2 tiles times size^2 positions;
    divided by the whole call's time it is a lower bound on the search
    kernel's rate (rocprofv3 --kernel-trace gives the kernel's own time)."""
    import gbm3d_amd
    from gbm3d_amd import synth

    size, d, K = args.size or 512, 8, 16
    _, noisy = synth.make_pair(SEED, size, size, SIGMA)
    ctx = gbm3d_amd.Context(0)
    for _ in range(args.warmup):
        ctx.match_global(noisy, SIGMA, d, K)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.match_global(noisy, SIGMA, d, K)
    dt = (time.perf_counter() - t0) / args.steps
    G = (size // d) ** 2
    pairs = G * size * size
    print(json.dumps({"workload": f"matchBlocksGPU, {size}x{size} synthetic image, sigma 25, 8x8 blocks, numMax 16",
                      "ms_per_call": round(dt * 1e3, 3), "tile_position_pairs": pairs,
                      "pairs_per_s": round(pairs / dt, 1),
                      "search_tflops_lower_bound": round(2 * d * d * pairs / dt / 1e12, 3),
                      "peak_fp32_valu_tflops": PEAK_FP32_VALU_TFLOPS}))


# kernel symbol prefix -> roofline key; the largest dispatch of a kernel is the
# one measured (the finest level for the wavelet synthesis)
TRAFFIC_KERNELS = {"void gbm3d::ht_fine_kernel<2": "ht_syn0", "void gbm3d::match_kernel": "match2",
                   "void gbm3d::wiener_kernel": "wiener", "void gbm3d::aggregate_batched_kernel<1": "agg1",
                   "void gbm3d::aggregate_batched_kernel<2": "agg2", "void gbm3d::agg1_prep_wave_kernel": "agg1_prep"}


def traffic_from(fetch_dir, write_dir, copy_mib, out_path):
    """HBM bytes per launch from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs
    (KiB per dispatch), each scaled so the copy kernel of copy_mib MiB reads
    and writes exactly copy_mib MiB."""
    import csv
    import glob

    def load(d, counter):
        rows = []
        for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            rows += [r for r in csv.DictReader(open(f)) if r["Counter_Name"] == counter]
        return rows

    res = {}
    for counter, d in (("FETCH_SIZE", fetch_dir), ("WRITE_SIZE", write_dir)):
        rows = load(d, counter)
        copy = [float(r["Counter_Value"]) for r in rows if "stream_copy_kernel" in r["Kernel_Name"]]
        scale = copy_mib * 1024.0 / copy[0]
        res[counter + "_scale"] = scale
        for pre, key in TRAFFIC_KERNELS.items():
            sel = [r for r in rows if r["Kernel_Name"].startswith(pre)]
            if not sel:
                continue
            gmax = max(int(r["Grid_Size"]) for r in sel)
            vals = [float(r["Counter_Value"]) for r in sel if int(r["Grid_Size"]) == gmax]
            res.setdefault(key, {})[counter] = float(np.mean(vals)) * 1024.0 * scale
    out = {k: v.get("FETCH_SIZE", 0.0) + v.get("WRITE_SIZE", 0.0) for k, v in res.items() if isinstance(v, dict)}
    out["_detail"] = res
    json.dump(out, open(out_path, "w"), indent=1)
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--profile-only", action="store_true", help="warmup + steps only (for rocprofv3 runs)")
    ap.add_argument("--stage-timing", action="store_true",
                    help="with --profile-only: run the steps with the stage timers on (match spins in series), "
                         "the schedule whose kernel durations the roofline figures use")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="with --profile-only: also run one device copy of MIB MiB (byte-counter calibration)")
    ap.add_argument("--form", default="plain", choices=["plain", "bf16_match", "learned"],
                    help="with --profile-only: the form whose steps run (for kernel profiles)")
    ap.add_argument("--workload", default="plain", choices=["plain", "precision", "batched", "learned",
                                                          "distributed", "global_match"])
    ap.add_argument("--size", type=int, default=0, help="image side for --workload precision / global_match")
    ap.add_argument("--traffic-from", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="derive profiles/traffic.json from two rocprofv3 --pmc runs made with --calibrate-copy 256")
    args = ap.parse_args()
    if args.traffic_from:
        traffic_from(args.traffic_from[0], args.traffic_from[1], 256, os.path.join(HERE, "profiles", "traffic.json"))
        return
    if args.workload == "precision":
        precision_report(args)
        return
    if args.workload == "batched":
        batched_report(args)
        return
    if args.workload == "learned":
        learned_report(args)
        return
    if args.workload == "distributed":
        distributed_report(args)
        return
    if args.workload == "global_match":
        global_match_report(args)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only (one image per launch); --gpus must be 1", file=sys.stderr)
        sys.exit(2)

    import gbm3d_amd
    from gbm3d_amd import device, synth

    clean, noisy = synth.make_pair(SEED, SIZE, SIZE, SIGMA)
    ctx = gbm3d_amd.Context(0)
    stream = device.Stream()
    ctx.set_stream(stream.handle)
    dI = device.DeviceArray.from_host(noisy.astype(np.float32))
    dU = device.DeviceArray((SIZE, SIZE))
    dU1 = device.DeviceArray((SIZE, SIZE))
    e0, e1 = device.Event(), device.Event()

    cur = {"dI": dI, "sigma": SIGMA}  # the resident input the steps denoise

    def step():
        ctx.denoise_device(cur["dI"].ptr, SIZE, SIZE, cur["sigma"], OPTS, dU.ptr, dU1.ptr)

    def timed(steps):
        for _ in range(args.warmup):
            step()
        stream.synchronize()
        device.synchronize()
        t0 = time.perf_counter()
        e0.record(stream)
        for _ in range(steps):
            step()
        e1.record(stream)
        stream.synchronize()
        device.synchronize()
        return e0.elapsed_ms(e1) / steps, time.perf_counter() - t0

    if args.profile_only and args.form != "plain":
        ctx.set_variant(**{"bf16_match": {"match_bf16": True}, "learned": {"learned": True}}[args.form])
    if args.profile_only and args.stage_timing:
        ctx.set_timing(True)  # the stage-timed (serial) schedule the roofline figures come from
    ms, wall = timed(args.steps)
    if args.profile_only:
        if args.calibrate_copy:
            nb = args.calibrate_copy << 20
            a, b = device.DeviceArray((nb // 4,)), device.DeviceArray((nb // 4,))
            ctx.device_copy(b.ptr, a.ptr, nb)
            stream.synchronize()
        print(json.dumps({"ms_per_step": ms}))
        return
    ro = gbm3d_amd.resolve_opts(OPTS)

    def staged(reps=8, warm=8):
        """Steps with the stage timers on (same work, outside the timed region):
        `warm` steps first (the serial schedule's kernels take ~8 steps to
        settle: the finest-level kernel runs 15 % slower in the first), then
        per-stage ms per step and launches per step averaged over `reps` steps."""
        ctx.set_timing(True)
        for _ in range(warm):
            step()
        tot_ms, tot_n = {}, {}
        for _ in range(reps):
            step()
            for k, (v, n) in ctx.stage_times().items():
                tot_ms[k] = tot_ms.get(k, 0.0) + v
                tot_n[k] = tot_n.get(k, 0) + n
        ctx.set_timing(False)
        stage_ms = {k: v / reps for k, v in tot_ms.items()}
        stage_n = {k: v // reps if v % reps == 0 else v / reps for k, v in tot_n.items()}
        rl = kernel_rooflines(stage_ms, stage_n, SIZE, SIZE, ro["numMax"], ro["blockSize"], ro["matchSize"],
                              len(ro["matchSpins"]), len(ro["matchSpinsWie"]), ro["blockSizeWie"])
        return stage_ms, rl

    U, U1 = dU.to_host(), dU1.to_host()
    psnr = synth.psnr(clean, U)
    stage_ms, rl = staged()
    forms = {"plain": {"ms": ms, "wall": wall, "psnr": psnr, "pilot": synth.psnr(clean, U1), "stage_ms": stage_ms,
                       "rl": rl, "U": U}}

    # forms beside the plain path, timed the same way on the same resident input
    variants = {}
    for name, kw in (("bf16_match", {"match_bf16": True}), ("learned", {"learned": True}),
                     ("learned_bf16_match", {"learned": True, "match_bf16": True})):
        ctx.set_variant(**kw)
        msv, wallv = timed(args.steps)
        Uv, U1v = dU.to_host(), dU1.to_host()
        smv, rlv = staged()
        pv = synth.psnr(clean, Uv)
        forms[name] = {"ms": msv, "wall": wallv, "psnr": pv, "pilot": synth.psnr(clean, U1v), "stage_ms": smv,
                       "rl": rlv, "U": Uv}
        variants[name] = {"mpix_s": round(SIZE * SIZE / (msv * 1e-3) / 1e6, 3), "ms_per_step": round(msv, 4),
                          "psnr_db": round(pv, 4), "dpsnr_db": round(pv - psnr, 4),
                          "psnr_pilot_db": round(forms[name]["pilot"], 4),
                          "max_abs_vs_plain": float(np.abs(Uv - U).max()),
                          "stages_ms": {k: round(v, 4) for k, v in smv.items() if v > 0}}
        if kw.get("match_bf16") and "match1" in rlv:
            # the distance kernel against the BF16 MFMA peak: the useful distance
            # work 2*m*n*d2^2 per launch, and the MFMA work it issues (three
            # 16x16x32 MFMAs per 16x16 candidate tile and block column: hi*hi,
            # hi*lo, lo*hi)
            k = rlv["match1"]
            d1 = ro["blockSize"]
            issued = 3.0 * 16 * 16 * 32 * 2 * (ro["matchSize"] // 16) ** 2 * d1 * (SIZE // d1) ** 2
            variants[name]["match_kernel"] = {
                "achieved_tflops": round(k["achieved"], 2), "peak": PEAK_BF16_TFLOPS,
                "frac": round(k["achieved"] / PEAK_BF16_TFLOPS, 4), "avg_launch_ms": round(k["ms"], 4),
                "mfma_issued_tflops": round(issued / (k["ms"] * 1e-3) / 1e12, 2),
                "mfma_issued_frac": round(issued / (k["ms"] * 1e-3) / 1e12 / PEAK_BF16_TFLOPS, 4)}
            variants[name]["index_agreement"] = bf16_index_agreement(ctx, noisy, U1)
            ctx.set_variant(**kw, match_stats=True)
            ctx.bf16_match_stats(reset=True)
            step()
            variants[name]["match_paths"] = ctx.bf16_match_stats(reset=True)
            ctx.set_variant(**kw)
        if kw.get("learned"):
            el, eh, fell = ctx.shrink_stats(0)
            variants[name].update({"removed_energy_vs_noise": round(el, 4),
                                   "hard_threshold_removed_energy_vs_noise": round(eh, 4), "fell_back_spins": fell})
            if "shrink" in rlv:
                k = rlv["shrink"]
                variants[name]["shrink_kernel"] = {"achieved_tflops": round(k["achieved"], 2), "peak": k["peak"],
                                                   "frac": round(k["frac"], 4), "avg_launch_ms": round(k["ms"], 4)}
    ctx.set_variant()

    # the plain path at the other noise levels of the 2048^2 configuration
    # (same image construction and seed; inputs resident, timed the same way)
    sigma_sweep = {}
    for sg in (10.0, 25.0, 50.0):
        if sg == SIGMA:
            msg, psg, png = ms, psnr, synth.psnr(clean, noisy)
        else:
            cl, no = synth.make_pair(SEED, SIZE, SIZE, sg)
            cur["dI"], cur["sigma"] = device.DeviceArray.from_host(no.astype(np.float32)), sg
            msg, _ = timed(args.steps)
            psg, png = synth.psnr(cl, dU.to_host()), synth.psnr(cl, no)
        sigma_sweep[f"{sg:g}"] = {"mpix_s": round(SIZE * SIZE / (msg * 1e-3) / 1e6, 3), "ms_per_step": round(msg, 4),
                                  "psnr_db": round(psg, 4), "psnr_noisy_db": round(png, 4)}
    cur["dI"], cur["sigma"] = dI, SIGMA

    # chosen form: the fastest whose PSNR stays within DPSNR_BOUND of the plain path
    eligible = [n for n, f in forms.items() if f["psnr"] - psnr >= DPSNR_BOUND]
    chosen = min(eligible, key=lambda n: forms[n]["ms"])
    cf = forms[chosen]
    rlc = cf["rl"]
    dom = max(rlc, key=lambda k: rlc[k]["ms"] * rlc[k]["launches"])
    d = rlc[dom]
    prof = os.path.join(HERE, "profiles", "traffic.json")
    traffic = json.load(open(prof)).get(dom) if os.path.exists(prof) else None
    roofline = {"bound": d["bound"], "achieved": round(d["achieved"], 3), "peak": d["peak"], "unit": d["unit"],
                "frac": round(d["frac"], 4), "traffic": traffic, "kernel": dom, "avg_launch_ms": round(d["ms"], 5)}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline()

    mpix = SIZE * SIZE / (cf["ms"] * 1e-3) / 1e6
    plain_mpix = SIZE * SIZE / (ms * 1e-3) / 1e6
    line = {
        "metric": METRIC, "value": round(mpix, 3), "unit": "Mpix/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(cf["ms"], 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "GBM3D two-step denoise, one 2048x2048 FP32 synthetic image (seed 2048), sigma=25, "
                               "8x8 blocks both steps, K=16, matchSize 32, spins [0 4]/[0 4], cycleSpin 2, 3 levels",
                   "form": chosen, "chosen_by": f"fastest form with PSNR >= plain {DPSNR_BOUND} dB",
                   "size": SIZE, "sigma": SIGMA},
        "plain_mpix_s": round(plain_mpix, 3), "plain_ms_per_step": round(ms, 4),
        "speedup_vs_plain": round(ms / cf["ms"], 4),
        "psnr_db": round(cf["psnr"], 4), "psnr_plain_db": round(psnr, 4),
        "dpsnr_db": round(cf["psnr"] - psnr, 4),
        "psnr_noisy_db": round(synth.psnr(clean, noisy), 4), "psnr_pilot_db": round(cf["pilot"], 4),
        "wall_s_timed": round(cf["wall"], 4),
        "stages_ms": {k: round(v, 4) for k, v in cf["stage_ms"].items() if v > 0},
        "roofline": roofline,
        "kernels": {k: {kk: (round(vv, 5) if isinstance(vv, float) else vv) for kk, vv in v.items()}
                    for k, v in rlc.items()},
        "variants": variants,
        "sigma_sweep": sigma_sweep,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
