"""Learned shrinkage of step-1 groups: weights file and GPU training driver.

The network (4 conv layers of width 3 along the 16 group positions, 64
channels = the 8x8 DCT coefficients, sigma/50 input channel, BF16 weights with
FP32 accumulation) and every kernel it runs on live in libgbm3d
(csrc/k_learn.hip); this module only orchestrates them:

  data      seeded synthetic clean/noisy pairs generated on the device
            (gbm3d_synth_pair), sigma ~ U[5, 50]; step-1 matching on the noisy
            image and, for a random subset of groups per image, the features
            c_noisy/sigma, the oracle Wiener gains of the clean group and the
            loss weights (gbm3d_shrink_features);
  training  Adam on sum(Wt (g - g*)^2)/G with the hand-written forward /
            backward GEMM kernels (gbm3d_shrink_grad, gbm3d_shrink_adam).

    python -m gbm3d_amd.shrink --images 20000 --groups 64 --out weights.npz

The trained parameters ship as data/shrink_weights.npz (float32 vector of
GBM3D_SHRINK_PARAMS values plus a JSON description of the run).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

from ._abi import GBM3D_SHRINK_PARAMS

WEIGHTS_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "shrink_weights.npz")

CH, POS, TAPS, NL = 64, 16, 3, 4
KW = TAPS * CH


def _off_w(l: int) -> int:
    return 0 if l == 0 else KW * CH + TAPS * CH + CH + (l - 1) * (KW * CH + CH)


def _off_b(l: int) -> int:
    return KW * CH + TAPS * CH if l == 0 else _off_w(l) + KW * CH


def init_params(seed: int = 0) -> np.ndarray:
    """He-scaled initial weights; the last layer starts near gain 1/2."""
    rng = np.random.default_rng(seed)
    p = np.zeros(GBM3D_SHRINK_PARAMS, dtype=np.float32)
    for l in range(NL):
        std = np.sqrt(2.0 / KW) if l < NL - 1 else np.sqrt(0.5 / KW)
        p[_off_w(l):_off_w(l) + KW * CH] = rng.normal(0.0, std, KW * CH)
    p[KW * CH:KW * CH + TAPS * CH] = rng.normal(0.0, 0.1, TAPS * CH)
    return p


def save_weights(path: str, params, meta: dict) -> None:
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    np.savez(path, params=np.asarray(params, dtype=np.float32), meta=np.array(json.dumps(meta)))


def load_weights(path: str | None = None) -> np.ndarray:
    p = path or WEIGHTS_PATH
    if not os.path.exists(p):
        raise FileNotFoundError(f"{p}: no trained shrinkage weights (python -m gbm3d_amd.shrink)")
    with np.load(p, allow_pickle=False) as z:
        params = np.asarray(z["params"], dtype=np.float32)
    if params.size != GBM3D_SHRINK_PARAMS:
        raise ValueError(f"{p}: {params.size} values, expected {GBM3D_SHRINK_PARAMS}")
    return params


def weights_meta(path: str | None = None) -> dict:
    with np.load(path or WEIGHTS_PATH, allow_pickle=False) as z:
        return json.loads(str(z["meta"]))


class Dataset:
    """Training rows of many images resident in device memory."""

    def __init__(self, ctx, images: int, groups: int, size: int = 512, seed: int = 1, smin: float = 5.0,
                 smax: float = 50.0, log=None, target: str = "energy"):
        from .device import DeviceArray

        self.images, self.gpi, self.size = images, groups, size
        G_img = (size // 8) ** 2
        if groups > G_img:
            raise ValueError("more groups per image than the image has")
        rng = np.random.default_rng(seed)
        self.sigmas = rng.uniform(smin, smax, images).astype(np.float32)
        sel = np.stack([np.sort(rng.choice(G_img, groups, replace=False)) for _ in range(images)]).astype(np.uint32)
        self.ngroups = images * groups
        rows = self.ngroups * POS
        self.X = DeviceArray((rows * CH,))
        self.T = DeviceArray((rows * CH,))
        self.Wt = DeviceArray((rows * CH,))
        self.sig = DeviceArray.from_host(np.repeat(self.sigmas, groups))
        dsel = DeviceArray.from_host(sel.ravel())
        clean = DeviceArray((size * size,))
        noisy = DeviceArray((size * size,))
        stride = groups * POS * CH * 4
        for i in range(images):
            ctx.synth_pair(seed * 1_000_003 + i, size, size, float(self.sigmas[i]), clean.ptr, noisy.ptr)
            ctx.shrink_features(clean.ptr, noisy.ptr, size, size, float(self.sigmas[i]), dsel.ptr + 4 * groups * i,
                                groups, self.X.ptr + stride * i, self.T.ptr + stride * i, self.Wt.ptr + stride * i,
                                target=target)
            if log and (i + 1) % 2000 == 0:
                ctx.synchronize()
                log(json.dumps({"images": i + 1}))
        ctx.synchronize()
        self._keep = (dsel, clean, noisy)

    def slice(self, g0: int, count: int):
        """Device pointers of groups [g0, g0 + count)."""
        o = g0 * POS * CH * 4
        return self.X.ptr + o, self.T.ptr + o, self.Wt.ptr + o, self.sig.ptr + 4 * g0


def baseline_losses(val: Dataset) -> dict:
    """Validation loss of fixed shrinkage rules on the same rows: hard
    threshold at 2.7 sigma and the empirical Wiener gain max(x^2 - 1, 0)/x^2."""
    X, T, Wt = val.X.to_host(), val.T.to_host(), val.Wt.to_host()
    x2 = X.astype(np.float64) ** 2
    out = {}
    for name, g in (("hard_threshold_2.7", (x2 > 2.7 ** 2).astype(np.float64)),
                    ("empirical_wiener", np.maximum(x2 - 1.0, 0.0) / np.maximum(x2, 1e-30))):
        out[name] = float(np.sum(Wt * (g - T) ** 2) / val.ngroups)
    return out


def train(ctx, data: Dataset, val: Dataset, epochs: int = 6, batch: int = 2048, lr: float = 1e-3,
          seed: int = 0, log=print):
    """Adam with cosine-decayed step size over contiguous batches in shuffled order."""
    from .device import DeviceArray

    p = DeviceArray.from_host(init_params(seed))
    m = DeviceArray.from_host(np.zeros(GBM3D_SHRINK_PARAMS, np.float32))
    v = DeviceArray.from_host(np.zeros(GBM3D_SHRINK_PARAMS, np.float32))
    grad = DeviceArray((GBM3D_SHRINK_PARAMS,))
    nb = data.ngroups // batch
    total = epochs * nb
    rng = np.random.default_rng(seed)
    step = 0
    hist = []

    def val_loss():
        tot = 0.0
        vb = min(batch, val.ngroups)
        k = val.ngroups // vb
        for i in range(k):
            X, T, W, S = val.slice(i * vb, vb)
            tot += ctx.shrink_grad(p.ptr, X, T, W, S, vb, grad.ptr)
        return tot / k

    t0 = time.time()
    for ep in range(epochs):
        run = 0.0
        for b in rng.permutation(nb):
            step += 1
            lr_t = 1e-5 + 0.5 * (lr - 1e-5) * (1 + np.cos(np.pi * (step - 1) / total))
            X, T, W, S = data.slice(int(b) * batch, batch)
            run += ctx.shrink_grad(p.ptr, X, T, W, S, batch, grad.ptr)
            ctx.shrink_adam(p.ptr, m.ptr, v.ptr, grad.ptr, step, float(lr_t))
            if step % 200 == 0:
                log(json.dumps({"step": step, "of": total, "s": round(time.time() - t0, 1)}))
        vl = val_loss()
        hist.append({"epoch": ep + 1, "train_loss": run / nb, "val_loss": vl, "s": round(time.time() - t0, 1)})
        log(json.dumps(hist[-1]))
    ctx.synchronize()
    return p.to_host(), hist


def main(argv=None):
    ap = argparse.ArgumentParser(description="train the step-1 learned shrinkage on the GPU")
    ap.add_argument("--images", type=int, default=20000)
    ap.add_argument("--groups", type=int, default=64, help="groups sampled per image")
    ap.add_argument("--val-images", type=int, default=256)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--epochs", type=int, default=6)
    ap.add_argument("--batch", type=int, default=2048)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default=WEIGHTS_PATH)
    ap.add_argument("--target", default="energy", choices=["energy", "wiener"],
                    help="gain targets: 'energy' min(|c|/|y|, 1), the pilot magnitudes step 2's Wiener gains "
                         "need; 'wiener' c^2/(c^2 + sigma^2)")
    a = ap.parse_args(argv)
    from .api import Context

    ctx = Context(0)
    t0 = time.time()
    data = Dataset(ctx, a.images, a.groups, a.size, seed=1, log=lambda s: print(s, flush=True), target=a.target)
    val = Dataset(ctx, a.val_images, a.groups, a.size, seed=2, target=a.target)
    t_data = time.time() - t0
    print(json.dumps({"data_s": round(t_data, 1), "groups": data.ngroups, "val_groups": val.ngroups}), flush=True)
    base = baseline_losses(val)
    print(json.dumps({"baseline_val_loss": base}), flush=True)
    params, hist = train(ctx, data, val, a.epochs, a.batch, a.lr, a.seed,
                         log=lambda s: print(s, flush=True))
    meta = {"images": a.images, "groups_per_image": a.groups, "size": a.size, "sigma": [5, 50], "target": a.target,
            "val_images": a.val_images, "epochs": a.epochs, "batch": a.batch, "lr": a.lr, "seed": a.seed,
            "data_seconds": round(t_data, 1), "baseline_val_loss": base, "history": hist}
    save_weights(a.out, params, meta)
    print(json.dumps({"saved": a.out, "val_loss": hist[-1]["val_loss"]}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
