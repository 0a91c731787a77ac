"""Study (oracle only, CPU): where the learned pilot loses in step 2.

TEST INFRASTRUCTURE: uses the oracle, the NumPy restatement of the learned
shrinkage and the clean image; not part of the product.
    python -m oracle.pilot_group_study [size]

Step 2 (source/GBM3D.m:70-83) reads its pilot twice: matchBlocksWie groups
the blocks by distances on the pilot (source/matchBlocksWie.m), and the
Wiener gains p^2 / (p^2 + sigma^2) come from the pilot's grouped
coefficients (source/denoiseAndAggregateWieFast.m:49).  This study runs step
2 with the groups matched on one pilot and the gains taken from another
(hard-threshold pilot U1 of the plain path, learned pilot, clean image +
1e-3, the noisy image itself), so the two uses can be told apart, and prints
the final PSNR of every combination plus the mean bias of each pilot's gains
against the gains of the clean image on the same groups; on the synthetic
construction and on a textured image without flat regions.
"""
from __future__ import annotations

import sys

import numpy as np

import oracle as O
from gbm3d_amd import synth
from gbm3d_amd.shrink import load_weights
from oracle import shrink_oracle as so

OPTS = {"blockSize": 8, "blockSizeWie": 8, "matchSpinsWie": [0, 4]}


def step2_mixed(group_pilot, gain_pilot, noisy, sigma):
    """Step 2 with the groups matched on group_pilot and the Wiener gains
    computed from gain_pilot's blocks in those groups."""
    ro = O.resolve_opts(OPTS)
    M, N = noisy.shape
    Us, Ws = np.zeros((M, N)), np.zeros((M, N))
    for sp in ro["matchSpinsWie"]:
        A = np.roll(group_pilot, (sp, sp), (0, 1))
        B = np.roll(gain_pilot, (sp, sp), (0, 1))
        I = np.roll(noisy, (sp, sp), (0, 1))
        S, _, VB, _ = O.match_wie(A, B, 8, 32, 16, precision=64)
        _, _, VN, _ = O.match_wie(A, I, 8, 32, 16, precision=64)
        V2f, g2 = O.wiener(VB, VN, sigma, precision=64)
        U, W = O.aggregate_wie(V2f, g2, S, M, N, 32, precision=64)
        Us += np.roll(U, (-sp, -sp), (0, 1))
        Ws += np.roll(W, (-sp, -sp), (0, 1))
    return Us / Ws


def textured(seed: int, n: int) -> np.ndarray:
    """A clean image without flat regions (two sinusoidal patterns and a
    smooth random field), for contrast with the synthetic construction, whose
    rectangles and disks are exactly flat."""
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:n, 0:n] / n
    img = 128 + 60 * np.sin(2 * np.pi * (3 * x + 2 * y)) + 40 * np.cos(2 * np.pi * 7 * x * y)
    f = np.fft.fft2(rng.standard_normal((n, n)))
    ky, kx = np.meshgrid(np.fft.fftfreq(n), np.fft.fftfreq(n), indexing="ij")
    field = np.real(np.fft.ifft2(f * np.exp(-(kx ** 2 + ky ** 2) / (2 * 0.03 ** 2))))
    return np.clip(img + 30 * field / np.std(field), 0, 255)


def main(size: int = 512) -> dict:
    params = load_weights()
    out = {}
    cases = [("synthetic", 10.0, 2048), ("synthetic", 25.0, 2048), ("synthetic", 25.0, 7), ("synthetic", 50.0, 2048),
             ("textured", 25.0, 5)]
    for kind, sigma, seed in cases:
        if kind == "synthetic":
            clean, noisy = synth.make_pair(seed, size, size, sigma)
        else:
            clean = textured(seed, size)
            noisy = clean + sigma * np.random.default_rng(seed + 1).standard_normal((size, size))
        clean, noisy = clean.astype(np.float64), noisy.astype(np.float64)
        _, U1 = O.gbm3d(noisy, sigma, OPTS, precision=64)
        pilots = {"ht": U1, "learned": so.learned_pilot(noisy, sigma, params, OPTS, precision=64), "clean": clean + 1e-3,
                  "noisy": noisy}
        final = {(a, b): synth.psnr(clean, step2_mixed(pilots[a], pilots[b], noisy, sigma)) for a in pilots for b in pilots}
        _, _, Vc, _ = O.match_wie(U1, clean, 8, 32, 16, precision=64)
        cc = so.coefficients(Vc)
        wc = cc ** 2 / (cc ** 2 + sigma ** 2)
        bias = {}
        for name in ("ht", "learned"):
            _, _, Vp, _ = O.match_wie(U1, pilots[name], 8, 32, 16, precision=64)
            cp = so.coefficients(Vp)
            bias[name] = float(np.mean(cp ** 2 / (cp ** 2 + sigma ** 2) - wc))
        res = {"pilot_psnr": {k: synth.psnr(clean, v) for k, v in pilots.items()},
               "final_psnr": {f"{a}/{b}": v for (a, b), v in final.items()}, "gain_bias": bias}
        out[(kind, sigma, seed)] = res
        print(f"{kind} sigma {sigma} seed {seed}: pilots", {k: round(v, 3) for k, v in res["pilot_psnr"].items()})
        print("  final PSNR, groups from the row's pilot, gains from the column's:")
        print("  " + " " * 9 + "".join(f"{b:>10s}" for b in pilots))
        for a in pilots:
            print(f"  {a:>9s}" + "".join(f"{final[(a, b)]:10.3f}" for b in pilots))
        print("  mean gain bias against the clean image's gains:", {k: round(v, 4) for k, v in bias.items()}, flush=True)
    return out


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 512)
