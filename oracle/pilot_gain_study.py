"""Study (oracle only, CPU): the final PSNR of step 2 run on pilots built from
*oracle* gains on the step-1 groups -- what a perfect learned shrinkage would
output -- against the plain path (global wavelet hard threshold).

TEST INFRASTRUCTURE: uses the oracle and the clean image; not part of the
product.    python -m oracle.pilot_gain_study

Gains on the DCT-Haar coefficients y of each step-1 group (clean c):
  mmse       c^2 / (c^2 + sigma^2)       (the Wiener target)
  energy     min(|c| / |y|, 1)           (pilot magnitude |c|: step 2 then
                                          computes the oracle Wiener gain)
  ht2.7      |y| > 2.7 sigma             (a hard threshold per group)
  sqrt_mmse  sqrt(c^2 / (c^2 + sigma^2))
Prints, per sigma, (pilot PSNR, final PSNR) of the plain path and of each rule.
"""
from __future__ import annotations

import numpy as np

import oracle as O
from gbm3d_amd import synth
from oracle import shrink_oracle as so

OPTS = {"blockSize": 8, "blockSizeWie": 8, "matchSpinsWie": [0, 4]}

GAINS = {
    "mmse": lambda y, c, s: c ** 2 / (c ** 2 + s ** 2),
    "energy": lambda y, c, s: np.minimum(np.abs(c) / np.maximum(np.abs(y), 1e-9), 1.0),
    "ht2.7": lambda y, c, s: (np.abs(y) > 2.7 * s).astype(float),
    "sqrt_mmse": lambda y, c, s: np.sqrt(c ** 2 / (c ** 2 + s ** 2)),
}


def pilot_with(I, Ic, sigma, gainfn):
    """Step 1 with the group filter replaced by gain rule gainfn (cf.
    shrink_oracle.learned_pilot): match, gains on the DCT-Haar coefficients,
    inverse, clip, the reference aggregation, spins summed."""
    ro = O.resolve_opts(OPTS)
    M, N = I.shape
    Us, Ws = np.zeros((M, N)), np.zeros((M, N))
    for sp in ro["matchSpins"]:
        Is, Ics = np.roll(I, (sp, sp), (0, 1)), np.roll(Ic, (sp, sp), (0, 1))
        S, _, _ = O.match(Is, 8, 32, 16, precision=64)
        cn = so.coefficients(so.gather_groups(O.search_pad(Is, 32), S))
        cc = so.coefficients(so.gather_groups(O.search_pad(Ics, 32), S))
        g = gainfn(cn, cc, sigma)
        g[:, 0, 14:] = 1.0
        Vf = np.maximum(so.inverse(cn * g), 0)
        Mt, Nt = M // 8, N // 8
        mosaic = Vf.reshape(8, 8, 16, Mt, Nt, order="F").transpose(3, 0, 4, 1, 2).reshape(M, N, 16)
        U, W = O.aggregate_ht(mosaic, S, 8, 32, precision=64)
        Us += np.roll(U, (-sp, -sp), (0, 1))
        Ws += np.roll(W, (-sp, -sp), (0, 1))
    return Us / Ws


def main(size: int = 256) -> dict:
    out = {}
    for sigma in (10.0, 25.0, 50.0):
        clean, noisy = synth.make_pair(9000 + int(sigma), size, size, sigma)
        U, U1 = O.gbm3d(noisy, sigma, OPTS, precision=64)
        res = {"plain": (synth.psnr(clean, U1), synth.psnr(clean, U))}
        for name, fn in GAINS.items():
            P = pilot_with(noisy.astype(np.float64), clean.astype(np.float64), sigma, fn)
            Uf = O.step2_from_pilot(P, noisy, sigma, OPTS, precision=64)
            res[name] = (synth.psnr(clean, P), synth.psnr(clean, Uf))
        out[sigma] = res
        print(sigma, {k: (round(a, 3), round(b, 3)) for k, (a, b) in res.items()}, flush=True)
    return out


if __name__ == "__main__":
    main()
