"""Writes the golden fixtures tests/golden/*.npz with the FP64 oracle.

Each fixture holds the noisy input (float32, the exact bytes the library
receives), the clean image, sigma, the resolved options and the FP64 oracle's
U, U1, the step-1 index sets of match spin 0 (S1) and the FP64 distances'
tie margin of those sets.  Run from the repository root:
    python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from gbm3d_amd import synth  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))

CASES = {
    "p_def_64": dict(seed=101, m=64, n=64, sigma=25.0, opts=None),
    "p_8_64": dict(seed=102, m=64, n=64, sigma=25.0, opts={"blockSize": 8, "matchSpinsWie": [0, 4]}),
    "fast_72x90": dict(seed=103, m=72, n=90, sigma=10.0, opts={"profile": "fast"}),
    "p_def_128_s50": dict(seed=104, m=128, n=128, sigma=50.0, opts=None),
}


def tie_margin(dist, S, K):
    """Smallest gap between consecutive ranks 1..K+1 of any tile (normalised distance)."""
    d = np.sort(dist, axis=0)[: K + 1]
    return float(np.diff(d, axis=0).min())


def main():
    for name, c in CASES.items():
        clean, noisy = synth.make_pair(c["seed"], c["m"], c["n"], c["sigma"])
        ro = O.resolve_opts(c["opts"])
        U, U1 = O.gbm3d(noisy, c["sigma"], c["opts"], precision=64)
        bs = ro["blockSize"]
        M = -(-c["m"] // bs) * bs
        N = -(-c["n"] // bs) * bs
        Ip = O.mirror_pad(noisy.astype(np.float64), M, N)
        S1, _, dist = O.match(Ip, bs, ro["matchSize"], ro["numMax"], precision=64, want_dist=True)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), noisy=noisy, clean=clean, sigma=c["sigma"],
                            opts=json.dumps(c["opts"]), resolved=json.dumps(ro), U=U, U1=U1, S1=S1,
                            S1_margin=tie_margin(dist, S1, ro["numMax"]))
        print(name, "psnr", synth.psnr(clean, U), "noisy", synth.psnr(clean, noisy))


if __name__ == "__main__":
    main()
