"""The end-to-end parity bar shared by the GPU tests: a GPU denoise chained
against the FP32 restatement and the FP64 oracle, margin-aware in step 1."""
import numpy as np

import oracle as O
from gbm3d_amd import synth

NEAR = 1e-3   # coefficient margin of the near-threshold mask (absolute, [0, 255] scale)


def chained_vs_oracle(ctx, clean, noisy, sigma, opts, check_step2=True):
    """Returns the GPU's U.

    U1 against the FP32 restatement outside the pixels that a step-1
    coefficient within NEAR of its threshold can move (oracle.u1_near_mask:
    two FP32 operation orders may decide such a coefficient either way, which
    moves whole groups by units); U against the restatement's step 2 on the
    GPU's own U1 everywhere.  PSNR: within 0.01 dB of the FP32 restatement
    (the reference computes in single precision), and no further from the FP64
    oracle than the FP32 restatement is, plus 0.01 dB.  An image that is not a
    block multiple is chained on its mirror-padded form (myMirrorPadImage, the
    pipeline's own first step), whose U1 holds the full pilot that step 2
    reads; its crop must be the plain call's U bit for bit."""
    m, n = noisy.shape
    bs = O.resolve_opts(opts)["blockSize"]
    M, N = -(-m // bs) * bs, -(-n // bs) * bs
    U, out = ctx.denoise(noisy, sigma, opts)
    Ip = O.mirror_pad(noisy, M, N).astype(np.float32) if (M, N) != (m, n) else noisy
    if Ip is not noisy:
        Up, outp = ctx.denoise(Ip, sigma, opts)
        assert np.array_equal(Up[:m, :n], U) and np.array_equal(outp["U1"][:m, :n], out["U1"])
    else:
        Up, outp = U, out
    U32, U132 = O.gbm3d(Ip, sigma, opts, precision=32)
    mask, ncoef = O.u1_near_mask(Ip, sigma, opts, NEAR)
    d1 = np.abs(outp["U1"] - U132)
    e1 = float(np.max(np.where(mask, 0.0, d1)))
    Uc = O.step2_from_pilot(outp["U1"], Ip, sigma, opts) if check_step2 else Up
    e2 = float(np.max(np.abs(Up - Uc)))
    U64, _ = O.gbm3d(noisy, sigma, opts, precision=64)
    p_gpu, p32, p64 = synth.psnr(clean, U), synth.psnr(clean, U32[:m, :n]), synth.psnr(clean, U64)
    print(f"{noisy.shape} sigma {sigma}: U1 vs FP32 {e1:.2e} outside {mask.mean():.3f} masked ({ncoef} coefficients; "
          f"{float(d1.max()):.2e} inside), U vs chained step 2 {e2:.2e}, PSNR gpu {p_gpu:.4f} fp32 {p32:.4f} "
          f"fp64 {p64:.4f}")
    assert e1 <= 1e-3 and e2 <= 1e-3
    assert mask.mean() < 0.25
    assert abs(p_gpu - p32) <= 0.01
    assert abs(p_gpu - p64) <= abs(p32 - p64) + 0.01
    return U
