"""NumPy restatement of the learned-shrinkage variant (the checker of the GPU
kernels in gbm3d_amd/csrc/k_learn.hip).

TEST INFRASTRUCTURE ONLY: imported by tests/ and bench.py's checks, never by
the product path.

The variant has no reference counterpart; what it borrows from the reference
is pinned against the oracle's own pinned C functions:
  * the group transform is the one of the step-2 Wiener filter
    (source/denoiseAndAggregateWieFast.m:33-47: T = dct(eye(8)) on every
    slice, 3-level Haar along the slices, waveTransforms3D/myWavDec1D.m:10-26)
    -- pinned by reproducing oracle.wiener() from these coefficients
    (tests/test_shrink_oracle.py::test_transform_reproduces_wiener);
  * the target gain |c|^2/(|c|^2 + sigma^2) is the gain formula of
    denoiseAndAggregateWieFast.m:49 on the clean group;
  * BF16 rounding (round to nearest even) is pinned against torch.bfloat16;
  * the backward pass is pinned against central finite differences (FP64,
    rounding off) and Adam's first step against its closed form.

Layouts follow the kernels: a group is an (8, 8, 16) column-major stack,
coefficients c[t, k] with t = u + 8v; network rows are (group, position k),
columns the 64 coefficients.  Parameters: W0[3][64][64] | wsig[3][64] | b0[64]
| (W_l[3][64][64] | b_l[64]) l = 1..3, W[tap][cin][cout], tap 0 reads k-1.
"""
from __future__ import annotations

import numpy as np

from . import haar_dec as _c_haar_dec

CH, POS, TAPS, NL = 64, 16, 3, 4
KW = TAPS * CH
OFF_WSIG = KW * CH
OFF_B0 = OFF_WSIG + TAPS * CH


def off_w(l: int) -> int:
    return 0 if l == 0 else OFF_B0 + CH + (l - 1) * (KW * CH + CH)


def off_b(l: int) -> int:
    return OFF_B0 if l == 0 else off_w(l) + KW * CH


NPARAM = off_b(NL - 1) + CH  # 49600


def unpack(p):
    """params -> (W list of (3, 64, 64), wsig (3, 64), b list of (64,))."""
    p = np.asarray(p)
    W = [p[off_w(l):off_w(l) + KW * CH].reshape(TAPS, CH, CH) for l in range(NL)]
    b = [p[off_b(l):off_b(l) + CH] for l in range(NL)]
    wsig = p[OFF_WSIG:OFF_WSIG + TAPS * CH].reshape(TAPS, CH)
    return W, wsig, b


def bf16(x):
    """Round float32 values to BF16 (nearest even), returned as float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    r = np.where(nan, (u | 0x400000) & 0xFFFF0000, r)
    return r.astype(np.uint32).view(np.float32).reshape(a.shape)


# ------------------------------------------------------------ the transform
def dct_matrix(d: int = 8) -> np.ndarray:
    """T(k, n) = c_k cos(pi (2n + 1) k / (2d)) -- dct(eye(d)) columns."""
    k = np.arange(d)[:, None]
    n = np.arange(d)[None, :]
    T = np.cos(np.pi * (2 * n + 1) * k / (2.0 * d))
    T[0] *= np.sqrt(1.0 / d)
    T[1:] *= np.sqrt(2.0 / d)
    return T


def haar_matrix(K: int = 16, levels: int = 3) -> np.ndarray:
    """Matrix H with H @ x = myWavDec1D(x) (columns from the C oracle)."""
    return np.stack([_c_haar_dec(np.eye(K)[:, j], levels) for j in range(K)], axis=1)


def coefficients(V) -> np.ndarray:
    """Groups (8, 8, 16, G) -> c (G, 64, 16): DCT of every slice, then Haar-K."""
    V = np.asarray(V, dtype=np.float64)
    G = V.size // (8 * 8 * 16)
    V = V.reshape(8, 8, 16, G, order="F")
    T = dct_matrix(8)
    Y = np.einsum("au,uvkg,bv->abkg", T, V, T)          # T X T' per slice
    H = haar_matrix(16, 3)
    C = np.einsum("jk,abkg->abjg", H, Y)                 # Haar along k
    return C.reshape(64, 16, G, order="F").transpose(2, 0, 1)


def inverse(c) -> np.ndarray:
    """c (G, 64, 16) -> groups (8, 8, 16, G)."""
    c = np.asarray(c, dtype=np.float64)
    G = c.shape[0]
    C = c.transpose(1, 2, 0).reshape(8, 8, 16, G, order="F")
    Hi = np.linalg.inv(haar_matrix(16, 3))
    Y = np.einsum("kj,abjg->abkg", Hi, C)
    T = dct_matrix(8)
    return np.einsum("au,abkg,bv->uvkg", T, Y, T)        # T' X T per slice


# --------------------------------------------------------------- network
def _conv_in(H):
    """H (G, 16, Cin) -> im2col (G, 16, 3*Cin) with zero padding (tap t reads k+t-1)."""
    G, P, Cn = H.shape
    Z = np.zeros((G, P + 2, Cn), dtype=H.dtype)
    Z[:, 1:P + 1] = H
    return np.concatenate([Z[:, t:t + P] for t in range(TAPS)], axis=2)


def _sig_taps(wsig, dtype):
    """(16, 64): sum of the sigma-channel weights over the taps inside the group."""
    out = np.zeros((POS, CH), dtype=dtype)
    for pos in range(POS):
        a = wsig[1].astype(dtype).copy()
        if pos >= 1:
            a = a + wsig[0]
        if pos <= POS - 2:
            a = a + wsig[2]
        out[pos] = a
    return out


def forward(params, X, sig, exact=False):
    """Network on features X (G, 16, 64) = c/sigma with sigma per group.

    exact=False mirrors the kernels: BF16 operands, FP32 sums; exact=True is
    FP64 without rounding (finite-difference pin).  Returns (Z list, H list);
    H[-1] are the gains."""
    dt = np.float64 if exact else np.float32
    rnd = (lambda a: a.astype(dt)) if exact else bf16
    W, wsig, b = unpack(np.asarray(params, dtype=dt))
    Hin = np.asarray(X, dtype=dt)
    s = (np.asarray(sig, dtype=dt) * dt(0.02)).astype(dt)
    Zs, Hs = [], []
    for l in range(NL):
        A = rnd(_conv_in(Hin))                              # (G, 16, 192)
        B = rnd(W[l].reshape(KW, CH))
        z = np.matmul(A, B).astype(dt)
        if l == 0:
            z = z + (b[0][None, None, :] + s[:, None, None] * _sig_taps(wsig, dt)[None]).astype(dt)
        else:
            z = z + b[l][None, None, :]
        Zs.append(z)
        Hin = np.maximum(z, 0) if l < NL - 1 else 1.0 / (1.0 + np.exp(-z))
        Hs.append(Hin.astype(dt))
    return Zs, Hs


def loss_and_grad(params, X, T, Wt, sig, exact=False):
    """loss = sum(Wt (g - T)^2) / G and its gradient (same rounding points as
    the kernels: BF16 operands of every product, FP32 sums)."""
    dt = np.float64 if exact else np.float32
    rnd = (lambda a: a.astype(dt)) if exact else bf16
    G = X.shape[0]
    W, wsig, b = unpack(np.asarray(params, dtype=dt))
    Zs, Hs = forward(params, X, sig, exact)
    g = Hs[-1]
    d = g - T
    loss = float(np.sum(np.asarray(Wt, np.float64) * np.asarray(d, np.float64) ** 2) / G)
    dZ = (2.0 / G * Wt * d * g * (1 - g)).astype(dt)
    grad = np.zeros(NPARAM, dtype=dt)
    s = (np.asarray(sig, dtype=dt) * dt(0.02)).astype(dt)
    for l in range(NL - 1, -1, -1):
        Hin = np.asarray(X, dtype=dt) if l == 0 else Hs[l - 1]
        A = rnd(_conv_in(Hin)).reshape(-1, KW)               # (G*16, 192)
        dZr = rnd(dZ).reshape(-1, CH)
        grad[off_w(l):off_w(l) + KW * CH] = (A.T @ dZr).reshape(-1)
        grad[off_b(l):off_b(l) + CH] = rnd(np.ones(1, dt))[0] * dZr.sum(axis=0)
        if l == 0:
            for t in range(TAPS):
                pos = np.arange(POS) + t - 1
                valid = ((pos >= 0) & (pos < POS)).astype(dt)
                sv = rnd(s[:, None] * valid[None, :]).reshape(-1)   # (G*16,)
                grad[OFF_WSIG + t * CH:OFF_WSIG + (t + 1) * CH] = sv @ dZr
            break
        # input gradient through the taps: dH[k] = sum_t dZ[k - t + 1] W[t]^T
        dZp = np.zeros((G, POS + 2, CH), dtype=dt)
        dZp[:, 1:POS + 1] = rnd(dZ)
        dH = np.zeros((G, POS, CH), dtype=dt)
        Wl = rnd(W[l])
        for t in range(TAPS):
            dH = dH + np.matmul(dZp[:, 2 - t:2 - t + POS], Wl[t].T)
        dZ = np.where(Zs[l - 1] > 0, dH, 0).astype(dt)
    return loss, grad


def adam(p, m, v, g, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    lr_t = lr * np.sqrt(1 - b2 ** step) / (1 - b1 ** step)
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    return p - np.float32(lr_t) * m / (np.sqrt(v) + eps), m, v


def init_params(seed: int = 0) -> np.ndarray:
    """He-scaled random parameters (test inputs; the trained ones ship separately)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM, dtype=np.float32)
    for l in range(NL):
        p[off_w(l):off_w(l) + KW * CH] = rng.normal(0, np.sqrt(2.0 / KW), KW * CH)
        p[off_b(l):off_b(l) + CH] = rng.normal(0, 0.05, CH)
    p[OFF_WSIG:OFF_WSIG + TAPS * CH] = rng.normal(0, 0.3, TAPS * CH)
    return p


# ------------------------------------------------------------ whole stage
def shrink_groups(V, sigma, params, exact=False):
    """Filtered groups (8, 8, 16, G) and gains (G, 64, 16), as the fused kernel
    (exact=True: FP64 network without BF16 rounding)."""
    c = coefficients(V)
    G = c.shape[0]
    if exact:
        X = (c / sigma).transpose(0, 2, 1)
        _, Hs = forward(np.asarray(params, np.float64), X, np.full(G, float(sigma)), exact=True)
    else:
        c32 = c.astype(np.float32)
        X = (c32 * np.float32(1.0 / np.float32(sigma))).transpose(0, 2, 1)   # (G, 16, 64)
        _, Hs = forward(params, X, np.full(G, sigma, np.float32))
    g = Hs[-1].transpose(0, 2, 1).astype(np.float64)                     # (G, 64, 16)
    g[:, 0, 14:] = 1.0
    out = inverse(c * g)
    return np.maximum(out, 0.0), g


def features(Vc, Vn, sigma, target="wiener"):
    """Training rows of groups (clean Vc, noisy Vn): X, T, Wt each (G, 16, 64).
    target "wiener": T = c^2/(c^2 + sigma^2), Wt = min(1 + c^2/sigma^2, 100);
    "energy": T = min(|c|/|y|, 1), Wt = min(y^2/sigma^2, 100) (c clean, y noisy)."""
    cc = coefficients(Vc).astype(np.float32)
    cn = coefficients(Vn).astype(np.float32)
    s2 = np.float32(sigma) * np.float32(sigma)
    X = cn * np.float32(1.0 / np.float32(sigma))
    e2 = cc * cc
    if target == "energy":
        ay = np.abs(cn)
        with np.errstate(divide="ignore", invalid="ignore"):
            T = np.where(ay > 0, np.minimum(np.abs(cc) / np.where(ay > 0, ay, 1), 1), 1).astype(np.float32)
        Wt = np.minimum(X * X, 100).astype(np.float32)
    else:
        T = e2 / (e2 + s2)
        Wt = np.minimum(1 + e2 / s2, 100).astype(np.float32)
    Wt[:, 0, 14:] = 0
    tr = (0, 2, 1)
    return X.transpose(tr), T.transpose(tr), Wt.transpose(tr)


def gather_groups(pad, S, d=8, d2=32):
    """Groups (d, d, K, G) of the padded search image through S (1-based window
    indices a + d2*b + 1, source/matchBlocksCPU.m:103-111)."""
    K, Mt, Nt = S.shape
    V = np.zeros((d, d, K, Mt, Nt))
    for tj in range(Nt):
        for ti in range(Mt):
            for k in range(K):
                idx = int(S[k, ti, tj]) - 1
                a, b = idx % d2, idx // d2
                V[:, :, k, ti, tj] = pad[ti * d + a:ti * d + a + d, tj * d + b:tj * d + b + d]
    return V.reshape(d, d, K, Mt * Nt, order="F")


def learned_pilot(I, sigma, params, opts, precision=32, exact=False):
    """Step 1 with the learned shrinkage in place of the hard threshold, composed
    from the oracle's stages (source/GBM3D.m:50-57 with the filter replaced):
    per match spin circshift, matchBlocksCPU, learned shrinkage of every group,
    aggregation with the step-1 weights, shift back; U1 = sum U / sum W.
    I: M x N (block multiple)."""
    from . import aggregate_ht, match, resolve_opts, search_pad

    ro = resolve_opts(opts)
    M, N = I.shape
    Us, Ws = np.zeros((M, N)), np.zeros((M, N))
    for sp in ro["matchSpins"]:
        Is = np.roll(I, (sp, sp), (0, 1))
        S, _, _ = match(Is, 8, 32, 16, precision=precision)
        Vn = gather_groups(search_pad(Is, 32), S)
        Vf, _ = shrink_groups(Vn, sigma, params, exact=exact)
        Mt, Nt = M // 8, N // 8
        mosaic = Vf.reshape(8, 8, 16, Mt, Nt, order="F").transpose(3, 0, 4, 1, 2).reshape(M, N, 16)
        U, W = aggregate_ht(mosaic, S, 8, 32, precision=precision)
        Us += np.roll(U.astype(np.float64), (-sp, -sp), (0, 1))
        Ws += np.roll(W.astype(np.float64), (-sp, -sp), (0, 1))
    return Us / Ws
