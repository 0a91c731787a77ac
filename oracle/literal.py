"""Literal numpy transcription of the reference MATLAB path (FFT formulation).

TEST INFRASTRUCTURE ONLY (a pin for the C++ oracle, not a product path).

The C++ oracle (gbm3d_oracle.hpp) restates the algorithm with direct index
formulas (separable circular filters, direct block sums).  This module follows
the reference's own formulation instead -- fftn of zero-padded filters,
2x2-tiled + circshifted search images, fft2 correlations, ind2sub, 1-based
slicing turned into 0-based -- so that an index, sign or alignment slip in
either formulation makes the two disagree.  Everything runs in float64 and is
meant for small inputs (tens of pixels).

MATLAB -> numpy: A(1:2:end) -> A[::2]; circshift(A,[s,s]) -> np.roll(A,(s,s),(0,1));
fftn(P, siz) zero-pads/truncates like np.fft.fftn(P, s=siz); column-major
linear indices via order='F'.
"""
from __future__ import annotations

import numpy as np
import scipy.fft

CA = 0.016572815184060
CB = 0.121533978016438
CC = np.sqrt(2) / 2


def getWaveFilters3Dbior():
    """waveTransforms3D/getWaveFilters3Dbior.m:12-58 -> Psi, Psi2 (10,10,2,8)."""
    lo = np.array([CA, -CA, -CB, CB, CC, CC, CB, -CB, -CA, CA])
    hi = np.zeros(10)
    hi[4] = -CC
    hi[5] = CC
    lo_d = np.array([CC, CC])
    hi_d = np.array([-CC, CC])

    def build(Lo, Hi):
        P = np.zeros((10, 10, 2, 8))
        combos = [(Lo, Lo, lo_d), (Hi, Lo, lo_d), (Lo, Hi, lo_d), (Lo, Lo, hi_d),
                  (Hi, Hi, lo_d), (Hi, Lo, hi_d), (Lo, Hi, hi_d), (Hi, Hi, hi_d)]
        for j, (a, b, z) in enumerate(combos):
            P[:, :, :, j] = np.outer(a, b)[:, :, None] * z[None, None, :]
        return P

    Psi = build(lo, hi)
    Lo2 = np.abs(hi)
    Hi2 = np.array([-CA, -CA, CB, CB, -CC, CC, -CB, -CB, CA, CA])
    Psi2 = build(Lo2, Hi2)
    return Psi, Psi2


def myWavDec3DFFT(Psi, levels, U):
    """waveTransforms3D/myWavDec3DFFT.m:10-23.  C[i][j] (j = 0..6), C[levels][0]."""
    C = [[None] * 7 for _ in range(levels + 1)]
    for i in range(levels):
        FU = np.fft.fftn(U)
        for j in range(7):
            c = np.fft.ifftn(FU * np.fft.fftn(Psi[:, :, :, j + 1], s=U.shape, axes=(0, 1, 2)))
            C[i][j] = np.real(c[::2, ::2, ::2])
        U = np.real(np.fft.ifftn(FU * np.fft.fftn(Psi[:, :, :, 0], s=U.shape, axes=(0, 1, 2))))
        U = U[::2, ::2, ::2]
    C[levels][0] = U
    return C


def myWavRec3DGPUfast(Psi, levels, C):
    """waveTransforms3D/myWavRec3DGPUfast.m:10-40."""
    Cup = [None] * (levels + 1)
    for i in range(levels):
        sh = tuple(2 * s for s in C[i][0].shape)
        Cup[i] = np.zeros(sh)
        tmp = np.zeros(sh)
        for j in range(7):
            tmp[::2, ::2, ::2] = C[i][j]
            Cup[i] = Cup[i] + np.real(np.fft.ifftn(np.fft.fftn(tmp) * np.conj(np.fft.fftn(Psi[:, :, :, j + 1], s=sh, axes=(0, 1, 2)))))
    Cup[levels] = C[levels][0]
    for i in range(levels, 0, -1):
        tmp = np.zeros(Cup[i - 1].shape)
        tmp[::2, ::2, ::2] = Cup[i]
        FPsi = np.conj(np.fft.fftn(Psi[:, :, :, 0], s=tmp.shape, axes=(0, 1, 2)))
        Cup[i - 1] = Cup[i - 1] + np.real(np.fft.ifftn(np.fft.fftn(tmp) * FPsi))
    return Cup[0]


def getTau(levels, mode):
    """utilities/getTau.m:9-51 (mode 1 only is used on the path)."""
    tau = np.zeros((levels, 7))
    for i in range(levels):
        for j in range(7):
            base = 3.3 - 0.3 * i
            if mode == 1:
                tau[i, j] = base
            elif mode == 2:
                tau[i, j] = base if j < 3 else (base + 0.5 if j < 6 else base + 1)
            elif mode == 3:
                tau[i, j] = base if j < 3 else (base + 0.25 if j < 6 else base + 1.5)
            elif mode == 4:
                tau[i, j] = 2.9 if j < 3 else (2.9 + 0.5 if j < 6 else 2.9 + 1)
    return tau


def ht_filter(Volume, tau, levels, cycleSpin):
    """source/denoiseAndAggregateGPU.m:19-47 ('ht' branch)."""
    Psi, Psi2 = getWaveFilters3Dbior()
    Vrec = []
    for spin in range(cycleSpin):
        V = np.roll(Volume, (spin, spin, spin), axis=(0, 1, 2))
        C = myWavDec3DFFT(Psi, levels, V)
        for i in range(levels):
            for j in range(7):
                C[i][j] = C[i][j] * (np.abs(C[i][j]) > tau[i, j])
        R = myWavRec3DGPUfast(Psi2, levels, C)
        Vrec.append(np.roll(R, (-spin, -spin, -spin), axis=(0, 1, 2)))
    return np.maximum(np.mean(np.stack(Vrec, axis=3), axis=3), 0)


def _ind2sub(d2, idx1):
    """[a,b] = ind2sub([d2,d2], idx) (1-based in and out)."""
    i0 = int(idx1) - 1
    return i0 % d2 + 1, i0 // d2 + 1


def _mink_idx(x, k):
    """[~, idx] = mink(x, k): ascending, NaN last, ties by position (1-based)."""
    key = np.where(np.isnan(x), np.inf, x)
    nan_last = np.isnan(x).astype(np.int8)
    order = np.lexsort((np.arange(x.size), key, nan_last))
    return order[:k] + 1


def _search_pad(I, d2):
    m, n = I.shape
    Ipad = np.zeros((2 * m, 2 * n))
    Ipad[:m, :n] = I
    Ipad[:m, n:] = I
    Ipad[m:, :n] = I
    Ipad[m:, n:] = I
    Ipad = Ipad[:m + d2, :n + d2]
    return np.roll(Ipad, (d2 // 2, d2 // 2), axis=(0, 1))


def _corr_fft(a, b_shape_src, shape):
    """real(ifft2(conj(fft2(a, shape)) .* fft2(b)))."""
    return np.real(np.fft.ifft2(np.conj(np.fft.fft2(a, s=shape)) * np.fft.fft2(b_shape_src)))


def matchBlocksCPU(I, d, d2, K):
    """source/matchBlocksCPU.m:25-111 (numMax == numMin).  S: list of lists."""
    m, n = I.shape
    M, N = m // d, n // d
    Ipad = _search_pad(I, d2)
    g = np.ones((d, d))
    blockNorms = _corr_fft(g, I ** 2, (m, n))
    blockNormsPad = _corr_fft(g, Ipad ** 2, (m + d2, n + d2))
    Volume = np.zeros((m, n, K))
    S = [[None] * N for _ in range(M)]
    dists = {}
    for i in range(1, M + 1):
        indI = (i - 1) * d
        for j in range(1, N + 1):
            indJ = (j - 1) * d
            xR = I[(i - 1) * d:i * d, (j - 1) * d:j * d]
            Itmp = Ipad[indI:indI + d + d2, indJ:indJ + d + d2]
            bd = _corr_fft(xR, Itmp, (d + d2, d + d2))
            nb = blockNorms[(i - 1) * d, (j - 1) * d]
            bd = (nb + blockNormsPad[indI:indI + d2, indJ:indJ + d2] - 2 * bd[:d2, :d2]) / nb
            dists[(i - 1, j - 1)] = bd
            s = _mink_idx(bd.flatten(order="F"), K)
            s[0] = d2 * d2 // 2 + d2 // 2 + 1
            S[i - 1][j - 1] = s
            for k in range(K):
                a, b = _ind2sub(d2, s[k])
                Volume[(i - 1) * d:i * d, (j - 1) * d:j * d, k] = Itmp[a - 1:a - 1 + d, b - 1:b - 1 + d]
    return S, Volume, dists


def matchBlocksWie(I, In, d, d2, K):
    """source/matchBlocksWie.m:24-89.  V1, V2: (d, d, K, M, N)."""
    m, n = I.shape
    M, N = m // d, n // d
    Ipad = _search_pad(I, d2)
    Ipad2 = _search_pad(In, d2)
    g = np.zeros((m, n))
    g[:d, :d] = 1
    blockNorms = _corr_fft(g, I ** 2, (m, n))
    gp = np.zeros((d, d)) + 1
    blockNormsPad = _corr_fft(gp, Ipad ** 2, (m + d2, n + d2))
    V1 = np.zeros((d, d, K, M, N))
    V2 = np.zeros_like(V1)
    S = [[None] * N for _ in range(M)]
    for i in range(1, M + 1):
        indI = (i - 1) * d
        for j in range(1, N + 1):
            indJ = (j - 1) * d
            xR = I[(i - 1) * d:i * d, (j - 1) * d:j * d]
            Itmp = Ipad[indI:indI + d + d2, indJ:indJ + d + d2]
            Itmp2 = Ipad2[indI:indI + d + d2, indJ:indJ + d + d2]
            bd = _corr_fft(xR, Itmp, (d + d2, d + d2))
            nb = blockNorms[(i - 1) * d, (j - 1) * d]
            bd = (nb + blockNormsPad[indI:indI + d2, indJ:indJ + d2] - 2 * bd[:d2, :d2]) / nb
            s = _mink_idx(bd.flatten(order="F"), K)
            s[0] = d2 * d2 // 2 + d2 // 2 + 1
            S[i - 1][j - 1] = s
            for k in range(K):
                a, b = _ind2sub(d2, s[k])
                V1[:, :, k, i - 1, j - 1] = Itmp[a - 1:a - 1 + d, b - 1:b - 1 + d]
                V2[:, :, k, i - 1, j - 1] = Itmp2[a - 1:a - 1 + d, b - 1:b - 1 + d]
    return S, V1, V2


def getBM3DAggWeights(U):
    """utilities/getBM3DAggWeights.m (kx = 3, kz = 1)."""
    W = np.sum(np.abs(np.diff(U, 3, axis=1))) * 2.0 ** (1 - 3)
    W = W + np.sum(np.abs(np.diff(U, 3, axis=0))) * 2.0 ** (1 - 3)
    W = W + np.sum(np.abs(np.diff(U, 1, axis=2))) * 2.0 ** (1 - 1)
    return W


def _linspace(a, b, n):
    y = a + np.arange(n) * ((b - a) / (n - 1))
    y[-1] = b
    return y


def aggregate_ht(Vavg, S, d, d2):
    """source/denoiseAndAggregateGPU.m:58-92."""
    M, N = len(S), len(S[0])
    m, n = M * d, N * d
    U = np.zeros((m + d2, n + d2))
    W = np.zeros((m + d2, n + d2))
    s = np.sin(np.pi * _linspace(0.1, 0.9, d))
    mask = np.outer(s, s)
    for y in range(1, M + 1):
        for x in range(1, N + 1):
            Nm = len(S[y - 1][x - 1])
            blocks = Vavg[(y - 1) * d:y * d, (x - 1) * d:x * d, :Nm]
            weight = 1 / (getBM3DAggWeights(blocks) + np.finfo(float).eps)
            for i in range(Nm):
                a, b = _ind2sub(d2, S[y - 1][x - 1][i])
                a = a + (y - 1) * d
                b = b + (x - 1) * d
                U[a - 1:a - 1 + d, b - 1:b - 1 + d] += weight * blocks[:, :, i] * mask
                W[a - 1:a - 1 + d, b - 1:b - 1 + d] += weight * mask
    h = d2 // 2
    return U[h:m + h, h:n + h], W[h:m + h, h:n + h]


def myWavDec1D(Psi, Phi, levels, U):
    """waveTransforms3D/myWavDec1D.m (dim = 3) on a (p,q,r,...) array."""
    C = np.zeros(U.shape)
    cnt = 0
    for _ in range(levels):
        r = U.shape[2]
        FU = np.fft.fft(U, r, axis=2)
        dd = np.fft.ifft(FU * np.fft.fft(Psi, r).reshape(1, 1, r, *([1] * (U.ndim - 3))), r, axis=2)
        C[:, :, cnt:cnt + r // 2] = np.real(dd[:, :, ::2])
        cnt += r // 2
        U = np.real(np.fft.ifft(FU * np.fft.fft(Phi, r).reshape(1, 1, r, *([1] * (U.ndim - 3))), r, axis=2))
        U = U[:, :, ::2]
    C[:, :, cnt:] = U
    return C


def myWavRec1D(Psi, Phi, levels, C):
    """waveTransforms3D/myWavRec1D.m (dim = 3)."""
    r = C.shape[2]
    Cup = [None] * (levels + 1)
    cnt = 0
    for i in range(levels):
        R = r // 2 ** i
        Cup[i] = np.zeros(C.shape[:2] + (R,) + C.shape[3:])
        Cup[i][:, :, ::2] = C[:, :, cnt:cnt + r // 2 ** (i + 1)]
        cnt += r // 2 ** (i + 1)
        FPsi = np.conj(np.fft.fft(Psi, R)).reshape(1, 1, R, *([1] * (C.ndim - 3)))
        Cup[i] = np.real(np.fft.ifft(np.fft.fft(Cup[i], R, axis=2) * FPsi, R, axis=2))
    Cup[levels] = C[:, :, cnt:]
    for i in range(levels, 0, -1):
        R = Cup[i - 1].shape[2]
        tmp = np.zeros(Cup[i - 1].shape)
        tmp[:, :, ::2] = Cup[i]
        FPhi = np.conj(np.fft.fft(Phi, R)).reshape(1, 1, R, *([1] * (C.ndim - 3)))
        Cup[i - 1] = Cup[i - 1] + np.real(np.fft.ifft(np.fft.fft(tmp, R, axis=2) * FPhi, R, axis=2))
    return Cup[0]


def wiener_filter(V1, V2, sigma):
    """source/denoiseAndAggregateWieFast.m:19-66 -> (V2 filtered, filt)."""
    d, K = V1.shape[0], V1.shape[2]
    levels = 3 if K >= 16 else 2
    Psi = np.array([-CC, CC])
    Phi = np.abs(Psi)
    T = scipy.fft.dct(np.eye(d), axis=0, norm="ortho")  # dct(eye(d))
    Tinv = np.linalg.inv(T)
    V1 = np.einsum("pu,uv...->pv...", T, V1)
    V1 = np.einsum("pv...,qv->pq...", V1, T)
    V1 = myWavDec1D(Psi, Phi, levels, V1)
    filt = np.abs(V1) ** 2 / (np.abs(V1) ** 2 + sigma ** 2)
    filt[0, 0, 14:] = 1
    V2 = np.einsum("pu,uv...->pv...", T, V2)
    V2 = np.einsum("pv...,qv->pq...", V2, T)
    V2 = myWavDec1D(Psi, Phi, levels, V2)
    V2 = V2 * filt
    V2 = myWavRec1D(Psi, Phi, levels, V2)
    V2 = np.einsum("pv...,qv->pq...", V2, Tinv)
    V2 = np.einsum("pu,uv...->pv...", Tinv, V2)
    return np.maximum(V2, 0), filt


def wiener_filter_dct(V1, V2, sigma):
    """source/denoiseAndAggregateWie.m:19-55 (the dct() twin of WieFast):
    dct along dims 1 and 2 (orthonormal type II), Haar along dim 3, the same
    gains, idct along dims 2 and 1, clip to [0, 1] -> (V2 filtered, filt)."""
    d, K = V1.shape[0], V1.shape[2]
    levels = 3 if K >= 16 else 2
    Psi = np.array([-CC, CC])
    Phi = np.abs(Psi)
    V1 = scipy.fft.dct(V1, type=2, axis=0, norm="ortho")
    V1 = scipy.fft.dct(V1, type=2, axis=1, norm="ortho")
    V1 = myWavDec1D(Psi, Phi, levels, V1)
    filt = np.abs(V1) ** 2 / (np.abs(V1) ** 2 + sigma ** 2)
    filt[0, 0, 14:] = 1
    V2 = scipy.fft.dct(V2, type=2, axis=0, norm="ortho")
    V2 = scipy.fft.dct(V2, type=2, axis=1, norm="ortho")
    V2 = myWavDec1D(Psi, Phi, levels, V2)
    V2 = V2 * filt
    V2 = myWavRec1D(Psi, Phi, levels, V2)
    V2 = scipy.fft.idct(V2, type=2, axis=1, norm="ortho")
    V2 = scipy.fft.idct(V2, type=2, axis=0, norm="ortho")
    return np.minimum(np.maximum(V2, 0), 1), filt


def aggregate_wie(V2, filt, S, d, d2):
    """source/denoiseAndAggregateWieFast.m:68-100."""
    M, N = len(S), len(S[0])
    m, n = M * d, N * d
    U = np.zeros((m + d2, n + d2))
    W = np.zeros((m + d2, n + d2))
    s = np.sin(np.pi * _linspace(0.2, 0.8, d))
    w = np.outer(s, s)
    for y in range(1, M + 1):
        for x in range(1, N + 1):
            blocks = V2[:, :, :, y - 1, x - 1]
            weight = w / np.linalg.norm(filt[:, :, :, y - 1, x - 1].flatten(order="F")) ** 2
            for i in range(len(S[y - 1][x - 1])):
                a, b = _ind2sub(d2, S[y - 1][x - 1][i])
                a = a + (y - 1) * d
                b = b + (x - 1) * d
                U[a - 1:a - 1 + d, b - 1:b - 1 + d] += weight * blocks[:, :, i]
                W[a - 1:a - 1 + d, b - 1:b - 1 + d] += weight
    h = d2 // 2
    return U[h:m + h, h:n + h], W[h:m + h, h:n + h]


def myMirrorPadImage(I, M, N):
    """utilities/myMirrorPadImage.m:3-8."""
    m, n = I.shape
    I2 = np.zeros((M, N))
    I2[:m, :n] = I
    I2[m:M, :n] = np.flipud(I[m - (M - m):m, :n])
    I2[:m, n:N] = np.fliplr(I[:m, n - (N - n):n])
    I2[m:M, n:N] = np.fliplr(np.flipud(I[m - (M - m):m, n - (N - n):n]))
    return I2


def GBM3D(I, sigma, opts):
    """source/GBM3D.m:21-83 on a resolved opts dict."""
    m, n = I.shape
    bs = opts["blockSize"]
    M = int(np.ceil(m / bs) * bs)
    N = int(np.ceil(n / bs) * bs)
    I = myMirrorPadImage(np.asarray(I, dtype=np.float64), M, N)
    ns = len(opts["matchSpins"])
    U = np.zeros((M, N, ns))
    W = U.copy()
    U1 = U.copy()
    tau = sigma * getTau(opts["levels"], opts["tauMode"])
    K = opts["numMax"]

    def setslice(A, ii, val):
        if ii >= A.shape[2]:
            A = np.concatenate([A, np.zeros((M, N, ii + 1 - A.shape[2]))], axis=2)
        A[:, :, ii] = val
        return A

    for ii, sp in enumerate(opts["matchSpins"]):
        S, Vol, _ = matchBlocksCPU(np.roll(I, (sp, sp), (0, 1)), bs, opts["matchSize"], K)
        Vavg = ht_filter(Vol, tau, opts["levels"], opts["cycleSpin"])
        u, w = aggregate_ht(Vavg, S, bs, opts["matchSize"])
        U1 = setslice(U1, ii, np.roll(u, (-sp, -sp), (0, 1)))
        W = setslice(W, ii, np.roll(w, (-sp, -sp), (0, 1)))
    U1 = np.sum(U1, 2) / np.sum(W, 2)
    if not opts["Wiener"]:
        return U1[:m, :n], U1[:m, :n]
    d = opts["blockSizeWie"]
    for ii, sp in enumerate(opts["matchSpinsWie"]):
        S, V1, V2 = matchBlocksWie(np.roll(U1, (sp, sp), (0, 1)), np.roll(I, (sp, sp), (0, 1)), d,
                                   opts["matchSize"], K)
        V2f, filt = wiener_filter(V1, V2, sigma)
        u, w = aggregate_wie(V2f, filt, S, d, opts["matchSize"])
        U = setslice(U, ii, np.roll(u, (-sp, -sp), (0, 1)))
        W = setslice(W, ii, np.roll(w, (-sp, -sp), (0, 1)))
    Uo = np.sum(U, 2) / np.sum(W, 2)
    return Uo[:m, :n], U1[:m, :n]


def S_to_array(S, K):
    """list-of-lists S -> (K, M, N) uint32 array."""
    M, N = len(S), len(S[0])
    out = np.zeros((K, M, N), dtype=np.uint32)
    for i in range(M):
        for j in range(N):
            out[:, i, j] = S[i][j]
    return out


def S_from_array(Sa):
    K, M, N = Sa.shape
    return [[Sa[:, i, j] for j in range(N)] for i in range(M)]


def getBM3D_distribution(m, n, matchSpins, k):
    """utilities/getBM3D_distribution.m:10-48 -> R (cases x 5, 1-based), d0, k."""
    d0 = 256 + k
    Xcnt = int(np.ceil((n - k) / (d0 - k)))
    Ycnt = int(np.ceil((m - k) / (d0 - k)))
    R = np.zeros((Xcnt * Ycnt * len(matchSpins), 5), dtype=np.int64)
    cnt = 0
    for i in range(1, Ycnt + 1):
        indY = (i - 1) * (d0 - k)
        for j in range(1, Xcnt + 1):
            indX = (j - 1) * (d0 - k)
            for mS in range(len(matchSpins)):
                R[cnt, 0] = indY + 1
                R[cnt, 1] = indY + d0
                R[cnt, 2] = indX + 1
                R[cnt, 3] = indX + d0
                R[cnt, 4] = matchSpins[mS]
                if R[cnt, 3] >= n:
                    R[cnt, 2] = min(n - int(np.ceil((n - R[cnt, 2] + 1) / k)) * k + 1, n - 2 * k + 1)
                if R[cnt, 1] >= m:
                    R[cnt, 0] = min(m - int(np.ceil((m - R[cnt, 0] + 1) / k)) * k + 1, m - 2 * k + 1)
                cnt += 1
    R[:, 1] = np.minimum(R[:, 1], m)
    R[:, 3] = np.minimum(R[:, 3], n)
    return R, d0, k


def makeInterpWindow(a, b, k):
    """utilities/makeInterpWindow.m:8-13."""
    w = np.ones((a, b))
    ramp = _linspace(0.01, 1, k)
    w[:k, :] = w[:k, :] * np.tile(ramp[:, None], (1, b))
    w[:, :k] = w[:, :k] * np.tile(ramp[None, :], (a, 1))
    w[a - k:, :] = w[a - k:, :] * np.flipud(np.tile(ramp[:, None], (1, b)))
    w[:, b - k:] = w[:, b - k:] * np.fliplr(np.tile(ramp[None, :], (a, 1)))
    return w


def GBM3D_distributed(I, sigma, opts):
    """source/GBM3D_distributed.m:21-120 on a resolved opts dict (white noise):
    per case (tile, match spin) the stages on circshift(I(tile), [s, s]), then
    ramp-window stitching U2 = sum(U1 .* w) ./ sum(w .* W).  Returns (U, U2)."""
    I = np.asarray(I, dtype=np.float64)
    m, n = I.shape
    k = opts["blockSize"]
    if m < 256 + k or n < 256 + k:  # :24-27
        return GBM3D(I, sigma, opts)
    tau = sigma * getTau(opts["levels"], opts["tauMode"])
    R, d0, K = getBM3D_distribution(m, n, opts["matchSpins"], k)
    RW, d0, K = getBM3D_distribution(m, n, opts["matchSpinsWie"], k)
    K_ = opts["numMax"]
    d2 = opts["matchSize"]

    def tile(A, r):
        return A[r[0] - 1:r[1], r[2] - 1:r[3]]

    U1, W = [None] * len(R), [None] * len(R)
    for ii in range(len(R)):  # :54-62
        sp = int(R[ii, 4])
        Ilocal = np.roll(tile(I, R[ii]), (sp, sp), (0, 1))
        S, Volume, _ = matchBlocksCPU(Ilocal, k, d2, K_)
        Vavg = ht_filter(Volume, tau, opts["levels"], opts["cycleSpin"])
        u, w = aggregate_ht(Vavg, S, k, d2)
        U1[ii] = np.roll(u, (-sp, -sp), (0, 1))
        W[ii] = np.roll(w, (-sp, -sp), (0, 1))
    U2 = np.zeros((m, n))
    norm = np.zeros((m, n))
    mWindow = makeInterpWindow(d0, d0, K)
    for ii in range(len(R)):  # :69-80
        a, b = R[ii, 1] - R[ii, 0] + 1, R[ii, 3] - R[ii, 2] + 1
        tw = makeInterpWindow(a, b, K) if (a != d0 or b != d0) else mWindow
        U2[R[ii, 0] - 1:R[ii, 1], R[ii, 2] - 1:R[ii, 3]] += U1[ii] * tw
        norm[R[ii, 0] - 1:R[ii, 1], R[ii, 2] - 1:R[ii, 3]] += tw * W[ii]
    U2 = U2 / norm
    if not opts["Wiener"]:
        return U2, U2
    d = opts["blockSizeWie"]
    U1, W = [None] * len(RW), [None] * len(RW)
    for ii in range(len(RW)):  # :94-104
        sp = int(RW[ii, 4])
        Ilocal1 = np.roll(tile(I, RW[ii]), (sp, sp), (0, 1))
        Ilocal2 = np.roll(tile(U2, RW[ii]), (sp, sp), (0, 1))
        S, V1, V2 = matchBlocksWie(Ilocal2, Ilocal1, d, d2, K_)
        V2f, filt = wiener_filter(V1, V2, sigma)
        u, w = aggregate_wie(V2f, filt, S, d, d2)
        U1[ii] = np.roll(u, (-sp, -sp), (0, 1))
        W[ii] = np.roll(w, (-sp, -sp), (0, 1))
    Uf = np.zeros((m, n))
    norm = np.zeros((m, n))
    for ii in range(len(RW)):  # :107-120
        a, b = RW[ii, 1] - RW[ii, 0] + 1, RW[ii, 3] - RW[ii, 2] + 1
        tw = makeInterpWindow(a, b, K) if (a != d0 or b != d0) else mWindow
        Uf[RW[ii, 0] - 1:RW[ii, 1], RW[ii, 2] - 1:RW[ii, 3]] += U1[ii] * tw
        norm[RW[ii, 0] - 1:RW[ii, 1], RW[ii, 2] - 1:RW[ii, 3]] += tw * W[ii]
    return Uf / norm, U2


# ------------------------------------------------ whole-image matching (2-D)
def getWaveFilters2D(Lo_d, Hi_d):
    """waveTransforms2D/getWaveFilters2D.m:12-16 -> Psi (N, N, 4); the taps
    are the data file's Lo_d, Hi_d (row vectors, so Lo_d'*Hi_d is an outer
    product with the first factor along dimension 1)."""
    Lo, Hi = np.asarray(Lo_d, float), np.asarray(Hi_d, float)
    Psi = np.zeros((Lo.size, Lo.size, 4))
    Psi[:, :, 0] = np.outer(Lo, Lo)
    Psi[:, :, 1] = np.outer(Hi, Lo)
    Psi[:, :, 2] = np.outer(Lo, Hi)
    Psi[:, :, 3] = np.outer(Hi, Hi)
    return Psi


def myWavDec2DFFT(Psi, levels, U):
    """waveTransforms2D/myWavDec2DFFT.m:10-23; C[i][j] 0-based (C{i,j})."""
    C = [[None] * 3 for _ in range(levels + 1)]
    for i in range(levels):
        FU = np.fft.fftn(U)
        for j in range(3):
            c = np.fft.ifftn(FU * np.fft.fftn(Psi[:, :, j + 1], s=U.shape, axes=(0, 1)))
            C[i][j] = np.real(c[::2, ::2])
        U = np.real(np.fft.ifftn(FU * np.fft.fftn(Psi[:, :, 0], s=U.shape, axes=(0, 1))))
        U = U[::2, ::2]
    C[levels][0] = U
    return C


def myWavRec2DGPUfast(Psi, levels, C):
    """waveTransforms2D/myWavRec2DGPUfast.m:10-34."""
    Cup = [None] * (levels + 1)
    for i in range(levels):
        Cup[i] = np.zeros((2 * C[i][0].shape[0], 2 * C[i][0].shape[1]))
        tmp = Cup[i].copy()
        for j in range(3):
            tmp[::2, ::2] = C[i][j]
            Cup[i] = Cup[i] + np.real(np.fft.ifftn(np.fft.fftn(tmp) *
                                                   np.conj(np.fft.fftn(Psi[:, :, j + 1], s=tmp.shape, axes=(0, 1)))))
    Cup[levels] = C[levels][0]
    for i in range(levels, 0, -1):          # i = levels+1:-1:2
        tmp = np.zeros(Cup[i - 1].shape)
        tmp[::2, ::2] = Cup[i]
        FPsi = np.conj(np.fft.fftn(Psi[:, :, 0], s=tmp.shape, axes=(0, 1)))
        Cup[i - 1] = Cup[i - 1] + np.real(np.fft.ifftn(np.fft.fftn(tmp) * FPsi))
    return Cup[0]


def matchBlocksGPU_prefilter(I, sigma, Lo_d, Hi_d):
    """source/matchBlocksGPU.m:34-50 (the sigma > 30/255 branch)."""
    tau = 2 * sigma
    Psi = getWaveFilters2D(Lo_d, Hi_d)
    Iavg = np.zeros(I.shape)
    shiftCnt, levels = 5, 3
    for k in range(1, shiftCnt + 1):
        C = myWavDec2DFFT(Psi, levels, np.roll(I, (k - 1, k - 1), (0, 1)))
        for i in range(levels):
            for j in range(3):
                C[i][j] = C[i][j] * (np.abs(C[i][j]) > tau)
        Iavg = Iavg + np.roll(myWavRec2DGPUfast(Psi, levels, C), (1 - k, 1 - k), (0, 1)) / shiftCnt
    return Iavg


def matchBlocksGPU(I, sigma, opts, Lo_d, Hi_d):
    """source/matchBlocksGPU.m:18-91 -> (S as a list of lists of 1-based
    uint32 index arrays, the distance maps, the image searched)."""
    tol = .005 * sigma ** 2
    d = opts["blockSize"]
    m, n = I.shape
    tol = tol * d ** 2
    M, N = m // d, n // d
    I = np.asarray(I, float)
    if sigma > 30 / 255:
        I = matchBlocksGPU_prefilter(I, sigma, Lo_d, Hi_d)
    g = np.zeros((m, n))
    g[:d, :d] = 1
    blockNorms = np.real(np.roll(np.fft.ifft2(np.fft.fft2(g) * np.fft.fft2(I ** 2)), (-d + 1, -d + 1), (0, 1)))
    FI = np.fft.fft2(I)
    S = [[None] * N for _ in range(M)]
    dists = {}
    for i in range(M):
        for j in range(N):
            xR = I[i * d:(i + 1) * d, j * d:(j + 1) * d]
            bd = np.real(np.fft.ifft2(np.conj(np.fft.fft2(xR, s=(m, n))) * FI))
            nb = blockNorms[i * d, j * d]
            bd = (nb + blockNorms - 2 * bd) / nb
            dists[(i, j)] = bd
            flat = bd.flatten(order="F")
            if opts["numMax"] == opts["numMin"]:
                s = _mink_idx(flat, opts["numMax"])
            else:
                s = np.flatnonzero(flat < tol) + 1
                if s.size > opts["numMax"]:
                    s = _mink_idx(flat, opts["numMax"])
                elif s.size < opts["numMin"]:
                    s = _mink_idx(flat, opts["numMin"])
            S[i][j] = s.astype(np.uint32)
    return S, dists, I


def matchBlocksCPU_tol(I, sigma, d, d2, numMax, numMin):
    """source/matchBlocksCPU.m:21-28, 56-111 with numMax ~= numMin (the
    tolerance branch).  Returns (S as a list of lists, Volume); raises
    ValueError where MATLAB's `Volume(.., 1:numel(S{i,j})) = blocks` (:111,
    blocks = zeros(d, d, numMax), :70) fails on a size mismatch."""
    tol = .005 * sigma ** 2
    tol = tol * d ** 2
    m, n = I.shape
    M, N = m // d, n // d
    Ipad = _search_pad(I, d2)
    g = np.ones((d, d))
    blockNorms = _corr_fft(g, I ** 2, (m, n))
    blockNormsPad = _corr_fft(g, Ipad ** 2, (m + d2, n + d2))
    Volume = np.zeros((m, n, numMax))
    blocks = np.zeros((d, d, numMax))
    S = [[None] * N for _ in range(M)]
    for i in range(1, M + 1):
        indI = (i - 1) * d
        for j in range(1, N + 1):
            indJ = (j - 1) * d
            xR = I[(i - 1) * d:i * d, (j - 1) * d:j * d]
            Itmp = Ipad[indI:indI + d + d2, indJ:indJ + d + d2]
            bd = _corr_fft(xR, Itmp, (d + d2, d + d2))
            nb = blockNorms[(i - 1) * d, (j - 1) * d]
            bd = (nb + blockNormsPad[indI:indI + d2, indJ:indJ + d2] - 2 * bd[:d2, :d2]) / nb
            flat = bd.flatten(order="F")
            s = np.flatnonzero(flat < tol) + 1                      # find(blockDistances<tol)
            if s.size > numMax:
                s = _mink_idx(flat, numMax)
            elif s.size < numMin:
                s = _mink_idx(flat, numMin)
            s = s.astype(np.uint32)
            s[0] = d2 * d2 // 2 + d2 // 2 + 1                        # S{i,j}(1) = centre (:103)
            for k in range(s.size):
                a, b = _ind2sub(d2, s[k])
                if k >= blocks.shape[2]:                             # MATLAB grows blocks
                    blocks = np.concatenate([blocks, np.zeros((d, d, k + 1 - blocks.shape[2]))], axis=2)
                blocks[:, :, k] = Itmp[a - 1:a - 1 + d, b - 1:b - 1 + d]
            if blocks.shape[2] != s.size:
                raise ValueError(f"Unable to perform assignment: left side {d}x{d}x{s.size}, "
                                 f"right side {d}x{d}x{blocks.shape[2]}")
            Volume[(i - 1) * d:i * d, (j - 1) * d:j * d, :s.size] = blocks
            S[i - 1][j - 1] = s
    return S, Volume


# ------------------------------------------------------------ coloured noise
def makeGausPSF(d1, d2, omega):
    """Stand-in for makeGausPSF([d1 d2], omega), which the reference calls
    (utilities/getColoredWie.m:6) but does not contain: the circular Gaussian
    of standard deviation omega pixels centred at the origin, unit sum, and
    its DFT (used as a blur by multiplication with fft2)."""
    dx = np.minimum(np.arange(d1), d1 - np.arange(d1))[:, None].astype(float)
    dy = np.minimum(np.arange(d2), d2 - np.arange(d2))[None, :].astype(float)
    g = np.exp(-(dx ** 2 + dy ** 2) / (2.0 * omega ** 2))
    g /= g.sum()
    return g, np.fft.fft2(g)


def _imresize_weights(n_in, n_out):
    """MATLAB imresize's contributions for the 'bilinear' (triangle) kernel,
    antialiased when shrinking: (indices 0-based, weights) per output sample."""
    scale = n_out / n_in
    kw = 2.0
    tri = lambda x: (x + 1) * ((-1 <= x) & (x < 0)) + (1 - x) * ((0 <= x) & (x <= 1))
    if scale < 1:
        h = lambda x: scale * tri(scale * x)
        kw = kw / scale
    else:
        h = tri
    x = np.arange(1, n_out + 1, dtype=float)[:, None]
    u = x / scale + 0.5 * (1 - 1 / scale)
    left = np.floor(u - kw / 2)
    P = int(np.ceil(kw)) + 2
    idx = left + np.arange(P)[None, :]
    w = h(u - idx)
    w = w / w.sum(axis=1, keepdims=True)
    aux = np.concatenate([np.arange(1, n_in + 1), np.arange(n_in, 0, -1)])
    idx = aux[np.mod(idx.astype(int) - 1, len(aux))] - 1
    return idx, w


def imresize_bilinear(A, out_shape):
    """imresize(A, out_shape, 'bilinear') for a 2-D double array: the
    dimension with the smaller scale factor first (rows on a tie)."""
    A = np.asarray(A, dtype=float)
    s = (out_shape[0] / A.shape[0], out_shape[1] / A.shape[1])
    order = (0, 1) if s[0] <= s[1] else (1, 0)
    for dim in order:
        idx, w = _imresize_weights(A.shape[dim], out_shape[dim])
        if dim == 0:
            A = np.einsum("ip,ipj->ij", w, A[idx, :])
        else:
            A = np.einsum("jp,ijp->ij", w, A[:, idx])
    return A


def getColoredWie(sigmaPSD, d):
    """utilities/getColoredWie.m:6-19 as written (with the makeGausPSF stand-in)."""
    sigmaPSD = np.asarray(sigmaPSD, dtype=float)
    _, ghat = makeGausPSF(sigmaPSD.shape[0], sigmaPSD.shape[1], 3)
    sigmaPSD = np.real(np.fft.ifft2(ghat * np.fft.fft2(sigmaPSD)))
    sigmaPSD = imresize_bilinear(sigmaPSD, (d, d))
    sigmaWie = np.zeros((d, d))
    for i in range(d):
        for j in range(d):
            x = np.zeros((d, d))
            x[j, i] = 1
            y = scipy.fft.idctn(x, norm="ortho")
            y = np.abs(np.fft.fft2(y)) ** 2
            sigmaWie[j, i] = np.sqrt(np.sum(sigmaPSD * y)) / d
    return sigmaWie
