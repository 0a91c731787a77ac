"""ctypes wrapper of the oracle library (oracle/liboracle_gbm3d.so).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker / CPU baseline, never by the
product path (gbm3d_amd/ does not import this package).

Arrays are numpy arrays with MATLAB semantics: an m x n image is a numpy array
of shape (m, n) and is passed column-major; volumes are (m, n, K); S is
(K, Mt, Nt) uint32 with 1-based window indices.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from gbm3d_amd._abi import Opts, opts_to_struct, struct_to_opts

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle_gbm3d.so")
_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = C.CDLL(LIB)
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f(a, dt):
    return np.asfortranarray(a, dtype=dt)


def _check(st):
    if st != 0:
        raise RuntimeError(f"oracle status {st}")


def _dt(precision):
    return (np.float64, "f64") if precision == 64 else (np.float32, "f32")


def set_threads(n: int) -> int:
    return lib().orc_set_threads(C.c_int(n))


def resolve_opts(opts: dict | None) -> dict:
    o = Opts()
    src = opts_to_struct(opts)
    lib().orc_resolve_opts(C.byref(src) if src is not None else None, C.byref(o))
    return struct_to_opts(o)


def set_opts(profile: str) -> dict:
    o = Opts()
    lib().orc_set_opts(C.c_char_p(profile.encode()), C.byref(o))
    return struct_to_opts(o)


def check_opts(opts: dict) -> dict:
    o = opts_to_struct(opts)
    lib().orc_check_opts(C.byref(o))
    return struct_to_opts(o)


def get_tau(levels: int, mode: int, sigma: float) -> np.ndarray:
    t = np.zeros((levels, 7), dtype=np.float64, order="F")
    _check(lib().orc_get_tau(C.c_int(levels), C.c_int(mode), C.c_double(sigma), _p(t)))
    return t


def mirror_pad(I, M, N):
    I = _f(I, np.float64)
    out = np.zeros((M, N), dtype=np.float64, order="F")
    _check(lib().orc_mirror_pad_f64(_p(I), I.shape[0], I.shape[1], M, N, _p(out)))
    return out


def search_pad(X, d2):
    X = _f(X, np.float64)
    m, n = X.shape
    out = np.zeros((m + d2, n + d2), dtype=np.float64, order="F")
    _check(lib().orc_search_pad_f64(_p(X), m, n, d2, _p(out)))
    return out


def match(I, d, d2, K, precision=64, want_dist=False):
    """matchBlocksCPU -> (S (K,Mt,Nt) uint32, Volume (m,n,K), dist (d2*d2,Mt,Nt) or None)."""
    dt, suf = _dt(precision)
    I = _f(I, dt)
    m, n = I.shape
    Mt, Nt = m // d, n // d
    S = np.zeros((K, Mt, Nt), dtype=np.uint32, order="F")
    V = np.zeros((m, n, K), dtype=dt, order="F")
    dist = np.zeros((d2 * d2, Mt, Nt), dtype=dt, order="F") if want_dist else None
    fn = getattr(lib(), "orc_match_" + suf)
    _check(fn(_p(I), m, n, d, d2, K, _p(S), _p(V), _p(dist)))
    return S, V, dist


def match_tol(I, d, d2, numMax, numMin, sigma, precision=64):
    """matchBlocksCPU's tolerance branch (numMax != numMin, tol = 0.005
    sigma^2 d^2) -> (S, Volume); raises where the reference fails (a tile
    with fewer than numMax distances below tol)."""
    dt, suf = _dt(precision)
    I = _f(I, dt)
    m, n = I.shape
    S = np.zeros((numMax, m // d, n // d), dtype=np.uint32, order="F")
    V = np.zeros((m, n, numMax), dtype=dt, order="F")
    _check(getattr(lib(), "orc_match_tol_" + suf)(_p(I), m, n, d, d2, numMax, numMin, C.c_double(sigma), _p(S),
                                                   _p(V)))
    return S, V


def match_wie(I, In, d, d2, K, precision=64, want_dist=False):
    """matchBlocksWie -> (S, V1, V2, dist); V: (d, d, K, Mt, Nt)."""
    dt, suf = _dt(precision)
    I = _f(I, dt)
    In = _f(In, dt)
    m, n = I.shape
    Mt, Nt = m // d, n // d
    S = np.zeros((K, Mt, Nt), dtype=np.uint32, order="F")
    V1 = np.zeros((d, d, K, Mt, Nt), dtype=dt, order="F")
    V2 = np.zeros_like(V1)
    dist = np.zeros((d2 * d2, Mt, Nt), dtype=dt, order="F") if want_dist else None
    fn = getattr(lib(), "orc_match_wie_" + suf)
    _check(fn(_p(I), _p(In), m, n, d, d2, K, _p(S), _p(V1), _p(V2), _p(dist)))
    return S, V1, V2, dist


def sym4_taps():
    """The sym4 (Lo_d, Hi_d) the oracle's prefilter is built from."""
    lo, hi = np.zeros(8), np.zeros(8)
    lib().orc_sym4_taps(_p(lo), _p(hi))
    return lo, hi


def wav2_prefilter(I, tau, precision=64):
    """matchBlocksGPU.m:34-50 with threshold tau: the mean over 5 cycle shifts
    of the 3-level 2-D sym4 analysis, hard threshold, synthesis."""
    dt, suf = _dt(precision)
    I = _f(I, dt)
    out = np.zeros_like(I)
    _check(getattr(lib(), "orc_wav2_prefilter_" + suf)(_p(I), I.shape[0], I.shape[1], C.c_double(tau), _p(out)))
    return out


def match_global(I, sigma, d, numMax, numMin=None, precision=64, want_filtered=False):
    """S = matchBlocksGPU(I, sigma, opts) (source/matchBlocksGPU.m:18-91).

    Returns (S, counts[, filtered]): S (numMax, floor(m/d), floor(n/d)) uint32,
    1-based column-major indices into m x n, each tile's first counts[ti, tj]
    entries its list (the rest 0); filtered is the image the search ran on
    (the sym4 prefilter when sigma > 30/255)."""
    dt, suf = _dt(precision)
    I = _f(I, dt)
    m, n = I.shape
    numMin = numMax if numMin is None else numMin
    S = np.zeros((numMax, m // d, n // d), dtype=np.uint32, order="F")
    cnt = np.zeros((m // d, n // d), dtype=np.uint32, order="F")
    F = np.zeros_like(I)
    _check(getattr(lib(), "orc_match_global_" + suf)(_p(I), m, n, C.c_double(sigma), d, numMax, numMin, _p(S),
                                                      _p(cnt), _p(F)))
    return (S, cnt, F) if want_filtered else (S, cnt)


def ht_filter(V, levels, cycle_spin, tau, precision=64, want_margin=False):
    dt, suf = _dt(precision)
    V = _f(V, dt)
    m, n, K = V.shape
    out = np.zeros_like(V)
    tau = np.asfortranarray(tau, dtype=np.float64)
    mm = C.c_double(0.0)
    fn = getattr(lib(), "orc_ht_filter_" + suf)
    _check(fn(_p(V), m, n, K, levels, cycle_spin, _p(tau), _p(out), C.byref(mm) if want_margin else None))
    return (out, mm.value) if want_margin else out


def ht_near_mask(V, levels, cycle_spin, tau, tol=1e-3):
    """Voxels of ht_filter's output in the synthesis support of a detail
    coefficient within tol (absolute) of its threshold (FP32 restatement):
    (mask (m, n, K) bool, number of such coefficients)."""
    V = _f(V, np.float32)
    m, n, K = V.shape
    mask = np.zeros((m, n, K), dtype=np.uint8, order="F")
    tau = np.asfortranarray(tau, dtype=np.float64)
    f = lib().orc_ht_near_mask_f32
    f.restype = C.c_long
    cnt = f(_p(V), m, n, K, levels, cycle_spin, _p(tau), C.c_double(tol), _p(mask))
    if cnt < 0:
        raise RuntimeError("oracle ht_near_mask failed")
    return mask.astype(bool), int(cnt)


def u1_near_mask(I, sigma, opts=None, tol=1e-3):
    """Pixels of GBM3D's first estimate U1 that a step-1 coefficient within
    tol (absolute) of its threshold can move (FP32 restatement; every pixel covered
    by a group whose TV weight such a coefficient changes):
    (mask (m, n) bool, number of such coefficients)."""
    I = _f(I, np.float32)
    m, n = I.shape
    o = Opts()
    src = opts_to_struct(opts)
    lib().orc_resolve_opts(C.byref(src) if src is not None else None, C.byref(o))
    mask = np.zeros((m, n), dtype=np.uint8, order="F")
    f = lib().orc_u1_near_mask_f32
    f.restype = C.c_long
    cnt = f(_p(I), m, n, C.c_float(sigma), C.byref(o), C.c_double(tol), _p(mask))
    if cnt < 0:
        raise RuntimeError("oracle u1_near_mask failed")
    return mask.astype(bool), int(cnt)


def median3(V, precision=64):
    """medfilt3(Volume, [3 3 3]) of the step-1 'median' branch."""
    dt, suf = _dt(precision)
    V = _f(V, dt)
    m, n, K = V.shape
    out = np.zeros_like(V)
    _check(getattr(lib(), "orc_median3_" + suf)(_p(V), m, n, K, _p(out)))
    return out


def agg_weights(Vavg, d, precision=64):
    dt, suf = _dt(precision)
    Vavg = _f(Vavg, dt)
    m, n, K = Vavg.shape
    w = np.zeros((m // d, n // d), dtype=dt, order="F")
    _check(getattr(lib(), "orc_agg_weights_" + suf)(_p(Vavg), m, n, d, K, _p(w)))
    return w


def aggregate_ht(Vavg, S, d, d2, precision=64, gather=True):
    dt, suf = _dt(precision)
    Vavg = _f(Vavg, dt)
    S = np.asfortranarray(S, dtype=np.uint32)
    m, n, K = Vavg.shape
    U = np.zeros((m, n), dtype=dt, order="F")
    W = np.zeros_like(U)
    _check(getattr(lib(), "orc_agg_ht_" + suf)(_p(Vavg), _p(S), m, n, d, d2, K, _p(U), _p(W), int(gather)))
    return U, W


def wiener(V1, V2, sigma, precision=64, sigma_wie=None, clip01=False):
    """sigma_wie (d x d): per-coefficient sigma (coloured noise) instead of sigma.
    clip01: denoiseAndAggregateWie.m's clip to [0, 1] (source/denoiseAndAggregateWie.m:55)."""
    dt, suf = _dt(precision)
    V1 = _f(V1, dt)
    V2 = _f(V2, dt)
    d, K = V1.shape[0], V1.shape[2]
    G = V1.size // (d * d * K)
    V2f = np.zeros_like(V2)
    g2 = np.zeros(V1.shape[3:], dtype=dt, order="F")
    if clip01:
        _check(getattr(lib(), "orc_wiener01_" + suf)(_p(V1), _p(V2), d, K, C.c_long(G), C.c_double(sigma), _p(V2f),
                                                     _p(g2)))
    elif sigma_wie is None:
        _check(getattr(lib(), "orc_wiener_" + suf)(_p(V1), _p(V2), d, K, C.c_long(G), C.c_double(sigma), _p(V2f),
                                                   _p(g2)))
    else:
        sw = _f(sigma_wie, np.float64)
        _check(getattr(lib(), "orc_wiener_tab_" + suf)(_p(V1), _p(V2), d, K, C.c_long(G), _p(sw), _p(V2f), _p(g2)))
    return V2f, g2


def aggregate_wie(V2f, gain2, S, m, n, d2, precision=64, gather=True):
    dt, suf = _dt(precision)
    V2f = _f(V2f, dt)
    gain2 = _f(gain2, dt)
    S = np.asfortranarray(S, dtype=np.uint32)
    d, K = V2f.shape[0], V2f.shape[2]
    U = np.zeros((m, n), dtype=dt, order="F")
    W = np.zeros_like(U)
    _check(getattr(lib(), "orc_agg_wie_" + suf)(_p(V2f), _p(gain2), _p(S), m, n, d, d2, K, _p(U), _p(W), int(gather)))
    return U, W


def wav_dec_level(X):
    X = _f(X, np.float64)
    p, q, r = X.shape
    C8 = np.zeros((p // 2, q // 2, r // 2, 8), dtype=np.float64, order="F")
    _check(lib().orc_wav_dec_level_f64(_p(X), p, q, r, _p(C8)))
    return C8


def wav_rec_level(C8):
    C8 = _f(C8, np.float64)
    p, q, r = 2 * C8.shape[0], 2 * C8.shape[1], 2 * C8.shape[2]
    X = np.zeros((p, q, r), dtype=np.float64, order="F")
    _check(lib().orc_wav_rec_level_f64(_p(C8), p, q, r, _p(X)))
    return X


def haar_dec(x, levels):
    x = np.array(x, dtype=np.float64)
    lib().orc_haar_dec_f64(_p(x), x.size, levels)
    return x


def haar_rec(x, levels):
    x = np.array(x, dtype=np.float64)
    lib().orc_haar_rec_f64(_p(x), x.size, levels)
    return x


def gbm3d(I, sigma, opts=None, precision=64, nthreads=0):
    """[U, out.U1] = GBM3D(I, sigma, opts) on the CPU restatement."""
    dt, suf = _dt(precision)
    I = _f(I, dt)
    m, n = I.shape
    o = Opts()
    src = opts_to_struct(opts)
    lib().orc_resolve_opts(C.byref(src) if src is not None else None, C.byref(o))
    U = np.zeros((m, n), dtype=dt, order="F")
    U1 = np.zeros_like(U)
    if precision == 64:
        st = lib().orc_gbm3d_f64(_p(I), m, n, C.c_double(sigma), C.byref(o), _p(U), _p(U1))
    else:
        st = lib().orc_gbm3d_f32(_p(I), m, n, C.c_float(sigma), C.byref(o), _p(U), _p(U1), nthreads)
    _check(st)
    return U, U1


def step2_from_pilot(pilot, noisy, sigma, opts=None, precision=32, sigma_wie=None):
    """Step 2 of GBM3D (source/GBM3D.m:70-83) run on a given pilot U1.

    pilot, noisy: M x N (already block-multiple sizes).  Composes the oracle's
    stage functions: circshift, matchBlocksWie, Wiener filter, aggregation,
    shift back, sum(U,3)./sum(W,3).  Profiles whose step-2 spin list is
    shorter than the step-1 list (the 'fast' slice quirk) are not covered.
    """
    ro = resolve_opts(opts)
    spins1, spins2 = ro["matchSpins"], ro["matchSpinsWie"]
    if len(spins2) < len(spins1):
        raise ValueError("step2_from_pilot: numel(matchSpinsWie) < numel(matchSpins) is not covered")
    dt = np.float64 if precision == 64 else np.float32
    pilot = np.asarray(pilot, dtype=dt)
    noisy = np.asarray(noisy, dtype=dt)
    M, N = pilot.shape
    d, d2, K = ro["blockSizeWie"], ro["matchSize"], ro["numMax"]
    Us = np.zeros((M, N))
    Ws = np.zeros((M, N))
    for sp in spins2:
        P = np.roll(pilot, (sp, sp), (0, 1))
        I = np.roll(noisy, (sp, sp), (0, 1))
        S, V1, V2, _ = match_wie(P, I, d, d2, K, precision=precision)
        V2f, g2 = wiener(V1, V2, sigma, precision=precision, sigma_wie=sigma_wie)
        U, W = aggregate_wie(V2f, g2, S, M, N, d2, precision=precision)
        Us += np.roll(U.astype(np.float64), (-sp, -sp), (0, 1))
        Ws += np.roll(W.astype(np.float64), (-sp, -sp), (0, 1))
    return Us / Ws


def distribution(m, n, spins, k):
    """getBM3D_distribution -> (cases, 5) int rows (y1, y2, x1, x2, spin), 1-based inclusive."""
    sp = np.ascontiguousarray(np.atleast_1d(spins), dtype=np.int32)
    cap = 4096
    rows = np.zeros((cap, 5), dtype=np.int32)
    cnt = lib().orc_distribution(m, n, _p(sp), sp.size, k, _p(rows), cap)
    if cnt < 0:
        raise RuntimeError("too many cases")
    return rows[:cnt].copy()


def interp_window(a, b, k):
    w = np.zeros((a, b), dtype=np.float64, order="F")
    _check(lib().orc_interp_window(a, b, k, _p(w)))
    return w


def gbm3d_distributed(I, sigma, opts=None, precision=64, pilot=None):
    """[U, out.U2] = GBM3D_distributed(I, sigma, opts); pilot: skip step 1."""
    dt, suf = _dt(precision)
    I = _f(I, dt)
    m, n = I.shape
    o = Opts()
    src = opts_to_struct(opts)
    lib().orc_resolve_opts(C.byref(src) if src is not None else None, C.byref(o))
    U = np.zeros((m, n), dtype=dt, order="F")
    U2 = np.zeros_like(U)
    P = None if pilot is None else _f(pilot, dt)
    fn = getattr(lib(), "orc_gbm3d_distributed_" + suf)
    sig = C.c_double(sigma) if precision == 64 else C.c_float(sigma)
    _check(fn(_p(I), m, n, sig, C.byref(o), _p(U), _p(U2), _p(P)))
    return U, U2


def colored_tau(levels, psd):
    psd = _f(psd, np.float64)
    d1, d2 = psd.shape
    tau = np.zeros((levels, 7), dtype=np.float64, order="F")
    _check(lib().orc_colored_tau(levels, _p(psd), d1, d2, _p(tau)))
    return tau


def gbm3d_colored(I, psd, sigma_wie, opts=None, precision=64):
    """[U, out.U1] = GBM3D(I, sigmaPSD, opts) with the caller's getColoredWie output."""
    dt, suf = _dt(precision)
    I = _f(I, dt)
    m, n = I.shape
    psd = _f(psd, np.float64)
    sw = None if sigma_wie is None else _f(sigma_wie, np.float64)
    o = Opts()
    src = opts_to_struct(opts)
    lib().orc_resolve_opts(C.byref(src) if src is not None else None, C.byref(o))
    U = np.zeros((m, n), dtype=dt, order="F")
    U1 = np.zeros_like(U)
    fn = getattr(lib(), "orc_gbm3d_colored_" + suf)
    _check(fn(_p(I), m, n, _p(psd), psd.shape[0], psd.shape[1], _p(sw), C.byref(o), _p(U), _p(U1)))
    return U, U1
