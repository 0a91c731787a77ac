"""Python mirror of the reference's GBM3D call surface, backed by libgbm3d.so.

    U, out = GBM3D(I, sigma, opts)          # source/GBM3D.m:1
    opts = setBM3Dopts('default')           # utilities/setBM3Dopts.m:1
    opts = checkBM3Dopts(opts)              # utilities/checkBM3Dopts.m:1

Images are numpy arrays of shape (m, n) holding values on any scale (the
path is scale-equivariant; sigma must be on the image's scale).  opts is a
dict playing the MATLAB struct: absent keys are absent fields.  Every call runs
the HIP kernels on the GPU; there is no CPU path.

Stage functions mirror the reference stages for stage-isolated use:
matchBlocksCPU, denoiseAndAggregateGPU (ht filter + aggregation),
matchBlocksWie, denoiseAndAggregateWieFast.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi
from ._abi import (GBM3D_ERR_ARG, GBM3D_ERR_UNSUPPORTED, GBM3DError, Opts, check, load_library, opts_to_struct, ptr,
                   struct_to_opts)


def setBM3Dopts(profile: str = "default") -> dict:
    """utilities/setBM3Dopts.m: profile table; unknown names mean 'default'."""
    lib = load_library()
    o = Opts()
    check(lib.gbm3d_set_opts(C.c_char_p(str(profile).encode()), C.byref(o)))
    return struct_to_opts(o)


def checkBM3Dopts(opts: dict) -> dict:
    """utilities/checkBM3Dopts.m: fill absent fields with defaults."""
    lib = load_library()
    o = opts_to_struct(opts)
    check(lib.gbm3d_check_opts(C.byref(o)))
    out = dict(opts)
    out.update(struct_to_opts(o))
    return out


def resolve_opts(opts: dict | None) -> dict:
    """The option rule at the top of GBM3D (source/GBM3D.m:15-19)."""
    lib = load_library()
    src = opts_to_struct(opts)
    o = Opts()
    check(lib.gbm3d_resolve_opts(C.byref(src) if src is not None else None, C.byref(o)))
    return struct_to_opts(o)


def validate(opts: dict | None, m: int, n: int) -> int:
    """Check opts and an m x n image against GBM3D's rules and the built
    kernels (host only); returns the device workspace bytes of one image with
    its match spins run in series.  Raises GBM3DError as gbm3d_denoise would."""
    lib = load_library()
    src = opts_to_struct(opts)
    nb = C.c_size_t(0)
    check(lib.gbm3d_validate(C.byref(src) if src is not None else None, m, n, C.byref(nb)))
    return int(nb.value)


def check_sigma_wie(opts: dict | None, sigma_wie) -> None:
    """sigmaWie (getColoredWie's output) scales the blockSizeWie x blockSizeWie
    DCT coefficients of every step-2 slice (denoiseAndAggregateWieFast.m:49):
    with the Wiener step on it must be exactly that size."""
    o = resolve_opts(opts)
    if not o.get("Wiener") or sigma_wie is None:  # None: the library computes getColoredWie
        return
    bw = o["blockSizeWie"]
    if np.shape(sigma_wie) != (bw, bw):
        raise GBM3DError(GBM3D_ERR_ARG, f"coloured noise: sigmaWie must be blockSizeWie x blockSizeWie "
                                        f"({bw} x {bw}), got {' x '.join(map(str, np.shape(sigma_wie)))}")


def getTau(levels: int, mode: int, sigma: float = 1.0) -> np.ndarray:
    """sigma*getTau(levels, mode) as a (levels, 7) array."""
    lib = load_library()
    t = np.zeros((levels, 7), dtype=np.float64, order="F")
    check(lib.gbm3d_get_tau(levels, mode, C.c_double(sigma), ptr(t)))
    return t


def getColoredTau(levels: int, sigmaPSD) -> np.ndarray:
    """utilities/getColoredTau.m: thresholds (levels, 7) from a d1 x d2 sigma^2 x PSD."""
    lib = load_library()
    psd = np.asfortranarray(sigmaPSD, dtype=np.float64)
    t = np.zeros((levels, 7), dtype=np.float64, order="F")
    check(lib.gbm3d_get_colored_tau(levels, ptr(psd), psd.shape[0], psd.shape[1], ptr(t)))
    return t


def getColoredWie(sigmaPSD, d: int) -> np.ndarray:
    """utilities/getColoredWie.m: the d x d per-DCT-coefficient Wiener sigma of
    a d1 x d2 sigma^2 x PSD (makeGausPSF, absent from the reference, taken as
    the unit-sum circular Gaussian of standard deviation 3 px)."""
    lib = load_library()
    psd = np.asfortranarray(sigmaPSD, dtype=np.float64)
    w = np.zeros((d, d), dtype=np.float64, order="F")
    check(lib.gbm3d_get_colored_wie(ptr(psd), psd.shape[0], psd.shape[1], d, ptr(w)))
    return w


def myPSNR(Iref, Irec, maxI=None) -> float:
    """utilities/myPSNR.m: 10*log10(maxI^2/MSE), maxI defaults to max(Iref)."""
    Iref = np.asarray(Iref, dtype=np.float64)
    Irec = np.asarray(Irec, dtype=np.float64)
    if maxI is None:
        maxI = Iref.max()
    mse = np.sum((Iref - Irec) ** 2) / Iref.size
    return float(10 * np.log10(maxI ** 2 / mse))


class Context:
    """One gbm3d_ctx: a device, a HIP stream and cached device workspace."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        h = C.c_void_p()
        check(self.lib.gbm3d_create(device, C.byref(h)))
        self.handle = h
        self.device = device

    def close(self):
        if self.handle:
            self.lib.gbm3d_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, stream_ptr: int | None):
        check(self.lib.gbm3d_set_stream(self.handle, C.c_void_p(stream_ptr) if stream_ptr else None))

    def synchronize(self):
        check(self.lib.gbm3d_synchronize(self.handle))

    def set_variant(self, match_bf16: bool = False, rescore: int = 0, learned: bool = False,
                    learned_bound: float = 1.0, weights=None, match_stats: bool = False):
        """Forms beside the reference path (all False = reference).

        match_bf16: BF16-MFMA distances with error bounds + exact FP32
        rescoring of every candidate the bounds cannot exclude (first set of
        `rescore` candidates per tile, numMax..64; 0: 2*numMax); index lists
        equal the FP32 path's bit for bit.
        learned: learned shrinkage replaces the step-1 hard threshold.  Per
        match spin a device-side judge compares the energy the learned gains
        removed with learned_bound x the noise energy; above it the spin's
        reference filter runs instead (gated launches, no host round trip), so
        an image whose spins all fall back equals the reference path bit for
        bit.  Loads `weights` (parameter vector) or the shipped trained
        weights."""
        if learned:
            from . import shrink

            self.load_shrink_weights(shrink.load_weights() if weights is None else weights)
        e = _abi.Ext()
        e.match_bf16 = int(bool(match_bf16))
        e.rescore = int(rescore)
        e.learned = int(bool(learned))
        e.learned_bound = int(round(learned_bound * 1000)) if learned else 0
        e.match_stats = int(bool(match_stats))
        check(self.lib.gbm3d_set_ext(self.handle, C.byref(e)))

    # ------------------------------------------------------ learned shrinkage
    def bf16_match_stats(self, reset: bool = True) -> dict:
        """Path counters of the BF16 matcher (tiles settled on the first
        rescored set, on a widened set, by the exact search; candidates rescored)."""
        out = (C.c_ulonglong * 8)()
        check(self.lib.gbm3d_bf16_match_stats(self.handle, out, int(reset)))
        return {"first_set": out[0], "widened": out[1], "exact_all": out[2], "rescored": out[3],
                "exact_all_why": {"first_set_over_128": out[4], "no_kth_key": out[5], "widened_over_128": out[6],
                                  "zero_norm_reference": out[7]}}

    def load_shrink_weights(self, params):
        p = np.ascontiguousarray(params, dtype=np.float32).ravel()
        check(self.lib.gbm3d_shrink_set_weights(self.handle, ptr(p), p.size))

    def shrink_stats(self, image: int = 0):
        """(energy removed by the learned gains, by a 2.7 sigma hard threshold,
        both / noise energy and averaged over the match spins; the number of
        spins that fell back) of image `image` of the last learned-form call
        (waits for the context stream)."""
        a, b, f = C.c_double(), C.c_double(), C.c_int()
        check(self.lib.gbm3d_shrink_stats(self.handle, image, C.byref(a), C.byref(b), C.byref(f)))
        return a.value, b.value, int(f.value)

    def shrink_groups(self, V, sigma: float, want_gains: bool = False):
        """Fused learned shrinkage of groups V (8, 8, 16, ...) -> (Vout, gains (G, 64, 16) or None)."""
        V = np.asfortranarray(V, dtype=np.float32)
        G = V.size // (8 * 8 * 16)
        out = np.zeros_like(V)
        g = np.zeros((8, 8, 16, G), dtype=np.float32, order="F") if want_gains else None
        check(self.lib.gbm3d_shrink_groups(self.handle, ptr(V), G, C.c_float(sigma), ptr(out), ptr(g)))
        gains = None if g is None else g.reshape(64, 16, G, order="F").transpose(2, 0, 1)
        return out, gains

    def synth_pair(self, seed: int, m: int, n: int, sigma: float, dClean: int | None, dNoisy: int | None):
        check(self.lib.gbm3d_synth_pair(self.handle, C.c_uint64(seed), m, n, C.c_float(sigma),
                                        C.c_void_p(dClean) if dClean else None,
                                        C.c_void_p(dNoisy) if dNoisy else None))

    def shrink_features(self, dClean: int, dNoisy: int, m: int, n: int, sigma: float, dSel: int, nsel: int,
                        dX: int, dT: int, dWt: int, target: str = "wiener"):
        t = {"wiener": 0, "energy": 1}[target]
        check(self.lib.gbm3d_shrink_features(self.handle, C.c_void_p(dClean), C.c_void_p(dNoisy), m, n,
                                             C.c_float(sigma), C.c_void_p(dSel), nsel, t, C.c_void_p(dX),
                                             C.c_void_p(dT), C.c_void_p(dWt)))

    def shrink_forward(self, dParams: int, dX: int, dSig: int, ngroups: int, dGains: int):
        check(self.lib.gbm3d_shrink_forward(self.handle, C.c_void_p(dParams), C.c_void_p(dX), C.c_void_p(dSig),
                                            ngroups, C.c_void_p(dGains)))

    def shrink_grad(self, dParams: int, dX: int, dT: int, dWt: int, dSig: int, ngroups: int, dGrad: int,
                    want_loss: bool = True):
        loss = C.c_double()
        check(self.lib.gbm3d_shrink_grad(self.handle, C.c_void_p(dParams), C.c_void_p(dX), C.c_void_p(dT),
                                         C.c_void_p(dWt), C.c_void_p(dSig), ngroups, C.c_void_p(dGrad),
                                         C.byref(loss) if want_loss else None))
        return loss.value if want_loss else None

    def shrink_adam(self, dParams: int, dM: int, dV: int, dGrad: int, step: int, lr: float, beta1: float = 0.9,
                    beta2: float = 0.999, eps: float = 1e-8):
        check(self.lib.gbm3d_shrink_adam(self.handle, C.c_void_p(dParams), C.c_void_p(dM), C.c_void_p(dV),
                                         C.c_void_p(dGrad), step, C.c_float(lr), C.c_float(beta1),
                                         C.c_float(beta2), C.c_float(eps)))

    def set_timing(self, on: bool):
        check(self.lib.gbm3d_set_timing(self.handle, int(bool(on))))

    def stage_times(self) -> dict:
        ms = np.zeros(len(_abi.STAGES), dtype=np.float64)
        cnt = np.zeros(len(_abi.STAGES), dtype=np.int32)
        check(self.lib.gbm3d_get_stage_times(self.handle, ptr(ms), ptr(cnt)))
        return {s: (float(ms[i]), int(cnt[i])) for i, s in enumerate(_abi.STAGES)}

    def device_copy(self, dst: int, src: int, nbytes: int):
        check(self.lib.gbm3d_device_copy(self.handle, C.c_void_p(dst), C.c_void_p(src), C.c_size_t(nbytes)))

    # ------------------------------------------------------------ full path
    def denoise(self, I, sigma: float, opts: dict | None = None):
        I = np.asfortranarray(I, dtype=np.float32)
        if I.ndim != 2:
            raise ValueError("GBM3D expects a 2-D grayscale image")
        m, n = I.shape
        U = np.zeros((m, n), dtype=np.float32, order="F")
        U1 = np.zeros_like(U)
        o = opts_to_struct(opts)
        check(self.lib.gbm3d_denoise(self.handle, ptr(I), m, n, C.c_float(sigma),
                                     C.byref(o) if o is not None else None, ptr(U), ptr(U1)))
        return U, {"U1": U1}

    def denoise_colored(self, I, sigma_psd, sigma_wie=None, opts: dict | None = None):
        """[U, out] = GBM3D(I, sigmaPSD, opts) for coloured noise; sigma_wie is
        getColoredWie(sigmaPSD, blockSizeWie) (blockSizeWie x blockSizeWie),
        computed by the library when None (source/GBM3D.m:35)."""
        I = np.asfortranarray(I, dtype=np.float32)
        m, n = I.shape
        psd = np.asfortranarray(sigma_psd, dtype=np.float64)
        if psd.ndim != 2:
            raise ValueError("sigmaPSD must be a 2-D array")
        sw = None if sigma_wie is None else np.asfortranarray(sigma_wie, dtype=np.float64)
        check_sigma_wie(opts, sw)
        wm, wn = (0, 0) if sw is None else sw.shape
        U = np.zeros((m, n), dtype=np.float32, order="F")
        U1 = np.zeros_like(U)
        o = opts_to_struct(opts)
        # the library checks wm x wn against the resolved blockSizeWie
        check(self.lib.gbm3d_denoise_colored(self.handle, ptr(I), m, n, ptr(psd), psd.shape[0], psd.shape[1],
                                             ptr(sw), wm, wn, C.byref(o) if o is not None else None, ptr(U),
                                             ptr(U1)))
        return U, {"U1": U1}

    def denoise_distributed(self, I, sigma: float, opts: dict | None = None):
        """[U, out] = GBM3D_distributed(I, sigma, opts): out["U2"] is the first estimate."""
        I = np.asfortranarray(I, dtype=np.float32)
        if I.ndim != 2:
            raise ValueError("GBM3D_distributed expects a 2-D grayscale image")
        m, n = I.shape
        U = np.zeros((m, n), dtype=np.float32, order="F")
        U2 = np.zeros_like(U)
        o = opts_to_struct(opts)
        check(self.lib.gbm3d_denoise_distributed(self.handle, ptr(I), m, n, C.c_float(sigma),
                                                 C.byref(o) if o is not None else None, ptr(U), ptr(U2)))
        return U, {"U2": U2}

    def denoise_distributed_device(self, dI: int, m: int, n: int, sigma: float, opts, dU: int,
                                   dU2: int | None = None):
        o = opts_to_struct(opts)
        check(self.lib.gbm3d_denoise_distributed_device(self.handle, C.c_void_p(dI), m, n, C.c_float(sigma),
                                                        C.byref(o) if o is not None else None, C.c_void_p(dU),
                                                        C.c_void_p(dU2) if dU2 else None))

    def denoise_device(self, dI: int, m: int, n: int, sigma: float, opts, dU: int, dU1: int | None = None):
        """Device pointers (ints, column-major float32); asynchronous."""
        o = opts_to_struct(opts)
        check(self.lib.gbm3d_denoise_device(self.handle, C.c_void_p(dI), m, n, C.c_float(sigma),
                                            C.byref(o) if o is not None else None, C.c_void_p(dU),
                                            C.c_void_p(dU1) if dU1 else None))

    def denoise_device_noalloc(self, dI: int, m: int, n: int, sigma: float, opts, dU: int, dU1: int | None = None):
        """gbm3d_denoise_device without any device allocation (static scratch array)."""
        o = opts_to_struct(opts)
        check(self.lib.gbm3d_denoise_device_noalloc(self.handle, C.c_void_p(dI), m, n, C.c_float(sigma),
                                                    C.byref(o) if o is not None else None, C.c_void_p(dU),
                                                    C.c_void_p(dU1) if dU1 else None))

    def denoise_batch_device(self, dIs, ms, ns, sigmas, opts, dUs, dU1s=None, graph=True, tune=False):
        """Batched GBM3D over device buffers (lists of int pointers); asynchronous."""
        cnt = len(dIs)
        P = C.c_void_p * cnt
        I32 = C.c_int * cnt
        F32 = C.c_float * cnt
        o = opts_to_struct(opts)
        flags = (_abi.GBM3D_BATCH_GRAPH if graph else 0) | (_abi.GBM3D_BATCH_TUNE if tune else 0)
        check(self.lib.gbm3d_denoise_batch_device(
            self.handle, cnt, P(*dIs), I32(*ms), I32(*ns), F32(*sigmas), C.byref(o) if o is not None else None,
            P(*dUs), P(*dU1s) if dU1s else None, flags))

    def batch_info(self):
        s, w, t = C.c_int(), C.c_int(), C.c_double()
        check(self.lib.gbm3d_batch_info(self.handle, C.byref(s), C.byref(w), C.byref(t)))
        return {"streams": s.value, "match_waves": w.value, "tuned_ms": t.value}

    def denoise_batch(self, images, sigmas, opts=None, graph=True, tune=False, repeats=1):
        """Host convenience: upload, run the batch `repeats` times, download U."""
        from . import device

        dI = [device.DeviceArray.from_host(np.asarray(I, dtype=np.float32)) for I in images]
        dU = [device.DeviceArray(np.shape(I)) for I in images]
        for _ in range(repeats):
            self.denoise_batch_device([d.ptr for d in dI], [np.shape(I)[0] for I in images],
                                      [np.shape(I)[1] for I in images], list(sigmas), opts, [d.ptr for d in dU],
                                      graph=graph, tune=tune)
        self.synchronize()
        return [d.to_host() for d in dU]

    # ------------------------------------------------------------ stages
    def match(self, I, d: int, d2: int, K: int, want_volume: bool = True):
        I = np.asfortranarray(I, dtype=np.float32)
        m, n = I.shape
        S = np.zeros((K, m // d, n // d), dtype=np.uint32, order="F")
        V = np.zeros((m, n, K), dtype=np.float32, order="F") if want_volume else None
        check(self.lib.gbm3d_match(self.handle, ptr(I), m, n, d, d2, K, ptr(S), ptr(V)))
        return S, V

    def match_tol(self, I, d: int, d2: int, numMax: int, numMin: int, sigma: float, want_volume: bool = True):
        """matchBlocksCPU's tolerance branch (numMax != numMin): (S, Volume);
        raises GBM3DError (ERR_ARG) where the reference fails (a tile with
        fewer than numMax matches below tol = 0.005 sigma^2 d^2)."""
        I = np.asfortranarray(I, dtype=np.float32)
        m, n = I.shape
        S = np.zeros((numMax, m // d, n // d), dtype=np.uint32, order="F")
        V = np.zeros((m, n, numMax), dtype=np.float32, order="F") if want_volume else None
        check(self.lib.gbm3d_match_tol(self.handle, ptr(I), m, n, d, d2, numMax, numMin, float(sigma), ptr(S),
                                       ptr(V)))
        return S, V

    def match_global(self, I, sigma: float, d: int, numMax: int, numMin: int | None = None,
                     want_filtered: bool = False):
        """matchBlocksGPU's whole-image search (source/matchBlocksGPU.m:18-91).

        Returns (S, counts[, filtered]): S (numMax, floor(m/d), floor(n/d))
        uint32 1-based column-major indices into m x n, each tile's first
        counts[ti, tj] entries its list and the rest 0; filtered, the image
        searched (the sym4 prefilter when sigma > 30/255)."""
        I = np.asfortranarray(I, dtype=np.float32)
        m, n = I.shape
        numMin = numMax if numMin is None else numMin
        S = np.zeros((numMax, m // d, n // d), dtype=np.uint32, order="F")
        cnt = np.zeros((m // d, n // d), dtype=np.uint32, order="F")
        F = np.zeros_like(I) if want_filtered else None
        check(self.lib.gbm3d_match_global(self.handle, ptr(I), m, n, float(sigma), d, numMax, numMin, ptr(S),
                                          ptr(cnt), ptr(F)))
        return (S, cnt, F) if want_filtered else (S, cnt)

    def ht_filter(self, V, levels: int, cycle_spin: int, tau):
        V = np.asfortranarray(V, dtype=np.float32)
        m, n, K = V.shape
        tau = np.asfortranarray(tau, dtype=np.float64)
        out = np.zeros_like(V)
        check(self.lib.gbm3d_ht_filter(self.handle, ptr(V), m, n, K, levels, cycle_spin, ptr(tau), ptr(out)))
        return out

    def median_filter(self, V):
        """Vavg = medfilt3(Volume, [3 3 3]) (the filtType 'median' branch)."""
        V = np.asfortranarray(V, dtype=np.float32)
        m, n, K = V.shape
        out = np.zeros_like(V)
        check(self.lib.gbm3d_median_filter(self.handle, ptr(V), m, n, K, ptr(out)))
        return out

    def aggregate_ht(self, Vavg, S, d: int, d2: int):
        Vavg = np.asfortranarray(Vavg, dtype=np.float32)
        S = np.asfortranarray(S, dtype=np.uint32)
        m, n, K = Vavg.shape
        U = np.zeros((m, n), dtype=np.float32, order="F")
        W = np.zeros_like(U)
        check(self.lib.gbm3d_aggregate_ht(self.handle, ptr(Vavg), ptr(S), m, n, d, d2, K, ptr(U), ptr(W)))
        return U, W

    def match_wie(self, I, In, d: int, d2: int, K: int, want_groups: bool = True):
        I = np.asfortranarray(I, dtype=np.float32)
        In = np.asfortranarray(In, dtype=np.float32)
        m, n = I.shape
        S = np.zeros((K, m // d, n // d), dtype=np.uint32, order="F")
        V1 = V2 = None
        if want_groups:
            V1 = np.zeros((d, d, K, m // d, n // d), dtype=np.float32, order="F")
            V2 = np.zeros_like(V1)
        check(self.lib.gbm3d_match_wie(self.handle, ptr(I), ptr(In), m, n, d, d2, K, ptr(S), ptr(V1), ptr(V2)))
        return S, V1, V2

    def wiener(self, V1, V2, sigma: float):
        V1 = np.asfortranarray(V1, dtype=np.float32)
        V2 = np.asfortranarray(V2, dtype=np.float32)
        d, K = V1.shape[0], V1.shape[2]
        G = V1.size // (d * d * K)
        V2f = np.zeros_like(V2)
        g2 = np.zeros(V1.shape[3:], dtype=np.float32, order="F")
        check(self.lib.gbm3d_wiener(self.handle, ptr(V1), ptr(V2), d, K, G, C.c_float(sigma), ptr(V2f), ptr(g2)))
        return V2f, g2

    def denoise_aggregate_wie(self, V1, V2, S, sigma: float, d2: int, clip01: bool = True):
        """[U, W] = denoiseAndAggregateWie(V1, V2, S, sigma, opts) (clip01, the
        dct() twin that clips to [0, 1]) or denoiseAndAggregateWieFast."""
        V1 = np.asfortranarray(V1, dtype=np.float32)
        V2 = np.asfortranarray(V2, dtype=np.float32)
        S = np.asfortranarray(S, dtype=np.uint32)
        d, K, Mt, Nt = V1.shape[0], V1.shape[2], V1.shape[3], V1.shape[4]
        m, n = Mt * d, Nt * d
        U = np.zeros((m, n), dtype=np.float32, order="F")
        W = np.zeros_like(U)
        check(self.lib.gbm3d_denoise_aggregate_wie(self.handle, ptr(V1), ptr(V2), ptr(S), m, n, d, d2, K,
                                                   C.c_float(sigma), int(bool(clip01)), ptr(U), ptr(W)))
        return U, W

    def aggregate_wie(self, V2f, gain2, S, m: int, n: int, d2: int):
        V2f = np.asfortranarray(V2f, dtype=np.float32)
        gain2 = np.asfortranarray(gain2, dtype=np.float32)
        S = np.asfortranarray(S, dtype=np.uint32)
        d, K = V2f.shape[0], V2f.shape[2]
        U = np.zeros((m, n), dtype=np.float32, order="F")
        W = np.zeros_like(U)
        check(self.lib.gbm3d_aggregate_wie(self.handle, ptr(V2f), ptr(gain2), ptr(S), m, n, d, d2, K, ptr(U),
                                           ptr(W)))
        return U, W


_default: Context | None = None


def default_context() -> Context:
    global _default
    if _default is None:
        _default = Context(0)
    return _default


def GBM3D(I, sigma, opts: dict | None = None, sigma_wie=None):
    """[U, out] = GBM3D(I, sigma, opts)  (source/GBM3D.m:1-84); out['U1'] is the pilot.

    sigma may be a matrix sigmaPSD (coloured noise, :29-36); then sigma_wie =
    getColoredWie(sigmaPSD, blockSizeWie), computed here when not given (with
    a stand-in for makeGausPSF, which the reference does not contain)."""
    if np.ndim(sigma) == 2 and np.size(sigma) > 1:
        return default_context().denoise_colored(I, sigma, sigma_wie, opts)
    return default_context().denoise(I, float(np.asarray(sigma).ravel()[0]), opts)


def GBM3D_distributed(I, sigma: float, opts: dict | None = None):
    """[U, out] = GBM3D_distributed(I, sigma, opts) (source/GBM3D_distributed.m)."""
    if opts is None:
        opts = {"profile": "default"}
    return default_context().denoise_distributed(I, sigma, opts)


def matchBlocksCPU(I, sigma: float, opts: dict):
    """[S, Volume] = matchBlocksCPU(I, sigma, opts) on the GPU (the tolerance
    branch when numMax != numMin)."""
    o = checkBM3Dopts(opts)
    if o["numMax"] != o["numMin"]:
        return default_context().match_tol(I, o["blockSize"], o["matchSize"], o["numMax"], o["numMin"], sigma)
    return default_context().match(I, o["blockSize"], o["matchSize"], o["numMax"])


def matchBlocksGPU(I, sigma: float, opts: dict):
    """S = matchBlocksGPU(I, sigma, opts): an object array (floor(m/d),
    floor(n/d)) of uint32 index columns (1-based, into m x n), like the
    reference's cell array; the tolerance branch when numMax != numMin."""
    o = checkBM3Dopts(opts)
    S, cnt = default_context().match_global(I, sigma, o["blockSize"], o["numMax"], o["numMin"])
    out = np.empty(cnt.shape, dtype=object)
    for i in range(cnt.shape[0]):
        for j in range(cnt.shape[1]):
            out[i, j] = S[:cnt[i, j], i, j].copy()
    return out


def denoiseAndAggregateGPU(Volume, S, tau, opts: dict):
    """[U, W] = denoiseAndAggregateGPU(Volume, S, tau, Psi, Psi2, opts) ('ht' or 'median')."""
    o = checkBM3Dopts(opts)
    ctx = default_context()
    if o.get("filtType", "ht") == "median":
        Vavg = ctx.median_filter(Volume)   # denoiseAndAggregateGPU.m:49-52
    elif o.get("filtType", "ht") == "ht":
        Vavg = ctx.ht_filter(Volume, o["levels"], o["cycleSpin"], tau)
    else:
        raise ValueError("unrecognized filtering")   # denoiseAndAggregateGPU.m:54
    return ctx.aggregate_ht(Vavg, S, o["blockSize"], o["matchSize"])


def matchBlocksWie(I, In, sigma: float, opts: dict):
    """[S, V1, V2] = matchBlocksWie(I, In, sigma, opts)."""
    o = checkBM3Dopts(opts)
    return default_context().match_wie(I, In, o["blockSizeWie"], o["matchSize"], o["numMax"])


def denoiseAndAggregateWieFast(V1, V2, S, sigma: float, opts: dict):
    """[U, W] = denoiseAndAggregateWieFast(V1, V2, S, sigma, opts)."""
    o = checkBM3Dopts(opts)
    ctx = default_context()
    V2f, g2 = ctx.wiener(V1, V2, sigma)
    d = o["blockSizeWie"]
    return ctx.aggregate_wie(V2f, g2, S, S.shape[1] * d, S.shape[2] * d, o["matchSize"])


def denoiseAndAggregateWie(V1, V2, S, sigma: float, opts: dict):
    """[U, W] = denoiseAndAggregateWie(V1, V2, S, sigma, opts): the dct() twin
    of denoiseAndAggregateWieFast whose filtered groups are clipped to [0, 1]
    (source/denoiseAndAggregateWie.m:34-55), so it expects images on [0, 1]."""
    o = checkBM3Dopts(opts)
    return default_context().denoise_aggregate_wie(V1, V2, S, sigma, o["matchSize"], clip01=True)
