"""ctypes mirror of include/gbm3d.h and the loader of libgbm3d.so.

The library is built in-tree (gbm3d_amd/libgbm3d.so) by gbm3d_amd.build;
loading fails loudly when it is missing -- there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

GBM3D_MAX_SPINS = 16
GBM3D_UNSET = -2147483648

GBM3D_BATCH_GRAPH = 1
GBM3D_BATCH_TUNE = 2

GBM3D_OK = 0
GBM3D_ERR_ARG = 1
GBM3D_ERR_UNSUPPORTED = 2
GBM3D_ERR_HIP = 3
GBM3D_ERR_NOMEM = 4
GBM3D_ERR_MISSING_FIELD = 5

FILT = {"ht": 0, "median": 1}
FILT_NAME = {v: k for k, v in FILT.items()}

STAGES = ("pad", "match1", "ht_lo", "ht_syn", "ht_syn0", "agg1_prep", "agg1", "match2", "wiener", "agg2", "norm",
          "shrink")

GBM3D_SHRINK_PARAMS = 49600

LIB_NAME = "libgbm3d.so"
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), LIB_NAME)


class Opts(C.Structure):
    """struct gbm3d_opts (include/gbm3d.h)."""

    _fields_ = [
        ("profile", C.c_char * 16),
        ("numMax", C.c_int32),
        ("numMin", C.c_int32),
        ("levels", C.c_int32),
        ("cycleSpin", C.c_int32),
        ("nMatchSpins", C.c_int32),
        ("matchSpins", C.c_int32 * GBM3D_MAX_SPINS),
        ("nMatchSpinsWie", C.c_int32),
        ("matchSpinsWie", C.c_int32 * GBM3D_MAX_SPINS),
        ("tauMode", C.c_int32),
        ("blockSize", C.c_int32),
        ("blockSizeWie", C.c_int32),
        ("matchSize", C.c_int32),
        ("Wiener", C.c_int32),
        ("filtType", C.c_int32),
    ]


class Ext(C.Structure):
    """struct gbm3d_ext (include/gbm3d.h)."""

    _fields_ = [("match_bf16", C.c_int32), ("rescore", C.c_int32), ("learned", C.c_int32),
                ("learned_bound", C.c_int32), ("match_stats", C.c_int32),
                ("reserved", C.c_int32 * 3)]


_INT_FIELDS = ("numMax", "numMin", "levels", "cycleSpin", "tauMode", "blockSize",
               "blockSizeWie", "matchSize")


def opts_to_struct(opts: dict | None) -> Opts | None:
    """MATLAB-struct-like dict -> gbm3d_opts; absent keys become GBM3D_UNSET.

    Keys the path never reads (wname, wnamez, order, orderz) are ignored, as in
    the reference (source/GBM3D.m:38-40).
    """
    if opts is None:
        return None
    o = Opts()
    prof = opts.get("profile")
    o.profile = (prof or "").encode()[:15]
    for f in _INT_FIELDS:
        setattr(o, f, int(opts[f]) if f in opts else GBM3D_UNSET)
    o.Wiener = int(bool(opts["Wiener"])) if "Wiener" in opts else GBM3D_UNSET
    if "filtType" in opts:
        ft = opts["filtType"]
        if isinstance(ft, str):
            if ft not in FILT:
                raise ValueError("unrecognized filtering")  # denoiseAndAggregateGPU.m:54
            ft = FILT[ft]
        o.filtType = int(ft)
    else:
        o.filtType = GBM3D_UNSET
    for name, cnt in (("matchSpins", "nMatchSpins"), ("matchSpinsWie", "nMatchSpinsWie")):
        if name in opts:
            vals = [int(v) for v in np.atleast_1d(opts[name])]
            if len(vals) > GBM3D_MAX_SPINS:
                raise ValueError(f"{name}: at most {GBM3D_MAX_SPINS} spins")
            setattr(o, cnt, len(vals))
            arr = getattr(o, name)
            for i, v in enumerate(vals):
                arr[i] = v
        else:
            setattr(o, cnt, GBM3D_UNSET)
    return o


def struct_to_opts(o: Opts) -> dict:
    """gbm3d_opts -> dict with only the present fields."""
    out: dict = {}
    if o.profile:
        out["profile"] = o.profile.decode()
    for f in _INT_FIELDS:
        v = getattr(o, f)
        if v != GBM3D_UNSET:
            out[f] = int(v)
    if o.Wiener != GBM3D_UNSET:
        out["Wiener"] = bool(o.Wiener)
    if o.filtType != GBM3D_UNSET:
        out["filtType"] = FILT_NAME.get(int(o.filtType), int(o.filtType))
    for name, cnt in (("matchSpins", "nMatchSpins"), ("matchSpinsWie", "nMatchSpinsWie")):
        n = getattr(o, cnt)
        if n != GBM3D_UNSET:
            out[name] = [int(getattr(o, name)[i]) for i in range(n)]
    return out


# Every exported symbol of libgbm3d.so: name -> (restype, argtypes).
_P = C.c_void_p
_I = C.c_int
_F = C.c_float
_D = C.c_double
SIGNATURES = {
    "gbm3d_opts_clear": (_I, [C.POINTER(Opts)]),
    "gbm3d_set_opts": (_I, [C.c_char_p, C.POINTER(Opts)]),
    "gbm3d_check_opts": (_I, [C.POINTER(Opts)]),
    "gbm3d_resolve_opts": (_I, [C.POINTER(Opts), C.POINTER(Opts)]),
    "gbm3d_get_tau": (_I, [_I, _I, _D, _P]),
    "gbm3d_get_colored_tau": (_I, [_I, _P, _I, _I, _P]),
    "gbm3d_get_colored_wie": (_I, [_P, _I, _I, _I, _P]),
    "gbm3d_denoise_device_noalloc": (_I, [_P, _P, _I, _I, C.c_float, C.POINTER(Opts), _P, _P]),
    "gbm3d_bf16_match_stats": (_I, [_P, C.POINTER(C.c_ulonglong), _I]),
    "gbm3d_denoise_aggregate_wie": (_I, [_P, _P, _P, _P, _I, _I, _I, _I, _I, C.c_float, _I, _P, _P]),
    "gbm3d_validate": (_I, [C.POINTER(Opts), _I, _I, C.POINTER(C.c_size_t)]),
    "gbm3d_denoise_colored": (_I, [_P, _P, _I, _I, _P, _I, _I, _P, _I, _I, C.POINTER(Opts), _P, _P]),
    "gbm3d_last_error": (C.c_char_p, []),
    "gbm3d_create": (_I, [_I, C.POINTER(_P)]),
    "gbm3d_set_ext": (_I, [_P, C.POINTER(Ext)]),
    "gbm3d_get_ext": (_I, [_P, C.POINTER(Ext)]),
    "gbm3d_destroy": (_I, [_P]),
    "gbm3d_set_stream": (_I, [_P, _P]),
    "gbm3d_synchronize": (_I, [_P]),
    "gbm3d_denoise": (_I, [_P, _P, _I, _I, _F, C.POINTER(Opts), _P, _P]),
    "gbm3d_denoise_device": (_I, [_P, _P, _I, _I, _F, C.POINTER(Opts), _P, _P]),
    "gbm3d_denoise_distributed": (_I, [_P, _P, _I, _I, _F, C.POINTER(Opts), _P, _P]),
    "gbm3d_denoise_distributed_device": (_I, [_P, _P, _I, _I, _F, C.POINTER(Opts), _P, _P]),
    "gbm3d_denoise_batch_device": (_I, [_P, _I, _P, _P, _P, _P, C.POINTER(Opts), _P, _P, _I]),
    "gbm3d_batch_info": (_I, [_P, C.POINTER(_I), C.POINTER(_I), C.POINTER(_D)]),
    "gbm3d_match": (_I, [_P, _P, _I, _I, _I, _I, _I, _P, _P]),
    "gbm3d_match_global": (_I, [_P, _P, _I, _I, _D, _I, _I, _I, _P, _P, _P]),
    "gbm3d_match_tol": (_I, [_P, _P, _I, _I, _I, _I, _I, _I, _D, _P, _P]),
    "gbm3d_ht_filter": (_I, [_P, _P, _I, _I, _I, _I, _I, _P, _P]),
    "gbm3d_median_filter": (_I, [_P, _P, _I, _I, _I, _P]),
    "gbm3d_aggregate_ht": (_I, [_P, _P, _P, _I, _I, _I, _I, _I, _P, _P]),
    "gbm3d_match_wie": (_I, [_P, _P, _P, _I, _I, _I, _I, _I, _P, _P, _P]),
    "gbm3d_wiener": (_I, [_P, _P, _P, _I, _I, _I, _F, _P, _P]),
    "gbm3d_aggregate_wie": (_I, [_P, _P, _P, _P, _I, _I, _I, _I, _I, _P, _P]),
    "gbm3d_set_timing": (_I, [_P, _I]),
    "gbm3d_device_copy": (_I, [_P, _P, _P, C.c_size_t]),
    "gbm3d_get_stage_times": (_I, [_P, _P, _P]),
    "gbm3d_shrink_set_weights": (_I, [_P, _P, _I]),
    "gbm3d_shrink_stats": (_I, [_P, _I, C.POINTER(_D), C.POINTER(_D), C.POINTER(_I)]),
    "gbm3d_shrink_groups": (_I, [_P, _P, _I, _F, _P, _P]),
    "gbm3d_shrink_features": (_I, [_P, _P, _P, _I, _I, _F, _P, _I, _I, _P, _P, _P]),
    "gbm3d_shrink_forward": (_I, [_P, _P, _P, _P, _I, _P]),
    "gbm3d_shrink_grad": (_I, [_P, _P, _P, _P, _P, _P, _I, _P, C.POINTER(_D)]),
    "gbm3d_shrink_adam": (_I, [_P, _P, _P, _P, _P, _I, _F, _F, _F, _F]),
    "gbm3d_synth_pair": (_I, [_P, C.c_uint64, _I, _I, _F, _P, _P]),
}

_lib = None


class GBM3DError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"gbm3d status {status}: {msg}")
        self.status = status


def load_library(path: str | None = None):
    """Load libgbm3d.so once; raise if it has not been built."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise ImportError(
            f"{p} is missing: build the HIP library first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def check(status: int) -> None:
    if status != GBM3D_OK:
        msg = load_library().gbm3d_last_error() or b""
        raise GBM3DError(status, msg.decode(errors="replace"))


def ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(C.c_void_p)
