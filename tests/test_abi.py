"""The C ABI (include/gbm3d.h) of libgbm3d.so: symbols, struct layout and the
host-only option logic, checked against the oracle's restatement.  CPU only:
no call here touches a GPU."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import gbm3d_amd
import oracle as O
from gbm3d_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gbm3d.h")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(gbm3d_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = gbm3d_amd.load_library()
    names = declared_functions()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", _abi.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (gbm3d_\w+)", out))
    assert set(names) <= exported
    assert set(_abi.SIGNATURES) == set(names)


def test_opts_struct_layout_matches_header():
    # profile[16] + 4 ints + n + 16 + n + 16 + 6 ints
    assert C.sizeof(_abi.Opts) == 16 + 4 * (4 + 1 + 16 + 1 + 16 + 6)
    src = "#include \"gbm3d.h\"\n#include <stdio.h>\n#include <stddef.h>\nint main(){printf(\"%zu %zu %zu\\n\"," \
          "sizeof(gbm3d_opts), offsetof(gbm3d_opts, matchSpinsWie), offsetof(gbm3d_opts, filtType));}\n"
    exe = os.path.join(ROOT, "gbm3d_amd", "_build", "abi_probe")
    os.makedirs(os.path.dirname(exe), exist_ok=True)
    r = subprocess.run(["gcc", "-x", "c", "-I", os.path.join(ROOT, "include"), "-", "-o", exe], input=src,
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    size, off_wie, off_ft = map(int, subprocess.run([exe], capture_output=True, text=True).stdout.split())
    assert size == C.sizeof(_abi.Opts)
    assert off_wie == _abi.Opts.matchSpinsWie.offset
    assert off_ft == _abi.Opts.filtType.offset


@pytest.mark.parametrize("profile", ["default", "fast", "superFast", "accuracy", "unknown"])
def test_set_opts_matches_oracle(profile):
    assert gbm3d_amd.setBM3Dopts(profile) == O.set_opts(profile)


@pytest.mark.parametrize("opts", [{}, {"blockSize": 8}, {"blockSize": 32, "blockSizeWie": 16},
                                  {"matchSpins": [0, 2, 4], "Wiener": False}, {"numMax": 8, "numMin": 8}])
def test_check_opts_matches_oracle(opts):
    assert gbm3d_amd.checkBM3Dopts(opts) == O.check_opts(opts)


@pytest.mark.parametrize("opts", [None, {"profile": "fast", "blockSize": 8}, {"blockSize": 8},
                                  {"profile": "superFast"}])
def test_resolve_opts_matches_oracle(opts):
    assert gbm3d_amd.resolve_opts(opts) == O.resolve_opts(opts)


@pytest.mark.parametrize("mode", [1, 2, 3, 4])
def test_get_tau_matches_oracle(mode):
    np.testing.assert_array_equal(gbm3d_amd.getTau(3, mode, 25.0), O.get_tau(3, mode, 25.0))


def test_unrecognized_filtering_raises():
    with pytest.raises(ValueError, match="unrecognized filtering"):
        _abi.opts_to_struct({"filtType": "gauss"})


def test_null_arguments_return_status_not_crash():
    lib = gbm3d_amd.load_library()
    assert lib.gbm3d_set_opts(b"default", None) == _abi.GBM3D_ERR_ARG
    assert b"NULL" in lib.gbm3d_last_error()
    assert lib.gbm3d_get_tau(0, 1, C.c_double(1.0), None) == _abi.GBM3D_ERR_ARG
    assert lib.gbm3d_destroy(None) == 0


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(ImportError):
        _abi.load_library(str(tmp_path / "libgbm3d.so"))


# Option sets GBM3D accepts, and those the built kernels reject before any
# launch (UNSUPPORTED), checked on the host through gbm3d_validate.
@pytest.mark.parametrize("opts", [None, {"profile": "superFast"}, {"profile": "accuracy"},
                                  {"blockSize": 8, "matchSpinsWie": [0, 4]},
                                  {"blockSizeWie": 4, "numMax": 32, "numMin": 32, "matchSpinsWie": [0]},
                                  {"blockSizeWie": 16, "numMax": 8, "numMin": 8, "matchSpinsWie": [0]},
                                  {"blockSize": 32, "matchSize": 48, "matchSpinsWie": [0]},
                                  {"blockSize": 4, "blockSizeWie": 4, "matchSize": 48, "matchSpinsWie": [0]},
                                  {"numMax": 16, "numMin": 8, "matchSpinsWie": [0]}])   # tolerance branch
def test_validate_accepts_built_combinations(opts):
    assert gbm3d_amd.validate(opts, 128, 96) > 0


@pytest.mark.parametrize("opts,what", [
    ({"numMax": 24, "numMin": 24, "matchSpinsWie": [0]}, "Wiener step"),        # Wiener group of 24 not built
    ({"numMax": 64, "numMin": 64, "matchSpinsWie": [0]}, "Wiener step"),
    ({"blockSizeWie": 32, "matchSpinsWie": [0]}, "blockSizeWie"),
    ({"matchSize": 64, "matchSpinsWie": [0]}, "matchSize"),
])
def test_validate_rejects_unbuilt_combinations(opts, what):
    with pytest.raises(gbm3d_amd.GBM3DError, match=what) as e:
        gbm3d_amd.validate(opts, 128, 96)
    assert e.value.status == _abi.GBM3D_ERR_UNSUPPORTED


def test_validate_rejects_num_min_above_num_max():
    with pytest.raises(gbm3d_amd.GBM3DError, match="numMin") as e:
        gbm3d_amd.validate({"numMax": 8, "numMin": 16, "matchSpinsWie": [0]}, 128, 96)
    assert e.value.status == _abi.GBM3D_ERR_ARG


def test_validate_reports_reference_errors_and_sizes():
    with pytest.raises(gbm3d_amd.GBM3DError, match="matchSpinsWie") as e:
        gbm3d_amd.validate({"blockSize": 8}, 64, 64)              # checkBM3Dopts has no matchSpinsWie default
    assert e.value.status == _abi.GBM3D_ERR_MISSING_FIELD
    with pytest.raises(gbm3d_amd.GBM3DError, match="matchSize"):
        gbm3d_amd.validate(None, 12, 12)                          # pads to 16 x 16 < matchSize
    # the workspace grows with the image
    assert gbm3d_amd.validate(None, 256, 256) > 3 * gbm3d_amd.validate(None, 128, 128)


def test_sigma_wie_must_be_block_size_wie_square():
    from gbm3d_amd.api import check_sigma_wie
    opts = {"blockSize": 8, "blockSizeWie": 8, "matchSpinsWie": [0]}
    check_sigma_wie(opts, np.ones((8, 8)))
    for bad in (np.ones((4, 4)), np.ones((8, 7)), np.ones(64)):
        with pytest.raises(gbm3d_amd.GBM3DError, match="blockSizeWie x blockSizeWie") as e:
            check_sigma_wie(opts, bad)
        assert e.value.status == _abi.GBM3D_ERR_ARG
    check_sigma_wie(opts, None)                                   # computed by the library (GBM3D.m:35)
    check_sigma_wie({"profile": "superFast"}, None)               # no Wiener step: not needed


def test_colored_abi_carries_sigma_wie_size():
    txt = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    decl = re.search(r"int gbm3d_denoise_colored\((.*?)\);", txt, flags=re.S).group(1)
    assert re.search(r"const double\* sigmaWie,\s*int wm,\s*int wn", decl)


# ---- the MATLAB gateway, compiled against a declarations-only mex.h and run
# on a minimal fake mxArray runtime (tests/mex_stub/): its option commands and
# argument checks, none of which needs a GPU.  Not a binding test against
# MATLAB itself (none is installed here).
STUB = os.path.join(ROOT, "tests", "mex_stub")
MEX_SRC = os.path.join(ROOT, "gbm3d_amd", "matlab", "gbm3d_mex.cpp")


def test_mex_gateway_compiles_against_the_mex_api():
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-fsyntax-only", "-I",
                        os.path.join(ROOT, "include"), "-I", STUB, MEX_SRC], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_mex_gateway_option_commands_and_errors(tmp_path):
    import json

    exe = str(tmp_path / "run_gateway")
    libdir = os.path.dirname(_abi.LIB_PATH)
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), "-I", STUB, MEX_SRC,
                        os.path.join(STUB, "fake_mx.cpp"), os.path.join(STUB, "run_gateway.cpp"), "-L", libdir,
                        "-lgbm3d", f"-Wl,-rpath,{libdir}", "-o", exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    res = {d["case"]: d for d in map(json.loads, out.stdout.splitlines())}
    for p in ("default", "fast", "superFast", "accuracy", "unknown"):
        assert res["set:" + p]["opts"] == O.set_opts(p), p
    assert res["set:none"]["opts"] == O.set_opts("default")
    assert res["check:blockSize8"]["opts"] == O.check_opts({"blockSize": 8})
    assert res["check:spins_median"]["opts"] == O.check_opts(
        {"matchSpins": [0, 2, 4], "Wiener": False, "filtType": "median", "numMax": 8, "numMin": 8})
    assert res["check:missing"]["error_id"] == "gbm3d:arg"
    assert res["check:badfilt"]["error_id"] == "gbm3d:filt"        # denoiseAndAggregateGPU.m:54
    assert res["unknown_command"]["error_id"] == "gbm3d:arg"
    assert res["denoise:matrix_sigma"]["error_id"] == "gbm3d:unsupported"
    assert res["denoise:complex"]["error_id"] == "gbm3d:arg"
    assert res["denoise:too_few"]["error_id"] == "gbm3d:arg"
