"""The committed golden fixtures (tests/golden/*.npz, written by the FP64
oracle) stay reproducible, and agree with the FFT-literal transcription."""
import glob
import json
import os

import numpy as np
import pytest

import oracle as O
from oracle import literal as L

FIX = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))


@pytest.mark.parametrize("path", FIX, ids=[os.path.basename(p) for p in FIX])
def test_oracle_reproduces_fixture(path):
    f = np.load(path, allow_pickle=False)
    opts = json.loads(str(f["opts"]))
    U, U1 = O.gbm3d(f["noisy"], float(f["sigma"]), opts, precision=64)
    np.testing.assert_allclose(U, f["U"], rtol=0, atol=1e-10)
    np.testing.assert_allclose(U1, f["U1"], rtol=0, atol=1e-10)


@pytest.mark.parametrize("path", [p for p in FIX if "64" in os.path.basename(p)])
def test_literal_transcription_reproduces_fixture(path):
    f = np.load(path, allow_pickle=False)
    ro = json.loads(str(f["resolved"]))
    U, U1 = L.GBM3D(f["noisy"].astype(np.float64), float(f["sigma"]), ro)
    np.testing.assert_allclose(U, f["U"], atol=1e-8)
    np.testing.assert_allclose(U1, f["U1"], atol=1e-8)
