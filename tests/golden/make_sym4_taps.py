"""Writes tests/golden/sym4_taps.json from the reference's filter data file
archive/waveletFilters/sym4Filters.mat (MATLAB v5 data, read with scipy's
data-only loader; the file getWaveFilters2D('sym', 4) loads,
waveTransforms2D/getWaveFilters2D.m:7-16).  Run where the reference is
mounted at /root/reference."""
import json
import os

import scipy.io

SRC = "/root/reference/archive/waveletFilters/sym4Filters.mat"
d = scipy.io.loadmat(SRC)
out = {k: [float(x) for x in d[k].ravel()] for k in ("Lo_d", "Hi_d")}
out["source"] = "archive/waveletFilters/sym4Filters.mat"
json.dump(out, open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "sym4_taps.json"), "w"), indent=1)
