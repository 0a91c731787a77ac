"""Writes tests/golden/bior15_taps.json from the reference's filter data file
archive/waveletFilters/bior15Filters.mat (MATLAB v5 data, read with scipy's
data-only loader).  Run where the reference is mounted at /root/reference."""
import json
import os

import scipy.io

SRC = "/root/reference/archive/waveletFilters/bior15Filters.mat"
d = scipy.io.loadmat(SRC)
out = {k: [float(x) for x in d[k].ravel()] for k in ("LoD", "HiD", "LoR", "HiR")}
out["source"] = "archive/waveletFilters/bior15Filters.mat"
json.dump(out, open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "bior15_taps.json"), "w"), indent=1)
