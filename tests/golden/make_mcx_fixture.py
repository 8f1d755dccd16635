"""Writes tests/golden/mcx_validation.json from the reference's examples/data/mcxValidationData.mat.

The .mat file is MATLAB v5 data (read with scipy.io.loadmat, which executes nothing from the
file). It holds the MCX ground truth ``mcx_res(t)`` that the reference's
examples/mcxValidation3D.m overlays on its own temporal-correlation curves, plus the scalars that
define that configuration (box dims, MFP, dynamic voxel size R, spacing h, D, source-sensor
separation). Only those numbers are copied.
"""
import json
import os
import sys

import numpy as np
import scipy.io

SRC = "/root/reference/examples/data/mcxValidationData.mat"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mcx_validation.json")

if __name__ == "__main__":
    src = sys.argv[1] if len(sys.argv) > 1 else SRC
    d = scipy.io.loadmat(src)
    rec = {
        "source": "examples/data/mcxValidationData.mat (used by examples/mcxValidation3D.m:11-158)",
        "t": np.asarray(d["t"], float).ravel().tolist(),
        "mcx_res": np.asarray(d["mcx_res"], float).ravel().tolist(),
        "D": float(np.asarray(d["D"]).ravel()[0]),
        "MFP": float(np.asarray(d["MFP"]).ravel()[0]),
        "R": float(np.asarray(d["R"]).ravel()[0]),
        "h": float(np.asarray(d["h"]).ravel()[0]),
        "dims": [float(x) for x in np.asarray(d["dims"]).ravel()],
        "viSep": float(np.asarray(d["viSep"]).ravel()[0]),
    }
    with open(OUT, "w") as f:
        json.dump(rec, f, indent=1)
    print("wrote", OUT)
