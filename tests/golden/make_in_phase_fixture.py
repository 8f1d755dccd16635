"""Writes tests/golden/in_phase_function.npz from the reference's examples/data/inPhasefunction.mat.

The .mat file is MATLAB v5 data (read with scipy.io.loadmat, which executes nothing from the
file). It holds `mudiffPhaseFunction`, the complex tabular amplitude function (3601 samples over
the scattering angle on [0, 2 pi]) that the reference's examples/spatioTemporalCorrelation2D.m
gives both materials of its 2D correlation-sampler example. Only that array is copied.
"""
import os
import sys

import numpy as np
import scipy.io

SRC = "/root/reference/examples/data/inPhasefunction.mat"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "in_phase_function.npz")

if __name__ == "__main__":
    src = sys.argv[1] if len(sys.argv) > 1 else SRC
    d = scipy.io.loadmat(src)
    np.savez(OUT, amp=np.asarray(d["mudiffPhaseFunction"], np.complex128).ravel())
    print("wrote", OUT)
