"""Writes tests/golden/mudiff_cbs.npz from the reference's examples/data/mudiffCBS.mat.

The .mat file is MATLAB v5 data (read with scipy.io.loadmat, which executes nothing from the
file). It holds what the reference's examples/mudiffCBSValidation2D.m uses: the tabular amplitude
function `ampFunc` (complex, over the scattering angle on [0, 2 pi]), the view angles `degList`
(radians), the slab extents `x`, `y`, the mean free path `MFP`, and the reference solution
`mudiffRes`: the mean speckle intensity per view angle, including the coherent-backscattering
cone, that the example overlays on its Monte Carlo field results. Only those arrays are copied.
"""
import os
import sys

import numpy as np
import scipy.io

SRC = "/root/reference/examples/data/mudiffCBS.mat"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mudiff_cbs.npz")

if __name__ == "__main__":
    src = sys.argv[1] if len(sys.argv) > 1 else SRC
    d = scipy.io.loadmat(src)
    np.savez(OUT,
             amp=np.asarray(d["ampFunc"], np.complex128).ravel(),
             deg=np.asarray(d["degList"], np.float64).ravel(),
             mudiff=np.asarray(d["mudiffRes"], np.float64).ravel(),
             x=np.asarray(d["x"], np.float64).ravel(),
             y=np.asarray(d["y"], np.float64).ravel(),
             mfp=np.float64(np.asarray(d["MFP"]).ravel()[0]))
    print("wrote", OUT)
