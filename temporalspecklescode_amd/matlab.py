"""MATLAB front end: the reference's ``[u, iter] = sstmc2D(config)`` / ``sstmc3D(config)`` MEX call
(MatlabInterface.cu:1565-1825) driven by a ``.mat`` file.

A MATLAB user saves the struct the reference's MEX takes (``save('cfg.mat', 'config')``, MAT-file
version 5 or 7) and runs

    python -m temporalspecklescode_amd.matlab cfg.mat --out result.mat

which renders on the GPU and writes ``u`` (Nl x Nv x Nt x renderingsNum, complex) and ``iter``
(1 x renderingsNum) into ``result.mat`` for ``load('result.mat')``. ``load_config`` turns the
struct into the dict form of :mod:`temporalspecklescode_amd.config`: structs become dicts, cell
arrays lists, char arrays strings; numeric arrays keep MATLAB's index order (``materialGrid(i, j,
k)`` is ``grid[i - 1, j - 1, k - 1]``). As in the MEX (getCellArray of ``scattering.type`` and
``scattering.amplitudeFunction``, MatlabInterface.cu:719-866), the scattering fields are cells with
one entry per material; a one-material cell, which the MAT-file reader collapses to its content,
is restored here.
"""
from __future__ import annotations

import argparse
import sys

import numpy as np
import scipy.io as sio


def _convert(x):
    if isinstance(x, sio.matlab.mat_struct):
        return {k: _convert(getattr(x, k)) for k in x._fieldnames}
    if isinstance(x, np.ndarray):
        if x.dtype == object:                      # cell array (column-major, as MATLAB numbers it)
            return [_convert(e) for e in x.ravel(order="F")]
        if x.dtype.kind in "US":                   # char matrix / cell of equal-length strings
            return [str(e) for e in x.ravel(order="F")]
    if isinstance(x, np.generic):
        return x.item()
    return x


def _per_material(value, n, kinds):
    """A cell with one entry per material, from what the MAT-file reader returns for it."""
    if isinstance(value, list):
        return value
    if n == 1:
        return [value]                             # {x} came back as x
    if isinstance(value, np.ndarray) and value.ndim == 1 and value.size == n and kinds <= {"HG", "HenyeyGreenstein",
                                                                                          "Isotropic"}:
        return [float(v) for v in value]           # a numeric cell {g1, g2, ...}
    return [value]


def load_config(path: str, name: str = "config") -> dict:
    """The MATLAB struct variable `name` of a .mat file as a config dict."""
    try:
        mat = sio.loadmat(path, squeeze_me=True, struct_as_record=False, chars_as_strings=True)
    except NotImplementedError as e:               # MAT-file 7.3 is HDF5
        raise ValueError("%s: MAT-file 7.3 (HDF5) is not read; save with -v7" % path) from e
    if name not in mat:
        raise KeyError("%s holds no variable '%s' (variables: %s)"
                       % (path, name, ", ".join(k for k in mat if not k.startswith("__"))))
    cfg = _convert(mat[name])
    if not isinstance(cfg, dict):
        raise ValueError("'%s' in %s is not a struct" % (name, path))
    sc = cfg.get("scattering")
    if isinstance(sc, dict) and "type" in sc:
        types = sc["type"] if isinstance(sc["type"], list) else [sc["type"]]
        sc["type"] = types
        if "amplitudeFunction" in sc:
            sc["amplitudeFunction"] = _per_material(sc["amplitudeFunction"], len(types), set(types))
    return cfg


def dims_of(cfg: dict) -> int:
    """2 or 3: sstmc3D when the medium has a zAxis, else sstmc2D."""
    return 3 if "zAxis" in (cfg.get("medium") or {}) else 2


def run(path: str, out: str | None = None, dims: int | None = None, name: str = "config"):
    """Render the config of a .mat file; returns (u, iter) and, with `out`, saves them as MATLAB
    variables u and iter."""
    from . import sstmc2D, sstmc3D
    cfg = load_config(path, name)
    d = dims or dims_of(cfg)
    u, it = (sstmc3D if d == 3 else sstmc2D)(cfg)
    if out:
        sio.savemat(out, {"u": u, "iter": np.asarray(it, np.float64).reshape(1, -1)})
    return u, it


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m temporalspecklescode_amd.matlab",
                                 description="Render a MATLAB config struct saved in a .mat file.")
    ap.add_argument("mat")
    ap.add_argument("--out", help="write u and iter into this .mat file")
    ap.add_argument("--dims", type=int, choices=(2, 3), help="default: 3 when medium.zAxis is present")
    ap.add_argument("--name", default="config", help="the struct variable (default: config)")
    ap.add_argument("--check", action="store_true", help="validate only, without a device")
    a = ap.parse_args(argv)
    try:
        cfg = load_config(a.mat, a.name)
        d = a.dims or dims_of(cfg)
        if a.check:
            from . import _native
            from .config import CConfig
            rc = _native.lib().sstmc_validate(CConfig(cfg, d).ptr())
            print('{"valid": %s, "code": %d, "dims": %d}' % ("true" if rc == 0 else "false", rc, d))
            return rc
        u, it = run(a.mat, a.out, d, a.name)
    except (KeyError, ValueError) as e:
        print("sstmc: %s" % e, file=sys.stderr)
        return 7
    print('{"dims": %d, "shape": [%s], "iter": [%s]}' % (d, ", ".join(str(s) for s in u.shape),
                                                          ", ".join("%.17g" % x for x in np.ravel(it))))
    return 0


if __name__ == "__main__":
    sys.exit(main())
