"""MATLAB-style config dict -> ``sstmc_config`` C struct (include/sstmc.h).

The dict mirrors the reference's ``config`` struct field for field (README.md:59-349, parsed by
MatlabInterface.cu:274-1563): ``simulation``, ``medium``, ``scattering``, ``illumination``,
``view``, ``sampler``, ``tracer``.  Numeric arrays may be lists or numpy arrays;
``medium.materialGrid`` may be an (nx, ny[, nz]) array indexed like MATLAB (x fastest in memory
after Fortran-order flattening) or an already-flat vector.  Extra, MI355X-specific keys live
under ``config['precision']`` (``terms``: 'fp32'|'fp64', ``contraction``: 'bf16x3'|'fp32'|'fp64',
``tauCut``: the optical depth past which an NEE term is dropped, 0 = exact)
and ``config['tracer']['output']`` ('field'|'correlation').  ``sampler.type`` is
'TemporalSampler' (default) or 'TemporalCorrelationSampler', which takes ``illumination2`` and
``view2`` as the second source pair.

Validation follows the reference's MEX checks where they exist (material ids in 1..M-1,
MatlabInterface.cu:975-992; D >= 0, :1293-1297; batchSize > 0, :1497-1509) and raises
``ValueError`` with the reference's message text.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

# enum values, kept equal to include/sstmc.h
MEDIUM_HETEROGENEOUS = 1
SCATTERING_ISOTROPIC = 1
SCATTERING_TABULAR = 2
SCATTERING_HG = 3
SCATTERING_VMF = 4
SOURCE_NONE, SOURCE_POINT, SOURCE_FARFIELD, SOURCE_GAUSSIAN_BEAM = 0, 1, 2, 3
SAMPLER_TEMPORAL, SAMPLER_TEMPORAL_CORRELATION = 1, 2
TRACER_NEE = 1
OUTPUT_FIELD, OUTPUT_CORRELATION = 0, 1
TERMS_FP32, TERMS_FP64 = 0, 1
CONTRACT_BF16X3, CONTRACT_FP32, CONTRACT_FP64 = 0, 1, 2

_DP = C.POINTER(C.c_double)
_U32P = C.POINTER(C.c_uint32)
_I32P = C.POINTER(C.c_int32)


class Simulation(C.Structure):
    _fields_ = [("batchSize", C.c_uint32), ("fullIteration", C.c_uint32),
                ("renderingsNum", C.c_uint32), ("deviceNumber", C.c_uint32),
                ("seed", C.c_uint64), ("lambda_", C.c_double), ("graphs", C.c_uint32)]


class Medium(C.Structure):
    _fields_ = [("type", C.c_int32), ("xAxisSize", C.c_uint32), ("yAxisSize", C.c_uint32),
                ("zAxisSize", C.c_uint32), ("xAxis", _DP), ("yAxis", _DP), ("zAxis", _DP),
                ("materialGrid", _U32P), ("materialNum", C.c_uint32), ("sigs", _DP), ("siga", _DP)]


class Scattering(C.Structure):
    _fields_ = [("materialNum", C.c_uint32), ("type", _I32P), ("g", _DP), ("tabularSize", _U32P),
                ("tabular", _DP), ("vmfMixtures", _U32P), ("vmfMu", _DP), ("vmfC", _DP), ("vmfAlpha", _DP)]


class Source(C.Structure):
    _fields_ = [("type", C.c_int32), ("num", C.c_uint32), ("x", _DP), ("y", _DP), ("z", _DP),
                ("isAreaSource", C.c_int32), ("dirX", _DP), ("dirY", _DP), ("dirZ", _DP), ("aperture", C.c_double)]


class Sampler(C.Structure):
    _fields_ = [("type", C.c_int32), ("tNum", C.c_uint32), ("t", _DP), ("D", _DP),
                ("Ux", _DP), ("Uy", _DP), ("Uz", _DP), ("isTemporalAmplitude", C.c_int32)]


class Tracer(C.Structure):
    _fields_ = [("type", C.c_int32), ("isCBS", C.c_int32), ("output", C.c_int32)]


class Precision(C.Structure):
    _fields_ = [("terms", C.c_int32), ("contraction", C.c_int32), ("tauCut", C.c_double)]


class Config(C.Structure):
    _fields_ = [("dims", C.c_uint32), ("simulation", Simulation), ("medium", Medium),
                ("scattering", Scattering), ("illumination", Source), ("view", Source),
                ("sampler", Sampler), ("tracer", Tracer), ("precision", Precision),
                ("illumination2", Source), ("view2", Source)]


class CConfig:
    """Owns an ``sstmc_config`` plus the numpy buffers its pointers reference."""

    def __init__(self, cfg: dict, dims: int):
        self._keep = []
        self.struct = Config()
        self.dims = dims
        self._fill(cfg, dims)

    # -- helpers
    def _arr(self, a, dtype):
        arr = np.ascontiguousarray(np.asarray(a, dtype=dtype).ravel(order="F"))
        self._keep.append(arr)
        return arr

    def _dptr(self, a):
        arr = self._arr(a, np.float64)
        return arr.ctypes.data_as(_DP), arr.size

    def ptr(self):
        """Pointer to the C struct. It holds this object, and with it every array the struct points
        into, for as long as the pointer itself is referenced (so `CConfig(cfg, d).ptr()` passed
        straight to a call stays valid during the call)."""
        p = C.pointer(self.struct)
        p._owner = self
        return p

    def _fill(self, cfg, dims):
        s = self.struct
        s.dims = dims
        sim = cfg.get("simulation", {}) or {}
        bs = sim.get("batchSize", 1024)
        if bs <= 0:
            raise ValueError("Batch size must be a positive integer.")
        s.simulation.batchSize = int(bs)
        fi = sim.get("fullIteration", 1)
        if fi <= 0:
            raise ValueError("Full iteraion type must be non-complex positive double.")
        s.simulation.fullIteration = int(fi)
        rn = sim.get("renderingsNum", 1)
        if rn <= 0:
            raise ValueError("Renderings num type must be non-complex positive double.")
        s.simulation.renderingsNum = int(rn)
        s.simulation.deviceNumber = int(sim.get("deviceNumber", 0))
        s.simulation.seed = int(abs(sim.get("seed", 0)))
        lam = np.atleast_1d(np.asarray(sim.get("lambda", 1.0), dtype=np.float64))
        if lam.size != 1:
            raise ValueError("only one wavelength is supported (Nw = 1)")
        s.simulation.lambda_ = float(lam[0])
        s.simulation.graphs = 1 if sim.get("graphs", False) else 0

        # scattering (MatlabInterface.cu:712-888)
        sc = cfg["scattering"]
        types = sc["type"]
        if isinstance(types, str):
            types = [types]
        amps = sc.get("amplitudeFunction", [0.0] * len(types))
        if not isinstance(amps, (list, tuple)):
            amps = [amps]
        elif len(types) == 1 and len(amps) > 1 and np.isscalar(amps[0]):
            amps = [amps]   # one material given its table as a plain list
        tcode, gvals, tsize, tables = [], [], [], []
        vnum, vmu, vc, valpha = [], [], [], []
        for i, t in enumerate(types):
            g, nt, nv = 0.0, 0, 0
            if t == "Isotropic":
                code = SCATTERING_ISOTROPIC
            elif t in ("HenyeyGreenstein", "HG"):
                code = SCATTERING_HG
                g = float(np.asarray(amps[i]).ravel()[0])
            elif t == "Tabular":
                # complex amplitude samples over the scattering angle (Scattering.cuh:272-429)
                f = np.asarray(amps[i], dtype=np.complex128).ravel(order="F")
                if f.size < 2 or not np.all(np.isfinite(f)) or not np.any(f != 0):
                    raise ValueError("Tabular amplitude function of material %d must hold at least two finite "
                                     "samples, not all zero" % i)
                code, nt = SCATTERING_TABULAR, f.size
                tables.append(np.stack([f.real, f.imag], axis=1).ravel())
            elif t in ("VonMisesFisher", "vMF"):
                # mixture of von Mises-Fisher lobes (MatlabInterface.cu:827-866)
                a = amps[i] if isinstance(amps[i], dict) else {}
                mu = np.asarray(a.get("mixtureMu", []), np.float64).ravel()
                cc = np.asarray(a.get("mixtureC", []), np.float64).ravel()
                al = np.asarray(a.get("mixtureAlpha", []), np.float64).ravel()
                if mu.size == 0 or cc.size != mu.size or al.size != mu.size:
                    raise ValueError("vMF amplitude function of material %d needs mixtureMu, mixtureC and "
                                     "mixtureAlpha of equal, non-zero length" % i)
                if not np.all(al > 0) or not np.all(np.isfinite(mu)) or not np.all(np.isfinite(cc)):
                    raise ValueError("vMF amplitude function of material %d: mixtureAlpha must be positive and "
                                     "mixtureMu, mixtureC finite" % i)
                code, nv = SCATTERING_VMF, mu.size
                vmu.append(mu)
                vc.append(cc)
                valpha.append(al)
            else:
                raise ValueError(
                    "Unsupported scattering type in material %d. Type entered: %s. Available "
                    "scattering types are Isotropic, Tabular, HenyeyGreenstein (HG) and VonMisesFisher (vMF)"
                    % (i, t))
            tcode.append(code)
            gvals.append(g)
            tsize.append(nt)
            vnum.append(nv)
        n_sc = len(tcode)
        material_num = 2 if n_sc == 1 else n_sc
        ta = self._arr(tcode, np.int32)
        s.scattering.materialNum = n_sc
        s.scattering.type = ta.ctypes.data_as(_I32P)
        s.scattering.g, _ = self._dptr(gvals)
        if tables:
            sz = self._arr(tsize, np.uint32)
            s.scattering.tabularSize = sz.ctypes.data_as(_U32P)
            s.scattering.tabular, _ = self._dptr(np.concatenate(tables))
        if vmu:
            vn = self._arr(vnum, np.uint32)
            s.scattering.vmfMixtures = vn.ctypes.data_as(_U32P)
            s.scattering.vmfMu, _ = self._dptr(np.concatenate(vmu))
            s.scattering.vmfC, _ = self._dptr(np.concatenate(vc))
            s.scattering.vmfAlpha, _ = self._dptr(np.concatenate(valpha))

        # medium (MatlabInterface.cu:890-1060)
        md = cfg["medium"]
        if md.get("type") != "Heterogeneous":
            raise ValueError("Unsupported medium type. Available medium type is Heterogeneous")
        s.medium.type = MEDIUM_HETEROGENEOUS
        axes = [np.asarray(md["xAxis"], np.float64).ravel(), np.asarray(md["yAxis"], np.float64).ravel()]
        if dims == 3:
            axes.append(np.asarray(md["zAxis"], np.float64).ravel())
        for a in axes:
            if a.size < 2 or np.any(np.diff(a) <= 0):
                raise ValueError("medium axes must hold at least two increasing values")
        s.medium.xAxis, s.medium.xAxisSize = self._dptr(axes[0])
        s.medium.yAxis, s.medium.yAxisSize = self._dptr(axes[1])
        if dims == 3:
            s.medium.zAxis, s.medium.zAxisSize = self._dptr(axes[2])
        ncell = int(np.prod([a.size - 1 for a in axes]))
        grid = np.asarray(md["materialGrid"])
        grid = grid.ravel(order="F") if grid.ndim > 1 else grid.ravel()
        if grid.size != ncell:
            raise ValueError("materialGrid must hold %d elements" % ncell)
        if np.any(grid == 0):
            raise ValueError("Material 0 is only allowed outside the grid. Allowed material numbers "
                             "are from 1 to %d." % (material_num - 1))
        if np.any(grid >= material_num):
            raise ValueError("Material %d is not defined. Allowed material numbers are from 1 to %d."
                             % (int(grid.max()), material_num - 1))
        ga = self._arr(grid, np.uint32)
        s.medium.materialGrid = ga.ctypes.data_as(_U32P)
        sigs = np.asarray(md["sigs"], np.float64).ravel()
        if sigs.size != material_num:
            raise ValueError("sigs must hold %d elements" % material_num)
        s.medium.sigs, _ = self._dptr(sigs)
        s.medium.materialNum = material_num
        if md.get("siga") is not None:
            siga = np.asarray(md["siga"], np.float64).ravel()
            if siga.size != material_num:
                raise ValueError("siga must hold %d elements" % material_num)
            s.medium.siga, _ = self._dptr(siga)

        # sources (MatlabInterface.cu:1062-1215)
        for name in ("illumination", "view"):
            self._source(getattr(s, name), cfg[name], dims, name)

        # sampler (MatlabInterface.cu:1217-1368)
        sp = cfg["sampler"]
        stype = sp.get("type", "TemporalSampler")
        if stype == "TemporalSampler":
            s.sampler.type = SAMPLER_TEMPORAL
        elif stype == "TemporalCorrelationSampler":
            s.sampler.type = SAMPLER_TEMPORAL_CORRELATION
            # MatlabInterface.cu:1313-1331
            if cfg.get("illumination2") is None or cfg.get("view2") is None:
                raise ValueError("For correlation sampler, source.illuminationsHandle2 and "
                                 "source.viewsHandle2 must be entered")
            self._source(s.illumination2, cfg["illumination2"], dims, "illumination2")
            self._source(s.view2, cfg["view2"], dims, "view2")
            if s.illumination2.num != s.illumination.num:
                raise ValueError("Illumination 1 in size of %d, Illumination 2 in size of %d. Both illuminations "
                                 "must be at the same size" % (s.illumination.num, s.illumination2.num))
            if s.view2.num != s.view.num:
                raise ValueError("View 1 in size of %d, View 2 in size of %d. Both views must be at the same size"
                                 % (s.view.num, s.view2.num))
        else:
            raise ValueError("Unsupported sapmler type. Available sampler types are TemporalSampler and "
                             "TemporalCorrelationSampler")
        t = np.asarray(sp.get("t", [0.0]), np.float64).ravel()
        s.sampler.t, s.sampler.tNum = self._dptr(t)
        D = np.asarray(sp.get("D", np.zeros(material_num)), np.float64).ravel()
        if D.size != material_num:
            raise ValueError("D must hold %d elements" % material_num)
        if np.any(D < 0):
            raise ValueError("D values in temporal sampler must be non-negative")
        s.sampler.D, _ = self._dptr(D)
        U = sp.get("U", {}) or {}
        s.sampler.Ux, _ = self._dptr(np.broadcast_to(np.asarray(U.get("x", 0.0), np.float64), (material_num,)))
        s.sampler.Uy, _ = self._dptr(np.broadcast_to(np.asarray(U.get("y", 0.0), np.float64), (material_num,)))
        s.sampler.Uz, _ = self._dptr(np.broadcast_to(np.asarray(U.get("z", 0.0), np.float64), (material_num,)))
        s.sampler.isTemporalAmplitude = 1 if sp.get("isTemporalAmplitude", True) else 0

        tr = cfg.get("tracer", {"type": "nee"})
        if tr.get("type", "nee") != "nee":
            raise ValueError("Unsupported tracer type. Available tracer type is nee")
        s.tracer.type = TRACER_NEE
        s.tracer.isCBS = 1 if tr.get("isCBS", False) else 0
        out = tr.get("output", "field")
        s.tracer.output = {"field": OUTPUT_FIELD, "correlation": OUTPUT_CORRELATION}[out]

        pr = cfg.get("precision", {}) or {}
        s.precision.terms = {"fp32": TERMS_FP32, "fp64": TERMS_FP64}[pr.get("terms", "fp32")]
        s.precision.contraction = {"bf16x3": CONTRACT_BF16X3, "fp32": CONTRACT_FP32,
                                   "fp64": CONTRACT_FP64}[pr.get("contraction", "bf16x3")]
        cut = float(pr.get("tauCut", 0.0))
        if not (cut >= 0 and np.isfinite(cut)):
            raise ValueError("precision.tauCut must be a finite non-negative optical depth (0: exact)")
        s.precision.tauCut = cut

    def _source(self, dst, src, dims, name):
        ty = src["type"]
        if ty == "PointSource":
            dst.type = SOURCE_POINT
            loc = src["location"]
        elif ty == "FarField":
            dst.type = SOURCE_FARFIELD
            loc = src["direction"]
        elif ty == "GaussianBeam":
            # MatlabInterface.cu:1153-1190: focalPoint, focalDirection (one per focal point), aperture
            dst.type = SOURCE_GAUSSIAN_BEAM
            loc = src["focalPoint"]
            fd = src["focalDirection"]
            n = np.atleast_1d(np.asarray(loc["x"])).size
            for k, attr in (("x", "dirX"), ("y", "dirY"), ("z", "dirZ"))[:dims]:
                v = np.atleast_1d(np.asarray(fd[k], np.float64)).ravel()
                if v.size != n:
                    raise ValueError("%s focalDirection must hold one direction per focal point" % name)
                setattr(dst, attr, self._dptr(v)[0])
            ap = float(np.asarray(src["aperture"]).ravel()[0])
            if not ap > 0:
                raise ValueError("%s aperture must be positive" % name)
            dst.aperture = ap
        else:
            raise ValueError("Unsupported %s source type %s. Available types are PointSource, FarField and "
                             "GaussianBeam" % (name, ty))
        xs = np.atleast_1d(np.asarray(loc["x"], np.float64)).ravel()
        ys = np.atleast_1d(np.asarray(loc["y"], np.float64)).ravel()
        if xs.size != ys.size:
            raise ValueError("%s x and y must have the same size" % name)
        dst.x, dst.num = self._dptr(xs)
        dst.y, _ = self._dptr(ys)
        if dims == 3:
            zs = np.atleast_1d(np.asarray(loc["z"], np.float64)).ravel()
            if zs.size != xs.size:
                raise ValueError("%s x and z must have the same size" % name)
            dst.z, _ = self._dptr(zs)
        dst.isAreaSource = 1 if src.get("isAreaSource", False) else 0


def sizes(cfg: dict, dims: int):
    """(Nl, Nv, Nt, renderingsNum) of a config dict."""
    def n(src):
        loc = {"PointSource": "location", "GaussianBeam": "focalPoint"}.get(src["type"], "direction")
        loc = src[loc]
        return int(np.atleast_1d(np.asarray(loc["x"])).size)
    nt = int(np.asarray(cfg["sampler"].get("t", [0.0])).size)
    rn = int((cfg.get("simulation") or {}).get("renderingsNum", 1))
    return n(cfg["illumination"]), n(cfg["view"]), nt, rn
