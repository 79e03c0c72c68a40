"""Uniform drift velocity from the radial record's phase sums.

`estimate_drift` is estimateDrift (motility-ddm src/motility_fit.js:658-725):
the drift V that best explains the Fourier phase increments of every binned
mode and lag pair, from the 8 sums per (bin, lag) that `DdmEngine.readRadial`
returns in `Radial.drift` (computed on the GPU, src/ddm_ref.js:360-381).  With
a swimmer component, only the lag pairs where its self-ISF has decayed below
0.2 are used.  The arithmetic is the library's host routine
`ddm_estimate_drift` (the one the verdict layer runs), so Python, the Node
facade and `analyzeMotility` share one implementation.

`isf3d` / `isf2d` are the library's swimmer kernels (`ddm_swimmer_isf`): the
3D Schulz-averaged ISF (src/motility_fit.js:135-144) and the tabulated in-plane
one the fits use (isf2dFast, :168-219).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from .engine import DdmError, load_library
from .fit import Prepared

DRIFT_ISF_MAX = 0.2   # GATES.DRIFT_ISF_MAX (src/motility_fit.js:34)
FLOW_SIGMA = 3.0      # GATES.FLOW_SIGMA (:30)
FLOW_MIN_UM_S = 2.0   # GATES.FLOW_MIN_UM_S (:31)
_GEOMETRY = {"3d": 0, "2d": 1}


def _lib():
    from . import verdict as V
    L = V._declare(load_library())
    if not getattr(L, "_drift_declared", False):
        L.ddm_swimmer_isf.argtypes = [C.c_int32, C.c_void_p, C.c_int32, C.c_double, C.c_void_p]
        L.ddm_swimmer_isf.restype = C.c_int
        L.ddm_estimate_drift.argtypes = [C.POINTER(V.RadialRecord), C.POINTER(V.MotilityOpts), C.c_void_p,
                                         C.POINTER(V.DriftEstimate)]
        L.ddm_estimate_drift.restype = C.c_int
        L._drift_declared = True
    return L


def swimmer_isf(geometry: str):
    """g(x; Z) of the given geometry: scalar in, scalar out; arrays in, arrays out."""
    if geometry not in _GEOMETRY:
        raise DdmError(f"swimmer ISF: geometry must be '3d' or '2d', got {geometry!r}")

    def g(x, Z):
        xs = np.ascontiguousarray(np.atleast_1d(x), dtype=np.float64)
        out = np.empty_like(xs)
        L = _lib()
        if L.ddm_swimmer_isf(_GEOMETRY[geometry], xs.ctypes.data, xs.size, float(Z), out.ctypes.data) != 0:
            raise DdmError(L.ddm_verdict_last_error().decode())
        return float(out[0]) if np.ndim(x) == 0 else out.reshape(np.shape(x))
    return g


isf3d = swimmer_isf("3d")
isf2d = swimmer_isf("2d")


@dataclass
class DriftEstimate:
    Vx: float
    Vy: float
    mag: float
    sigma: float
    angleDeg: float
    significant: bool
    pairs: int
    samples: float


def estimate_drift(radial, pr: Prepared, dxUm: float, geometry: str = "3d", swimmer=None) -> DriftEstimate:
    """estimateDrift (src/motility_fit.js:671-725).

    radial: the Radial record with drift sums (readRadial(drift=True));
    pr: `fit.prepare(radial, dxUm, fps)` (its frame interval); swimmer: None or
    (vbar, Z) of the fitted swimmer component, which turns on the lag-pair gating.
    """
    from . import verdict as V
    if radial.drift is None or len(radial.drift) == 0:
        raise ValueError("estimate_drift: the Radial record carries no drift sums (readRadial(drift=True))")
    if geometry not in _GEOMETRY:
        raise DdmError(f"estimate_drift: geometry must be '3d' or '2d', got {geometry!r}")
    NB, L = int(radial.NB), int(radial.L)
    kc = np.ascontiguousarray(radial.kCenter, dtype=np.float64)
    kn = (np.ascontiguousarray(radial.kNodes, dtype=np.float64) if radial.kNodes is not None
          and len(radial.kNodes) else np.zeros(0))
    D = np.ascontiguousarray(radial.D, dtype=np.float64)
    dr = np.ascontiguousarray(radial.drift, dtype=np.float64)
    nm = np.ascontiguousarray(radial.nModes, dtype=np.uint32)
    if dr.size != NB * L * 8 or D.size != NB * L:
        raise DdmError("estimate_drift: Radial record arrays do not match NB and L")
    rec = V.RadialRecord(int(radial.N), NB, L, kn.size // NB, kc.ctypes.data, kn.ctypes.data if kn.size else None,
                         D.ctypes.data, dr.ctypes.data, nm.ctypes.data)
    opt = V.MotilityOpts(float(dxUm), 1.0 / pr.dt, _GEOMETRY[geometry], 1)
    sw = None if swimmer is None else np.array([float(swimmer[0]), float(swimmer[1])])
    out = V.DriftEstimate()
    lib = _lib()
    if lib.ddm_estimate_drift(C.byref(rec), C.byref(opt), None if sw is None else sw.ctypes.data, C.byref(out)) != 0:
        raise DdmError(lib.ddm_verdict_last_error().decode())
    return DriftEstimate(out.Vx, out.Vy, out.mag, out.sigma, out.angle_deg, bool(out.significant), int(out.pairs),
                         out.samples)
