"""analyzeMotility: the verdict layer over a Radial record (libddm.so `ddm_analyze_motility`).

Mirrors the reference's analyzeMotility (motility-ddm src/motility_fit.js:808-990):
same inputs (the Radial record readRadial returns, with its drift sums, and
AnalysisOptions dxUm / fps / geometry / exposureLocked) and the same result
fields (verdict, physicsVerdict, gates G0-G5, TSI, alpha, vbar, sigmaV, Z,
Ddiff, DdiffBrownian, dBIC, drift, mu, scaling, fit, q, bins, lags, tau, nData,
nEff, residualRho).  The fits run in the library's C++ host code, FP64.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .engine import DdmError, Radial, load_library
from .fit import lag_grid

VERDICTS = ("MOTILE", "NOT_MOTILE", "DIRECTED_TRANSPORT", "NO_DYNAMICS", "UNLOCKED_EXPOSURE")
GEOMETRY = {"3d": 0, "2d": 1}
# gate text of the reference (src/motility_fit.js:918-957); GATES thresholds (:28-38)
_GATE_TEXT = {
    0: ("G0", "Usable q-bins", ">= 4"),
    1: ("G1", "Temporal structure index", ">= 0.15"),
    2: ("G2", "Uniform flow", "not (|V| > 3 sigma and |V| > 2 um/s)"),
    3: ("G3", "Swimmer vs Brownian Delta BIC", "> 10"),
    4: ("G4", "Motile contrast share alpha", "> 0.05, swimmer decay resolved in >= 1 bin"),
    5: ("G5", "Exposure locked", "manual exposure readback"),
}


class RadialRecord(C.Structure):
    _fields_ = [("N", C.c_int32), ("NB", C.c_int32), ("L", C.c_int32), ("knodes", C.c_int32),
                ("kCenter", C.c_void_p), ("kNodes", C.c_void_p), ("D", C.c_void_p), ("drift", C.c_void_p),
                ("nModes", C.c_void_p)]


class MotilityOpts(C.Structure):
    _fields_ = [("dx_um", C.c_double), ("fps", C.c_double), ("geometry", C.c_int32), ("exposure_locked", C.c_int32)]


class DriftEstimate(C.Structure):
    _fields_ = [("Vx", C.c_double), ("Vy", C.c_double), ("mag", C.c_double), ("sigma", C.c_double),
                ("angle_deg", C.c_double), ("samples", C.c_double), ("significant", C.c_int32),
                ("pairs", C.c_int32)]


class Gate(C.Structure):
    _fields_ = [("id", C.c_int32), ("passed", C.c_int32), ("value", C.c_double), ("sigma", C.c_double),
                ("resolved_bins", C.c_int32), ("pad_", C.c_int32)]


class MotilityResult(C.Structure):
    _fields_ = [("verdict", C.c_int32), ("physics_verdict", C.c_int32), ("insufficient", C.c_int32),
                ("ngates", C.c_int32), ("gates", Gate * 6)] + [
        (n, C.c_double) for n in ("TSI", "alpha", "vbar", "sigma_v", "Z", "Ddiff", "Ddiff_brownian", "dBIC", "mu",
                                  "nData", "nEff", "residual_rho", "swim_rss", "swim_Vmag", "brown_rss",
                                  "brown_Vmag")] + [
        ("alpha_resolved", C.c_int32), ("resolved_bins", C.c_int32), ("nbins", C.c_int32), ("nscaling", C.c_int32),
        ("drift", DriftEstimate)]


class MotilityDetail(C.Structure):
    _fields_ = [("cap", C.c_int32), ("pad_", C.c_int32), ("bin", C.c_void_p), ("q", C.c_void_p),
                ("swim_amps", C.c_void_p), ("brown_amps", C.c_void_p), ("scaling_q", C.c_void_p),
                ("scaling_tau_half", C.c_void_p)]


def _declare(L):
    if getattr(L, "_verdict_declared", False):
        return L
    L.ddm_analyze_motility.argtypes = [C.POINTER(RadialRecord), C.POINTER(MotilityOpts), C.POINTER(MotilityResult),
                                       C.POINTER(MotilityDetail)]
    L.ddm_analyze_motility.restype = C.c_int
    L.ddm_verdict_last_error.argtypes = []
    L.ddm_verdict_last_error.restype = C.c_char_p
    L._verdict_declared = True
    return L


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def analyze_motility(radial: Radial, dxUm: float, fps: float, geometry: str = "3d",
                     exposureLocked: bool = True) -> dict:
    """analyzeMotility(radial, {dxUm, fps, geometry, exposureLocked}) -> the reference's MotilityResult."""
    if geometry not in GEOMETRY:
        raise DdmError(f"analyzeMotility: geometry must be '3d' or '2d', got {geometry!r}")
    L = _declare(load_library())
    NB, Lg = int(radial.NB), int(radial.L)
    kc = _f64(radial.kCenter)
    kn = _f64(radial.kNodes) if radial.kNodes is not None and len(radial.kNodes) else np.zeros(0)
    D = _f64(radial.D)
    dr = _f64(radial.drift) if radial.drift is not None else np.zeros(0)
    nm = np.ascontiguousarray(radial.nModes, dtype=np.uint32)
    if len(kc) != NB or len(nm) != NB or D.size != NB * Lg:
        raise DdmError("analyzeMotility: Radial record arrays do not match NB and L")
    if dr.size != NB * Lg * 8:
        raise DdmError("analyzeMotility: the Radial record has no drift sums (readRadial(drift=True))")
    rec = RadialRecord(int(radial.N), NB, Lg, len(kn) // NB if len(kn) else 0, kc.ctypes.data,
                       kn.ctypes.data if len(kn) else None, D.ctypes.data, dr.ctypes.data, nm.ctypes.data)
    opt = MotilityOpts(float(dxUm), float(fps), GEOMETRY[geometry], 1 if exposureLocked else 0)
    res = MotilityResult()
    bins = np.zeros(NB, np.int32)
    q = np.zeros(NB)
    sa = np.zeros(2 * NB)
    ba = np.zeros(2 * NB)
    sq = np.zeros(NB)
    st = np.zeros(NB)
    det = MotilityDetail(NB, 0, bins.ctypes.data, q.ctypes.data, sa.ctypes.data, ba.ctypes.data, sq.ctypes.data,
                         st.ctypes.data)
    rc = L.ddm_analyze_motility(C.byref(rec), C.byref(opt), C.byref(res), C.byref(det))
    if rc != 0:
        raise DdmError(L.ddm_verdict_last_error().decode())
    gates = []
    for g in res.gates[:res.ngates]:
        gid, name, thr = _GATE_TEXT[g.id]
        d = dict(id=gid, name=name, value=float(g.value), threshold=thr, **{"pass": bool(g.passed)})
        if g.id == 2:
            d["sigma"] = float(g.sigma)
        if g.id == 4:
            d["resolvedBins"] = int(g.resolved_bins)
        gates.append(d)
    out = dict(verdict=VERDICTS[res.verdict], physicsVerdict=VERDICTS[res.physics_verdict], gates=gates,
               TSI=float(res.TSI))
    if res.insufficient:
        return out
    nb, ns = int(res.nbins), int(res.nscaling)
    drift = res.drift
    amps = lambda a, alpha: [dict(A=float(a[2 * i]), B=float(a[2 * i + 1]), W=alpha * float(a[2 * i]),
                                  P=(1 - alpha) * float(a[2 * i])) for i in range(nb)]
    dt = 1.0 / fps
    lags = [int(x) for x in lag_grid(Lg)]
    out.update(
        alpha=float(res.alpha), alphaResolved=bool(res.alpha_resolved), resolvedBins=int(res.resolved_bins),
        vbar=float(res.vbar), sigmaV=float(res.sigma_v), Z=float(res.Z), Ddiff=float(res.Ddiff),
        DdiffBrownian=float(res.Ddiff_brownian), dBIC=float(res.dBIC),
        drift=dict(Vx=drift.Vx, Vy=drift.Vy, mag=drift.mag, sigma=drift.sigma, angleDeg=drift.angle_deg,
                   significant=bool(drift.significant), pairs=int(drift.pairs), samples=drift.samples),
        mu=float(res.mu),
        scaling=dict(q=sq[:ns].tolist(), tauHalf=st[:ns].tolist(), mu=float(res.mu)),
        geometry=geometry,
        fit=dict(swimmer=dict(vbar=float(res.vbar), Z=float(res.Z), Dd=float(res.Ddiff), Vmag=float(res.swim_Vmag),
                              alpha=float(res.alpha), rss=float(res.swim_rss), amps=amps(sa, float(res.alpha))),
                 brownian=dict(vbar=0.0, Z=1.0, Dd=float(res.Ddiff_brownian), Vmag=float(res.brown_Vmag), alpha=0.0,
                               rss=float(res.brown_rss), amps=amps(ba, 0.0))),
        q=q[:nb].tolist(), bins=bins[:nb].tolist(), lags=lags, tau=[l * dt for l in lags],
        nData=float(res.nData), nEff=float(res.nEff), residualRho=float(res.residual_rho))
    return out
