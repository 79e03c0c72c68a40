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
    L.ddm_analyze_motility_batch.argtypes = [C.POINTER(RadialRecord), C.POINTER(MotilityOpts), C.c_int32,
                                             C.POINTER(MotilityResult), C.POINTER(MotilityDetail), C.c_int32]
    L.ddm_analyze_motility_batch.restype = C.c_int
    L._verdict_declared = True
    return L


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class _Prepared:
    """ctypes inputs and output buffers of one record (kept alive for the call)."""

    def __init__(self, radial: Radial, dxUm: float, fps: float, geometry: str, exposureLocked: bool):
        if geometry not in GEOMETRY:
            raise DdmError(f"analyzeMotility: geometry must be '3d' or '2d', got {geometry!r}")
        NB, Lg = int(radial.NB), int(radial.L)
        self.kc = _f64(radial.kCenter)
        self.kn = _f64(radial.kNodes) if radial.kNodes is not None and len(radial.kNodes) else np.zeros(0)
        self.D = _f64(radial.D)
        self.dr = _f64(radial.drift) if radial.drift is not None else np.zeros(0)
        self.nm = np.ascontiguousarray(radial.nModes, dtype=np.uint32)
        if len(self.kc) != NB or len(self.nm) != NB or self.D.size != NB * Lg:
            raise DdmError("analyzeMotility: Radial record arrays do not match NB and L")
        if self.dr.size != NB * Lg * 8:
            raise DdmError("analyzeMotility: the Radial record has no drift sums (readRadial(drift=True))")
        self.rec = RadialRecord(int(radial.N), NB, Lg, len(self.kn) // NB if len(self.kn) else 0,
                                self.kc.ctypes.data, self.kn.ctypes.data if len(self.kn) else None,
                                self.D.ctypes.data, self.dr.ctypes.data, self.nm.ctypes.data)
        self.opt = MotilityOpts(float(dxUm), float(fps), GEOMETRY[geometry], 1 if exposureLocked else 0)
        self.L, self.fps, self.geometry = Lg, float(fps), geometry
        self.bins = np.zeros(NB, np.int32)
        self.q = np.zeros(NB)
        self.sa = np.zeros(2 * NB)
        self.ba = np.zeros(2 * NB)
        self.sq = np.zeros(NB)
        self.st = np.zeros(NB)
        self.det = MotilityDetail(NB, 0, self.bins.ctypes.data, self.q.ctypes.data, self.sa.ctypes.data,
                                  self.ba.ctypes.data, self.sq.ctypes.data, self.st.ctypes.data)

    def result(self, res: MotilityResult) -> dict:
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

        def amps(a, alpha):
            return [dict(A=float(a[2 * i]), B=float(a[2 * i + 1]), W=alpha * float(a[2 * i]),
                         P=(1 - alpha) * float(a[2 * i])) for i in range(nb)]
        dt = 1.0 / self.fps
        lags = [int(x) for x in lag_grid(self.L)]
        out.update(
            alpha=float(res.alpha), alphaResolved=bool(res.alpha_resolved), resolvedBins=int(res.resolved_bins),
            vbar=float(res.vbar), sigmaV=float(res.sigma_v), Z=float(res.Z), Ddiff=float(res.Ddiff),
            DdiffBrownian=float(res.Ddiff_brownian), dBIC=float(res.dBIC),
            drift=dict(Vx=drift.Vx, Vy=drift.Vy, mag=drift.mag, sigma=drift.sigma, angleDeg=drift.angle_deg,
                       significant=bool(drift.significant), pairs=int(drift.pairs), samples=drift.samples),
            mu=float(res.mu),
            scaling=dict(q=self.sq[:ns].tolist(), tauHalf=self.st[:ns].tolist(), mu=float(res.mu)),
            geometry=self.geometry,
            fit=dict(swimmer=dict(vbar=float(res.vbar), Z=float(res.Z), Dd=float(res.Ddiff),
                                  Vmag=float(res.swim_Vmag), alpha=float(res.alpha), rss=float(res.swim_rss),
                                  amps=amps(self.sa, float(res.alpha))),
                     brownian=dict(vbar=0.0, Z=1.0, Dd=float(res.Ddiff_brownian), Vmag=float(res.brown_Vmag),
                                   alpha=0.0, rss=float(res.brown_rss), amps=amps(self.ba, 0.0))),
            q=self.q[:nb].tolist(), bins=self.bins[:nb].tolist(), lags=lags, tau=[l * dt for l in lags],
            nData=float(res.nData), nEff=float(res.nEff), residualRho=float(res.residual_rho))
        return out


def analyze_motility(radial: Radial, dxUm: float, fps: float, geometry: str = "3d",
                     exposureLocked: bool = True) -> dict:
    """analyzeMotility(radial, {dxUm, fps, geometry, exposureLocked}) -> the reference's MotilityResult."""
    L = _declare(load_library())
    p = _Prepared(radial, dxUm, fps, geometry, exposureLocked)
    res = MotilityResult()
    if L.ddm_analyze_motility(C.byref(p.rec), C.byref(p.opt), C.byref(res), C.byref(p.det)) != 0:
        raise DdmError(L.ddm_verdict_last_error().decode())
    return p.result(res)


def analyze_motility_batch(radials, dxUm, fps, geometry="3d", exposureLocked=True, threads: int = 0) -> list:
    """analyzeMotility of many records at once on the library's host threads (threads <= 0: one per
    hardware thread).  dxUm, fps, geometry and exposureLocked are scalars or one per record."""
    L = _declare(load_library())
    n = len(radials)
    per = lambda v: list(v) if isinstance(v, (list, tuple, np.ndarray)) else [v] * n
    ps = [_Prepared(r, d, f, g, x) for r, d, f, g, x in
          zip(radials, per(dxUm), per(fps), per(geometry), per(exposureLocked))]
    recs = (RadialRecord * max(n, 1))(*[p.rec for p in ps])
    opts = (MotilityOpts * max(n, 1))(*[p.opt for p in ps])
    dets = (MotilityDetail * max(n, 1))(*[p.det for p in ps])
    res = (MotilityResult * max(n, 1))()
    if L.ddm_analyze_motility_batch(recs, opts, n, res, dets, int(threads)) != 0:
        raise DdmError(L.ddm_verdict_last_error().decode())
    return [p.result(res[i]) for i, p in enumerate(ps)]
