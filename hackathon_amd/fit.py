"""Per-q motility fits on the GPU (libddm.so `ddm_fit`).

`prepare` and `bounds` restate the reference's data preparation
(motility-ddm src/motility_fit.js:383-401 and :480-491): bins with at least 3
modes and a positive maximum, the quasi-logarithmic lag grid (:358-369), q from
kCenter, q-nodes from kNodes, weight w = 1/max D.  `Fitter.fit` runs one
batched Levenberg-Marquardt launch over any number of (video, bin) problems;
`analyzeMotilityLM` is the per-q counterpart of `analyzeMotility`'s fitting
part (:870-911) for one Radial record.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from .engine import DdmError, Radial, load_library

MAXT = 128
MAXK = 8
BROWNIAN, SWIMMER_3D, SWIMMER_2D = 0, 1, 2   # SWIMMER_2D: the in-plane ISF (isf2dFast table)
SEED_GRID, SEED_GIVEN, SEED_NONLINEAR = 0, 1, 2
SWIM_MIN_UM_S = 2.0   # GATES.SWIM_MIN_UM_S (src/motility_fit.js:37)


class FitBatch(C.Structure):
    _fields_ = [("nfits", C.c_int32), ("ntau_max", C.c_int32), ("knodes", C.c_int32), ("model", C.c_int32),
                ("seed_mode", C.c_int32), ("nstarts", C.c_int32), ("max_iter", C.c_int32),
                ("ntau", C.c_void_p), ("tau", C.c_void_p), ("data", C.c_void_p), ("weight", C.c_void_p),
                ("qnodes", C.c_void_p), ("bounds", C.c_void_p), ("seeds", C.c_void_p), ("vmag", C.c_void_p)]


class GlobalFitBatch(C.Structure):
    _fields_ = [("nvideos", C.c_int32), ("ntau_max", C.c_int32), ("knodes", C.c_int32), ("model", C.c_int32),
                ("seed_mode", C.c_int32), ("nstarts", C.c_int32), ("max_iter", C.c_int32),
                ("bin_offset", C.c_void_p), ("ntau", C.c_void_p), ("tau", C.c_void_p), ("data", C.c_void_p),
                ("weight", C.c_void_p), ("qnodes", C.c_void_p), ("bounds", C.c_void_p), ("seeds", C.c_void_p),
                ("vmag", C.c_void_p)]


def _declare(L):
    if getattr(L, "_fit_declared", False):
        return L
    vp = C.c_void_p
    L.ddm_fit_global.argtypes = [vp, C.POINTER(GlobalFitBatch), vp, vp, vp, vp, vp, vp]
    L.ddm_fit_global.restype = C.c_int
    L.ddm_fitter_create.argtypes = [C.POINTER(vp)]
    L.ddm_fitter_create.restype = C.c_int
    L.ddm_fitter_destroy.argtypes = [vp]
    L.ddm_fitter_destroy.restype = None
    L.ddm_fitter_last_error.argtypes = [vp]
    L.ddm_fitter_last_error.restype = C.c_char_p
    L.ddm_fit.argtypes = [vp, C.POINTER(FitBatch), vp, vp, vp, vp]
    L.ddm_fit.restype = C.c_int
    L.ddm_isf2d_eval.argtypes = [vp, vp, vp, C.c_int32, C.c_int32, vp]
    L.ddm_isf2d_eval.restype = C.c_int
    L._fit_declared = True
    return L


def _helpers(L):
    if not getattr(L, "_fit_helpers_declared", False):
        L.ddm_lag_grid.argtypes = [C.c_int32, C.c_void_p, C.c_int32]
        L.ddm_lag_grid.restype = C.c_int
        L.ddm_usable_bins.argtypes = [C.c_void_p, C.c_void_p, C.c_int32]
        L.ddm_usable_bins.restype = C.c_int
        L._fit_helpers_declared = True
    return L


def lag_grid(L: int) -> np.ndarray:
    """The lags the fits sample (src/motility_fit.js:358-369): the library's ddm_lag_grid."""
    lib = _helpers(load_library())
    n = lib.ddm_lag_grid(int(L), None, 0)
    if n < 0:
        raise DdmError("lagGrid: L must be >= 2")
    out = np.zeros(n, np.int32)
    lib.ddm_lag_grid(int(L), out.ctypes.data, n)
    return out.astype(np.int64)


@dataclass
class PreparedBin:
    b: int
    q: float
    qn: np.ndarray
    d: np.ndarray
    w: float


@dataclass
class Prepared:
    bins: list
    lags: np.ndarray
    tau: np.ndarray
    dt: float
    L: int
    tauMax: float


def usable_bins(radial: Radial) -> np.ndarray:
    """prepare's bin selection (src/motility_fit.js:383-401), the library's ddm_usable_bins: nModes >= 3 and
    a finite positive maximum of D over the lag grid."""
    from .verdict import RadialRecord
    NB, L = int(radial.NB), int(radial.L)
    kc = np.ascontiguousarray(radial.kCenter, dtype=np.float64)
    D = np.ascontiguousarray(radial.D, dtype=np.float64)
    nm = np.ascontiguousarray(radial.nModes, dtype=np.uint32)
    rec = RadialRecord(int(radial.N), NB, L, 0, kc.ctypes.data, None, D.ctypes.data, None, nm.ctypes.data)
    lib = _helpers(load_library())
    out = np.zeros(NB, np.int32)
    n = lib.ddm_usable_bins(C.byref(rec), out.ctypes.data, NB)
    if n < 0:
        raise DdmError("prepare: invalid Radial record")
    return out[:n]


def prepare(radial: Radial, dxUm: float, fps: float) -> Prepared:
    """prepare (src/motility_fit.js:383-401): the usable bins with their data on the lag grid, q from kCenter,
    the q-nodes from kNodes (in rad/um) and the weight 1 / max D."""
    NB, L = radial.NB, radial.L
    dt = 1.0 / fps
    lags = lag_grid(L)
    D = np.asarray(radial.D, dtype=np.float64).reshape(NB, L)
    per_k = 2 * np.pi / (radial.N * dxUm)
    kn = radial.kNodes
    K = len(kn) // NB if kn is not None and len(kn) else 0
    nodes = np.asarray(kn, dtype=np.float64).reshape(NB, K) * per_k if K else None
    bins = [PreparedBin(int(b), per_k * float(radial.kCenter[b]),
                        nodes[b] if K else np.array([per_k * float(radial.kCenter[b])]), D[b, lags],
                        1.0 / float(D[b, lags].max()))
            for b in usable_bins(radial)]
    return Prepared(bins, lags, lags * dt, dt, L, (L - 1) * dt)


def bounds(pr: Prepared) -> np.ndarray:
    """vmin, vmax, Dmin, Dmax (src/motility_fit.js:480-491)."""
    qs = np.array([b.q for b in pr.bins])
    qmin, qmax = qs.min(), qs.max()
    return np.array([max(SWIM_MIN_UM_S, 0.3 / (qmax * pr.tauMax)), 3 / (qmin * pr.dt),
                     0.01 / (qmax * qmax * pr.tauMax), 10 / (qmin * qmin * pr.dt)])


@dataclass
class FitResult:
    theta: np.ndarray     # (nfits, 6): A, B, alpha, ln vbar, ln Z, ln D
    chi2: np.ndarray
    iters: np.ndarray
    kernel_ms: float

    @property
    def alpha(self):
        return self.theta[:, 2]

    @property
    def vbar(self):
        return np.exp(self.theta[:, 3])

    @property
    def Z(self):
        return np.exp(self.theta[:, 4])

    @property
    def D(self):
        return np.exp(self.theta[:, 5])


class Fitter:
    """A fitter context (one HIP stream, device buffers reused across calls)."""

    def __init__(self):
        self._lib = _declare(load_library())
        h = C.c_void_p()
        if self._lib.ddm_fitter_create(C.byref(h)) != 0:
            raise DdmError("ddm_fitter_create failed")
        self._h = h

    def close(self):
        if self._h:
            self._lib.ddm_fitter_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def isf2d(self, x, Z, exact=True) -> np.ndarray:
        """The in-plane swimmer ISF on the GPU in FP64: isf2d by the Legendre quadrature (exact) or
        isf2dFast, the device-built table the in-plane fits interpolate (src/motility_fit.js:120-220)."""
        x, Z = np.broadcast_arrays(np.asarray(x, np.float64), np.asarray(Z, np.float64))
        x, Z = np.ascontiguousarray(x.ravel()), np.ascontiguousarray(Z.ravel())
        out = np.zeros(len(x))
        if self._lib.ddm_isf2d_eval(self._h, x.ctypes.data, Z.ctypes.data, len(x), int(bool(exact)),
                                    out.ctypes.data) != 0:
            raise DdmError(self._lib.ddm_fitter_last_error(self._h).decode())
        return out

    def fit(self, tau, data, weight, qnodes, bnds, model=SWIMMER_3D, seeds=None, nstarts=3,
            max_iter=400, seed_ab=True, vmag=None) -> FitResult:
        """tau/data: lists (or 2D arrays) of per-fit lag times and D values; weight (nfits,);
        qnodes (nfits, K); bnds (nfits, 4); seeds (nfits, 6) for seeded LM, else grid seeding.
        seed_ab=False takes only the nonlinear parameters from the seeds and solves A, B by NNLS.
        vmag: a fixed uniform drift speed per fit (scalar or (nfits,)), E = exp(-D q^2 tau) J0(q vmag tau)."""
        n = len(data)
        ntau = np.array([len(x) for x in data], dtype=np.int32)
        if n == 0:
            return FitResult(np.zeros((0, 6)), np.zeros(0), np.zeros(0, np.int32), 0.0)
        tmax = int(ntau.max())
        if tmax > MAXT:
            raise DdmError(f"at most {MAXT} lags per fit")
        T = np.zeros((n, tmax))
        Dd = np.zeros((n, tmax))
        for i in range(n):
            T[i, :ntau[i]] = tau[i]
            Dd[i, :ntau[i]] = data[i]
        w = np.ascontiguousarray(weight, dtype=np.float64)
        qn = np.ascontiguousarray(qnodes, dtype=np.float64).reshape(n, -1)
        bd = np.ascontiguousarray(bnds, dtype=np.float64).reshape(n, 4)
        sd = None if seeds is None else np.ascontiguousarray(seeds, dtype=np.float64).reshape(n, 6)
        vm = None if vmag is None else np.ascontiguousarray(np.broadcast_to(np.asarray(vmag, dtype=np.float64), (n,)))
        mode = SEED_GRID if sd is None else (SEED_GIVEN if seed_ab else SEED_NONLINEAR)
        b = FitBatch(n, tmax, qn.shape[1], model, mode, nstarts, max_iter,
                     ntau.ctypes.data, T.ctypes.data, Dd.ctypes.data, w.ctypes.data, qn.ctypes.data, bd.ctypes.data,
                     None if sd is None else sd.ctypes.data, None if vm is None else vm.ctypes.data)
        th = np.zeros((n, 6))
        chi = np.zeros(n)
        it = np.zeros(n, dtype=np.int32)
        ms = C.c_double()
        rc = self._lib.ddm_fit(self._h, C.byref(b), th.ctypes.data, chi.ctypes.data, it.ctypes.data, C.byref(ms))
        if rc != 0:
            raise DdmError(self._lib.ddm_fitter_last_error(self._h).decode())
        return FitResult(th, chi, it, ms.value)


@dataclass
class GlobalFitResult:
    theta: np.ndarray     # (nvideos, 4): alpha, ln vbar, ln Z, ln D shared by the video's bins
    rss: np.ndarray       # the reference's objective sum_b w_b^2 rss_b
    iters: np.ndarray     # LM steps (all starts)
    brown: np.ndarray     # (nvideos, 2): D and RSS of the Brownian fit (plain seeding; NaN when seeded)
    amps: list            # per video: (nbins, 2) amplitudes A_b, B_b at theta
    kernel_ms: float

    @property
    def alpha(self):
        return self.theta[:, 0]

    @property
    def vbar(self):
        return np.exp(self.theta[:, 1])

    @property
    def Z(self):
        return np.exp(self.theta[:, 2])

    @property
    def D(self):
        return np.exp(self.theta[:, 3])


def fit_global(fitter: "Fitter", prs, model=SWIMMER_3D, seeds=None, nstarts=3, max_iter=200, vmag=None
               ) -> GlobalFitResult:
    """Global-parameter fits of prepared records (one per video): alpha, vbar, Z, D shared by a video's bins,
    A_b, B_b >= 0 per bin, the reference's objective (fitSwimmer / fitBrownian over evalModel,
    src/motility_fit.js:450-474, :500-611), all videos in one launch.  seeds: (nvideos, 4) = (alpha, ln vbar,
    ln Z, ln D) for LM from one start; None = the reference's grids.  vmag: drift speed per video."""
    nv = len(prs)
    if nv == 0:
        return GlobalFitResult(np.zeros((0, 4)), np.zeros(0), np.zeros(0, np.int32), np.zeros((0, 2)), [], 0.0)
    nbs = [len(p.bins) for p in prs]
    if min(nbs) < 1:
        raise DdmError("fit_global: every record needs at least one usable bin")
    off = np.concatenate([[0], np.cumsum(nbs)]).astype(np.int32)
    tmax = max(len(p.tau) for p in prs)
    if tmax > MAXT:
        raise DdmError(f"at most {MAXT} lags per fit")
    K = max(len(b.qn) for p in prs for b in p.bins)
    nb = int(off[-1])
    T = np.zeros((nv, tmax))
    Dd = np.zeros((nb, tmax))
    w = np.zeros(nb)
    qn = np.zeros((nb, K))
    bd = np.zeros((nv, 4))
    nt = np.array([len(p.tau) for p in prs], np.int32)
    for v, p in enumerate(prs):
        T[v, :nt[v]] = p.tau
        bd[v] = bounds(p)
        for i, b in enumerate(p.bins):
            Dd[off[v] + i, :nt[v]] = b.d
            w[off[v] + i] = b.w
            qn[off[v] + i] = np.resize(b.qn, K)
    sd = None if seeds is None else np.ascontiguousarray(seeds, dtype=np.float64).reshape(nv, 4)
    vm = None if vmag is None else np.ascontiguousarray(np.broadcast_to(np.asarray(vmag, np.float64), (nv,)))
    b = GlobalFitBatch(nv, tmax, K, model, SEED_GRID if sd is None else SEED_GIVEN, nstarts, max_iter,
                       off.ctypes.data, nt.ctypes.data, T.ctypes.data, Dd.ctypes.data, w.ctypes.data, qn.ctypes.data,
                       bd.ctypes.data, None if sd is None else sd.ctypes.data, None if vm is None else vm.ctypes.data)
    th = np.zeros((nv, 4))
    rss = np.zeros(nv)
    it = np.zeros(nv, np.int32)
    br = np.zeros((nv, 2))
    ab = np.zeros((nb, 2))
    ms = C.c_double()
    rc = fitter._lib.ddm_fit_global(fitter._h, C.byref(b), th.ctypes.data, rss.ctypes.data, it.ctypes.data,
                                    br.ctypes.data, ab.ctypes.data, C.byref(ms))
    if rc != 0:
        raise DdmError(fitter._lib.ddm_fitter_last_error(fitter._h).decode())
    return GlobalFitResult(th, rss, it, br, [ab[off[v]:off[v + 1]] for v in range(nv)], ms.value)


def problems_from(pr: Prepared):
    """Per-bin fit problems of one prepared Radial: (tau, data, weight, qnodes, bounds) lists."""
    bd = bounds(pr)
    K = max(len(b.qn) for b in pr.bins)
    tau = [pr.tau for _ in pr.bins]
    data = [b.d for b in pr.bins]
    w = np.array([b.w for b in pr.bins])
    qn = np.array([np.resize(b.qn, K) for b in pr.bins])
    return tau, data, w, qn, np.tile(bd, (len(pr.bins), 1))


def analyzeMotilityLM(radial: Radial, dxUm: float, fps: float, model=SWIMMER_3D, fitter: Fitter | None = None,
                      seeds=None, nstarts=3, max_iter=400, vmag=None):
    """Per-q LM fits of every usable bin of one Radial record -> (Prepared, FitResult).  vmag: a uniform
    drift speed (um/s, e.g. estimate_drift's mag) held fixed in every bin's model."""
    pr = prepare(radial, dxUm, fps)
    f = fitter or Fitter()
    res = f.fit(*problems_from(pr), model=model, seeds=seeds, nstarts=nstarts, max_iter=max_iter,
                vmag=None if not vmag else vmag)
    return pr, res
