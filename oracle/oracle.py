"""ctypes bindings of oracle/liboracle.so plus numpy helpers.

TEST INFRASTRUCTURE ONLY: the checker the HIP engine is compared against.
Every wrapped function restates a reference function; the C source names the
reference lines (motility-ddm/src/fft.js, src/ddm_ref.js, src/motility_fit.js).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def build() -> None:
    """Compile liboracle.so from oracle/ddm_oracle.c (gcc)."""
    subprocess.run(["make", "-s", "-B", "-C", _HERE, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_blackman_harris.argtypes = [C.c_int, _dp]
        L.orc_fft.argtypes = [C.c_int, _dp, C.c_int]
        L.orc_fft.restype = C.c_int
        L.orc_dft.argtypes = [C.c_int, _dp, _dp, C.c_int]
        L.orc_spectrum2d_half.argtypes = [_dp, C.c_int, _dp, _dp]
        L.orc_history_ingest_stack.argtypes = [_dp, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_double,
                                               C.c_int, C.c_int]
        L.orc_temporal_mode.argtypes = [_dp, C.c_long, C.c_int, _u32p, _fp, C.c_int, _dp, C.c_void_p]
        L.orc_temporal_mode.restype = C.c_double
        L.orc_temporal_mode_direct.argtypes = [_dp, C.c_long, C.c_int, _u32p, _fp, C.c_int, _dp, C.c_void_p]
        L.orc_temporal_mode_direct.restype = C.c_double
        L.orc_temporal_meta.argtypes = [C.c_int, C.c_int, _u32p, _fp, _fp, C.c_void_p]
        L.orc_v8_hypot.argtypes = [C.c_double, C.c_double]
        L.orc_v8_hypot.restype = C.c_double
        L.orc_make_bins.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int, _u32p, _i32p,
                                    C.c_long, _dp, _u32p, C.c_void_p]
        L.orc_make_bins.restype = C.c_long
        L.orc_analyze_radial.argtypes = [_dp, C.c_int, _u32p, _fp, _u32p, _i32p, C.c_int, C.c_int, _dp, _dp]
        L.orc_analyze_radial_split.argtypes = [_dp, C.c_int, _u32p, _fp, _u32p, _i32p, C.c_int, C.c_int, _dp, _dp]
        L.orc_radial_drift.argtypes = [_dp, C.c_int, _u32p, _fp, _u32p, _i32p, C.c_int, C.c_int, _dp]
        L.orc_drift_from_modes.argtypes = [_dp, _u32p, _i32p, C.c_int, C.c_int, _dp]
        L.orc_analyze_radial_direct.argtypes = [_dp, C.c_int, _u32p, _fp, _u32p, _i32p, C.c_int, C.c_int, _dp,
                                                _dp]
        for name in ("orc_bessel_j0", "orc_lngamma"):
            getattr(L, name).argtypes = [C.c_double]
            getattr(L, name).restype = C.c_double
        for name in ("orc_isf3d", "orc_isf2d", "orc_isf2d_fast", "orc_legendre_p"):
            getattr(L, name).argtypes = [C.c_double, C.c_double]
            getattr(L, name).restype = C.c_double
        L.orc_schulz_pdf.argtypes = [C.c_double, C.c_double, C.c_double]
        L.orc_schulz_pdf.restype = C.c_double
        L.orc_lag_grid.argtypes = [C.c_int, _i32p, C.c_int]
        L.orc_lag_grid.restype = C.c_int
        L.orc_nnls3.argtypes = [_dp, _dp, C.c_double, _dp, C.POINTER(C.c_double)]
        L.orc_model_column.argtypes = [C.c_int, _dp, C.c_int, _dp, C.c_double, C.c_double, C.c_double,
                                       C.c_double, C.c_double, C.c_int, C.c_int, _dp]
        _lib = L
    return _lib


# ------------------------------------------------------------------ spatial

def blackman_harris(n: int) -> np.ndarray:
    w = np.zeros(n)
    lib().orc_blackman_harris(n, w)
    return w


def fft(x: np.ndarray, inverse: bool = False) -> np.ndarray:
    """Complex FFT (unnormalised), radix-2 Stockham; x complex128 of pow2 length."""
    d = np.ascontiguousarray(np.asarray(x, dtype=np.complex128)).view(np.float64).copy()
    if lib().orc_fft(len(d) // 2, d, int(inverse)) != 0:
        raise ValueError("FFT length is not a power of two")
    return d.view(np.complex128)


def dft(x: np.ndarray, inverse: bool = False) -> np.ndarray:
    d = np.ascontiguousarray(np.asarray(x, dtype=np.complex128)).view(np.float64)
    out = np.zeros_like(d)
    lib().orc_dft(len(d) // 2, d, out, int(inverse))
    return out.view(np.complex128)


def spectrum2d_half(img: np.ndarray) -> np.ndarray:
    """(N, N) float image -> (N, N/2+1) complex half-plane spectrum (mode = ky*NH+kx)."""
    N = img.shape[0]
    out = np.zeros(2 * N * (N // 2 + 1))
    lib().orc_spectrum2d_half(np.ascontiguousarray(img, dtype=np.float64), N, blackman_harris(N), out)
    return out.view(np.complex128).reshape(N, N // 2 + 1)


_DTYPE_CODE = {np.dtype(np.uint8): 0, np.dtype(np.uint16): 1, np.dtype(np.float32): 2}
SCALE = {np.dtype(np.uint8): 1.0 / 255.0, np.dtype(np.uint16): 1.0 / 65535.0, np.dtype(np.float32): 1.0}


def history(frames: np.ndarray, T: int | None = None, first_slot: int = 0, scale: float | None = None) -> np.ndarray:
    """Ingest a frame stack (F, N, N) into a ring history, complex (modes, T)."""
    frames = np.ascontiguousarray(frames)
    F, N, _ = frames.shape
    T = F if T is None else T
    if scale is None:
        scale = SCALE[frames.dtype]
    modes = N * (N // 2 + 1)
    hist = np.zeros(modes * T * 2)
    lib().orc_history_ingest_stack(hist, N, T, frames.ctypes.data, _DTYPE_CODE[frames.dtype], scale, F, first_slot)
    return hist.view(np.complex128).reshape(modes, T)


# ------------------------------------------------------------------ temporal

def temporal_mode(hist: np.ndarray, mode: int, order, mask, L: int, direct: bool = False):
    """Returns (D[L], C[L] complex, P) for one mode; hist complex (modes, T)."""
    h = np.ascontiguousarray(hist, dtype=np.complex128).view(np.float64).ravel()
    T = hist.shape[1]
    D = np.zeros(L)
    Cc = np.zeros(2 * L)
    o = np.ascontiguousarray(order, dtype=np.uint32)
    m = np.ascontiguousarray(mask, dtype=np.float32)
    f = lib().orc_temporal_mode_direct if direct else lib().orc_temporal_mode
    P = f(h, mode, T, o, m, L, D, Cc.ctypes.data)
    return D, Cc.view(np.complex128), P


def temporal_meta(T: int, L: int, order, mask):
    npairs = np.zeros(L, dtype=np.float32)
    n = 1
    while n < 2 * T:
        n <<= 1
    mspec = np.zeros(2 * n, dtype=np.float32)
    lib().orc_temporal_meta(T, L, np.ascontiguousarray(order, dtype=np.uint32),
                            np.ascontiguousarray(mask, dtype=np.float32), npairs, mspec.ctypes.data)
    return npairs, mspec.view(np.complex64)


# ------------------------------------------------------------------ bins

def v8_hypot(a: float, b: float) -> float:
    return lib().orc_v8_hypot(a, b)


def make_bins(N: int, NB: int, kmin: float, kmax: float, hypot_mode: int = 0, knodes: int = 4,
              with_nodes: bool = True) -> dict:
    NH = N // 2 + 1
    cap = N * NH
    off = np.zeros(NB + 1, dtype=np.uint32)
    modes = np.zeros(4 * cap, dtype=np.int32)
    kc = np.zeros(NB)
    nm = np.zeros(NB, dtype=np.uint32)
    kn = np.zeros(NB * knodes)
    total = lib().orc_make_bins(N, NB, kmin, kmax, hypot_mode, knodes, off, modes, cap, kc, nm,
                                kn.ctypes.data if with_nodes else None)
    return dict(N=N, NB=NB, kmin=kmin, kmax=kmax, binOffsets=off, binModes=modes[:4 * total].reshape(total, 4),
                kCenter=kc, nModes=nm, kNodes=kn if with_nodes else None)


def analyze_radial(hist: np.ndarray, order, mask, bins: dict, L: int, direct: bool = False, split: bool = False):
    """Bin-averaged D[NB, L] and P[NB] (FP64); hist complex (modes, T).

    Default: the reference's algorithm (packed z = m + i p transform).  split:
    mask and p transformed separately (the FP64 truth at long series, where
    the packed form's p inherits ~ulp x the mask's energy).  direct: O(T L)
    sums."""
    h = np.ascontiguousarray(hist, dtype=np.complex128).view(np.float64).ravel()
    T = hist.shape[1]
    NB = bins["NB"]
    D = np.zeros(NB * L)
    P = np.zeros(NB)
    f = lib().orc_analyze_radial_direct if direct else (lib().orc_analyze_radial_split if split else
                                                         lib().orc_analyze_radial)
    f(h, T, np.ascontiguousarray(order, dtype=np.uint32), np.ascontiguousarray(mask, dtype=np.float32),
      bins["binOffsets"], np.ascontiguousarray(bins["binModes"].ravel()), NB, L, D, P)
    return D.reshape(NB, L), P


def radial_drift(hist: np.ndarray, order, mask, bins: dict, L: int) -> np.ndarray:
    """Phase-drift sums drift[NB, L, 8] (src/ddm_ref.js:360-381), FP64; hist complex (modes, T)."""
    h = np.ascontiguousarray(hist, dtype=np.complex128).view(np.float64).ravel()
    T = hist.shape[1]
    NB = bins["NB"]
    out = np.zeros(NB * L * 8)
    lib().orc_radial_drift(h, T, np.ascontiguousarray(order, dtype=np.uint32),
                           np.ascontiguousarray(mask, dtype=np.float32), bins["binOffsets"],
                           np.ascontiguousarray(bins["binModes"].ravel()), NB, L, out)
    return out.reshape(NB, L, 8)


def radial_from_full(D: np.ndarray, Cm: np.ndarray, P: np.ndarray, bins: dict, L: int) -> dict:
    """radialFromFull (src/ddm_ref.js:474-487, finalizeRadial :404-412) from per-mode arrays D[modes, L],
    complex C[modes, L], P[modes]: bin means of D and P, drift sums [NB, L, 8]."""
    NB = bins["NB"]
    off = bins["binOffsets"]
    bm = np.ascontiguousarray(bins["binModes"].ravel())
    Dr = np.zeros((NB, L))
    Pr = np.zeros(NB)
    for b in range(NB):
        ms = bins["binModes"][off[b]:off[b + 1], 0]
        if len(ms):
            Dr[b] = np.asarray(D)[ms].sum(axis=0) / len(ms)
            Pr[b] = np.asarray(P)[ms].sum() / len(ms)
    c = np.ascontiguousarray(Cm, dtype=np.complex128).view(np.float64).ravel()
    drift = np.zeros(NB * L * 8)
    lib().orc_drift_from_modes(c, off, bm, NB, L, drift)
    return dict(N=bins["N"], NB=NB, L=L, kCenter=bins["kCenter"], kNodes=bins["kNodes"], nModes=bins["nModes"],
                D=Dr.ravel(), P=Pr, drift=drift)


# ------------------------------------------------------------------ special functions / fit blocks

def bessel_j0(x): return lib().orc_bessel_j0(float(x))
def lngamma(x): return lib().orc_lngamma(float(x))
def schulz_pdf(v, vbar, Z): return lib().orc_schulz_pdf(float(v), float(vbar), float(Z))
def legendre_p(nu, c): return lib().orc_legendre_p(float(nu), float(c))
def isf3d(x, Z): return lib().orc_isf3d(float(x), float(Z))
def isf2d(x, Z): return lib().orc_isf2d(float(x), float(Z))
def isf2d_fast(x, Z): return lib().orc_isf2d_fast(float(x), float(Z))


def lag_grid(L: int) -> np.ndarray:
    out = np.zeros(max(L, 16), dtype=np.int32)
    n = lib().orc_lag_grid(L, out, len(out))
    return out[:n].copy()


def nnls3(G, h, dd):
    a = np.zeros(3)
    rss = C.c_double()
    lib().orc_nnls3(np.ascontiguousarray(G, dtype=np.float64).ravel(), np.ascontiguousarray(h, dtype=np.float64),
                    float(dd), a, C.byref(rss))
    return a, rss.value


def model_column(tau, qn, vbar, Z, Dd, Vmag=0.0, alpha=0.0, swim=True, geometry=0):
    tau = np.ascontiguousarray(tau, dtype=np.float64)
    qn = np.ascontiguousarray(qn, dtype=np.float64)
    out = np.zeros(len(tau))
    lib().orc_model_column(len(tau), tau, len(qn), qn, vbar, Z, Dd, Vmag, alpha, int(swim), geometry, out)
    return out


def linear_order(valid):
    """linearOrder; src/ddm_ref.js:493-502."""
    valid = np.asarray(valid, dtype=bool)
    return np.arange(len(valid), dtype=np.uint32), valid.astype(np.float32)


def mulberry32(seed: int):
    """Seeded uniform RNG; test/sim.mjs:14-23."""
    a = seed & 0xFFFFFFFF

    def imul(x, y):
        return (x * y) & 0xFFFFFFFF

    def u():
        nonlocal a
        a = (a + 0x6D2B79F5) & 0xFFFFFFFF
        t = a
        t = imul(t ^ (t >> 15), t | 1)
        t ^= (t + imul(t ^ (t >> 7), t | 61)) & 0xFFFFFFFF
        return ((t ^ (t >> 14)) & 0xFFFFFFFF) / 4294967296.0
    return u


def shuffle_order(order, mask, rand):
    """shuffleOrder (seeded Fisher-Yates of (slot, mask) pairs); src/ddm_ref.js:510-519."""
    order = np.array(order, dtype=np.uint32)
    mask = np.array(mask, dtype=np.float32)
    for i in range(len(order) - 1, 0, -1):
        j = int(np.floor(rand() * (i + 1)))
        order[i], order[j] = order[j], order[i]
        mask[i], mask[j] = mask[j], mask[i]
    return order, mask


# ------------------------------------------------------------------ fits

def _fit_decls():
    L = lib()
    if getattr(L, "_fit_declared", False):
        return L
    L.orc_eval_model.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp, C.c_double, C.c_double, C.c_double,
                                 C.c_double, C.c_double, C.c_int, C.c_int, C.c_void_p]
    L.orc_eval_model.restype = C.c_double
    L.orc_bounds.argtypes = [C.c_int, _dp, C.c_double, C.c_double, _dp]
    L.orc_fit_brownian.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, C.c_double,
                                   C.POINTER(C.c_double)]
    L.orc_fit_brownian.restype = C.c_double
    L.orc_fit_swimmer.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, C.c_int, C.c_double, C.c_double,
                                  _dp]
    L.orc_fit_swimmer.restype = C.c_double
    L.orc_lm_seed.argtypes = [C.c_int, _dp, _dp, C.c_double, C.c_int, _dp, _dp, C.c_int, C.c_int, C.c_double, _dp]
    L.orc_lm_fit.argtypes = [C.c_int, _dp, _dp, C.c_double, C.c_int, _dp, _dp, C.c_int, C.c_int, C.c_double, C.c_int,
                             _dp, C.POINTER(C.c_int)]
    L.orc_lm_fit.restype = C.c_double
    L._fit_declared = True
    return L


def prepare(radial: dict, dxUm: float, fps: float) -> dict:
    """prepare (src/motility_fit.js:383-401): bins with nModes >= 3 and max D > 0,
    q, q-nodes from kNodes, data on the lag grid, weight 1/max d."""
    N, NB, L = radial["N"], radial["NB"], radial["L"]
    dt = 1.0 / fps
    lags = lag_grid(L)
    tau = lags * dt
    D = np.asarray(radial["D"]).reshape(NB, L)
    qs = 2 * np.pi / (N * dxUm)
    kn = radial.get("kNodes")
    K = len(kn) // NB if kn is not None and len(kn) else 0
    out = []
    for b in range(NB):
        if radial["nModes"][b] < 3:
            continue
        d = D[b, lags]
        scale = d.max()
        if not (scale > 0) or not np.isfinite(scale):
            continue
        q = qs * radial["kCenter"][b]
        qn = np.asarray(kn[b * K:(b + 1) * K]) * qs if K else np.array([q])
        out.append(dict(b=b, q=q, qn=qn, d=d.astype(np.float64), w=1.0 / scale))
    return dict(bins=out, lags=lags, tau=tau, dt=dt, L=L, tauMax=(L - 1) * dt)


def _stack(pr):
    bins = pr["bins"]
    K = len(bins[0]["qn"])
    return (len(bins), len(pr["tau"]), K, np.ascontiguousarray(pr["tau"], dtype=np.float64),
            np.ascontiguousarray(np.concatenate([b["d"] for b in bins])),
            np.ascontiguousarray([b["w"] for b in bins], dtype=np.float64),
            np.ascontiguousarray(np.concatenate([b["qn"] for b in bins]), dtype=np.float64))


def bounds(pr) -> np.ndarray:
    q = np.array([b["q"] for b in pr["bins"]], dtype=np.float64)
    out = np.zeros(6)
    _fit_decls().orc_bounds(len(q), q, pr["dt"], pr["tauMax"], out)
    return out


def fit_brownian(pr, Vmag=0.0):
    """fitBrownian (src/motility_fit.js:500-536) -> (Dd, rss)."""
    nb, nt, K, tau, d, w, qn = _stack(pr)
    rss = C.c_double()
    Dd = _fit_decls().orc_fit_brownian(nb, nt, K, tau, d, w, qn, bounds(pr), Vmag, C.byref(rss))
    return Dd, rss.value


def fit_swimmer(pr, geometry="3d", Vmag=0.0, Dstart=0.0):
    """fitSwimmer (src/motility_fit.js:577-611) -> dict(vbar, Z, Dd, alpha, rss); 2d uses the table kernel."""
    nb, nt, K, tau, d, w, qn = _stack(pr)
    out = np.zeros(4)
    geo = 0 if geometry == "3d" else 2
    rss = _fit_decls().orc_fit_swimmer(nb, nt, K, tau, d, w, qn, bounds(pr), geo, Vmag, Dstart, out)
    return dict(vbar=out[0], Z=out[1], Dd=out[2], alpha=out[3], rss=rss)


GEOMETRY = {"3d": 0, "2d-exact": 1, "2d": 2}


def lm_seed(tau, d, w, qn, bounds4, swim=True, geometry="3d", vmag=0.0):
    th = np.zeros(6)
    _fit_decls().orc_lm_seed(len(tau), np.ascontiguousarray(tau, dtype=np.float64),
                             np.ascontiguousarray(d, dtype=np.float64), w, len(qn),
                             np.ascontiguousarray(qn, dtype=np.float64), np.ascontiguousarray(bounds4, dtype=np.float64),
                             int(swim), GEOMETRY[geometry], float(vmag), th)
    return th


def lm_fit(tau, d, w, qn, bounds4, seed, swim=True, geometry="3d", maxit=200, vmag=0.0):
    """Per-q LM (FP64) from theta = (A, B, alpha, ln vbar, ln Z, ln D) -> (theta, chi2, iterations).
    vmag: a fixed uniform drift speed (E = exp(-D q^2 tau) J0(q vmag tau))."""
    th = np.ascontiguousarray(seed, dtype=np.float64).copy()
    it = C.c_int()
    chi = _fit_decls().orc_lm_fit(len(tau), np.ascontiguousarray(tau, dtype=np.float64),
                                  np.ascontiguousarray(d, dtype=np.float64), w, len(qn),
                                  np.ascontiguousarray(qn, dtype=np.float64),
                                  np.ascontiguousarray(bounds4, dtype=np.float64), int(swim), GEOMETRY[geometry],
                                  float(vmag), maxit, th, C.byref(it))
    return th, chi, it.value


def lm_plain(tau, d, w, qn, bounds4, swim=True, geometry="3d", nstarts=3, maxit=100, vmag=0.0):
    """Grid-seeded per-q LM from the best `nstarts` grid starts -> (theta, chi2, total iterations)."""
    L = _fit_decls()
    if not hasattr(L, "_plain"):
        L.orc_lm_plain.argtypes = [C.c_int, _dp, _dp, C.c_double, C.c_int, _dp, _dp, C.c_int, C.c_int, C.c_double,
                                   C.c_int, C.c_int, _dp, C.POINTER(C.c_int)]
        L.orc_lm_plain.restype = C.c_double
        L._plain = True
    th = np.zeros(6)
    it = C.c_int()
    chi = L.orc_lm_plain(len(tau), np.ascontiguousarray(tau, dtype=np.float64), np.ascontiguousarray(d, dtype=np.float64),
                         w, len(qn), np.ascontiguousarray(qn, dtype=np.float64),
                         np.ascontiguousarray(bounds4, dtype=np.float64), int(swim), GEOMETRY[geometry], float(vmag),
                         nstarts, maxit, th, C.byref(it))
    return th, chi, it.value


def lm_starts(tau, d, w, qn, bounds4, swim=True, geometry="3d", vmag=0.0):
    """orc_lm_plain's start grid ranked by grid chi2 -> (starts (n, 6), chi2 (n,))."""
    L = _fit_decls()
    if not hasattr(L, "_starts"):
        L.orc_lm_starts.argtypes = [C.c_int, _dp, _dp, C.c_double, C.c_int, _dp, _dp, C.c_int, C.c_int,
                                    C.c_double, _dp, _dp]
        L.orc_lm_starts.restype = C.c_int
        L._starts = True
    st = np.zeros(64 * 6)
    sf = np.zeros(64)
    n = L.orc_lm_starts(len(tau), np.ascontiguousarray(tau, dtype=np.float64), np.ascontiguousarray(d, dtype=np.float64),
                        w, len(qn), np.ascontiguousarray(qn, dtype=np.float64),
                        np.ascontiguousarray(bounds4, dtype=np.float64), int(swim), GEOMETRY[geometry], float(vmag),
                        st, sf)
    return st[:6 * n].reshape(n, 6), sf[:n]


# ------------------------------------------------------------------ scenario simulator

class SimParams(C.Structure):
    _fields_ = [(n, C.c_int) for n in ("N", "T", "seed", "n_swim", "swim_2d", "n_passive", "n_fixed", "drop_exact")] + \
               [(n, C.c_double) for n in ("fps", "dx_um", "vbar", "sigma_v", "D_um2s", "drift_speed", "drift_angle_deg",
                                          "drop_frac", "blob_sigma", "amp", "background", "noise")]


def simulate(sim: dict):
    """simulate (test/sim.mjs:80-205) -> (frames u8 [T, N, N], valid bool [T], mean swimmer speed).
    sim: the reference's SimParams (N, T, fps, dxUm, seed, swimmers {n, vbar, sigma, geometry}, passive {n},
    Dum2s, fixed {n}, drift {speed, angleDeg}, dropFrac, dropExact, blobSigmaPx, amp, background, noise)."""
    sw = sim.get("swimmers") or {}
    p = SimParams(N=sim.get("N", 256), T=sim.get("T", 256), seed=sim.get("seed", 1), n_swim=sw.get("n", 0),
                  swim_2d=1 if sw.get("geometry") == "2d" else 0, n_passive=(sim.get("passive") or {}).get("n", 0),
                  n_fixed=(sim.get("fixed") or {}).get("n", 0), drop_exact=sim.get("dropExact", 0) or 0,
                  fps=sim.get("fps", 30), dx_um=sim.get("dxUm", 1), vbar=sw.get("vbar", 0), sigma_v=sw.get("sigma", 1),
                  D_um2s=sim.get("Dum2s", 0), drift_speed=(sim.get("drift") or {}).get("speed", 0),
                  drift_angle_deg=(sim.get("drift") or {}).get("angleDeg", 0), drop_frac=sim.get("dropFrac", 0) or 0,
                  blob_sigma=sim.get("blobSigmaPx", 1.5), amp=sim.get("amp", 0.15),
                  background=sim.get("background", 0.25), noise=sim.get("noise", 0.01))
    L = lib()
    L.orc_simulate.argtypes = [C.POINTER(SimParams), C.c_void_p, C.c_void_p]
    L.orc_simulate.restype = C.c_double
    frames = np.zeros((p.T, p.N, p.N), dtype=np.uint8)
    valid = np.zeros(p.T, dtype=np.uint8)
    ms = L.orc_simulate(C.byref(p), frames.ctypes.data, valid.ctypes.data)
    return frames, valid.astype(bool), ms


# ------------------------------------------------------------------ Node CPU path (baseline)

def cpu_path(frames: np.ndarray, bins: dict, L: int, threads: int = 1, timeout: float = 600.0) -> dict:
    """Run oracle/cpu_path.mjs (the reference's CPU path restated in JavaScript) under Node on `frames`
    (T x N x N, uint8 or uint16; frame t into slot t, every frame valid) with `threads` threads.  Returns
    its timings (JSON) plus D[NB, L], P[NB] and drift[NB, L, 8]."""
    import json
    import tempfile
    T, N = frames.shape[0], frames.shape[1]
    NB = bins["NB"]
    dt = {np.dtype(np.uint8): "u8", np.dtype(np.uint16): "u16"}[frames.dtype]
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, "spec.json"), "w") as f:
            json.dump(dict(N=N, T=T, L=L, NB=NB, dtype=dt, threads=int(threads)), f)
        np.ascontiguousarray(frames).tofile(os.path.join(d, "frames.bin"))
        np.ascontiguousarray(bins["binOffsets"], np.uint32).tofile(os.path.join(d, "binOffsets.bin"))
        np.ascontiguousarray(bins["binModes"], np.int32).ravel().tofile(os.path.join(d, "binModes.bin"))
        p = subprocess.run(["node", os.path.join(_HERE, "cpu_path.mjs"), d], check=True, capture_output=True,
                           text=True, timeout=timeout)
        out = json.loads(p.stdout.strip().splitlines()[-1])
        out["D"] = np.fromfile(os.path.join(d, "D.bin")).reshape(NB, L)
        out["P"] = np.fromfile(os.path.join(d, "P.bin"))
        out["drift"] = np.fromfile(os.path.join(d, "drift.bin")).reshape(NB, L, 8)
    return out
