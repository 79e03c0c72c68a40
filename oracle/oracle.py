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


def analyze_radial(hist: np.ndarray, order, mask, bins: dict, L: int, direct: bool = False):
    """Bin-averaged D[NB, L] and P[NB] (FP64); hist complex (modes, T)."""
    h = np.ascontiguousarray(hist, dtype=np.complex128).view(np.float64).ravel()
    T = hist.shape[1]
    NB = bins["NB"]
    D = np.zeros(NB * L)
    P = np.zeros(NB)
    f = lib().orc_analyze_radial_direct if direct else lib().orc_analyze_radial
    f(h, T, np.ascontiguousarray(order, dtype=np.uint32), np.ascontiguousarray(mask, dtype=np.float32),
      bins["binOffsets"], np.ascontiguousarray(bins["binModes"].ravel()), NB, L, D, P)
    return D.reshape(NB, L), P


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
