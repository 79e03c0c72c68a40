"""Python host mirror of the reference's GPU engine interface over libddm.so.

`DdmEngine` keeps the method names, argument meaning and error behaviour of
the reference's `DdmEngine` (motility-ddm src/ddm-engine.ts:62-426): `create`,
`ingestLuma`, `computeTemporal`, `readRadial`, `readHistory`, `readMean`,
`readTemporalFull`, `idle`, `destroy`, `bins`; errors raise `DdmError` with the
engine's message where the reference throws `Error`.  `ingestFrames` is the
batched form of `ingestLuma` (a stack of u8/u16/f32 frames in one call), and
`ingestDevice` takes a stack already resident in GPU memory.

All compute runs in libddm.so (HIP, gfx950).  If the library is missing this
module raises on import of the engine -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libddm.so")

DDM_OK, DDM_DROPPED = 0, 1
DDM_U8, DDM_U16, DDM_F32 = 0, 1, 2
K_ROWS, K_COLS, K_TEMPORAL, K_FINALIZE, K_COUNT = 0, 1, 2, 3, 4
KERNEL_NAMES = ("spatial_rows", "spatial_cols_cornerturn", "temporal_wk", "radial_finalize")

# Exported symbols of include/ddm.h (checked by tests/test_abi.py).
EXPORTS = ("ddm_create", "ddm_destroy", "ddm_last_error", "ddm_get_sizes", "ddm_get_bins", "ddm_ingest",
           "ddm_ingest_device", "ddm_compute_temporal", "ddm_read_radial", "ddm_read_history", "ddm_idle",
           "ddm_history_device", "ddm_set_profiling", "ddm_kernel_times", "ddm_fitter_create", "ddm_fitter_destroy",
           "ddm_fitter_last_error", "ddm_fit", "ddm_mlp_create", "ddm_mlp_destroy", "ddm_mlp_last_error",
           "ddm_mlp_set_params", "ddm_mlp_get_params", "ddm_mlp_forward", "ddm_mlp_loss_grad", "ddm_mlp_train_step",
           "ddm_synth_create", "ddm_synth_destroy", "ddm_synth_last_error", "ddm_synth_video", "ddm_synth_tracks",
           "ddm_create_tuned", "ddm_get_tuning", "ddm_batch_create", "ddm_batch_run", "ddm_batch_engine",
           "ddm_batch_last_error", "ddm_batch_destroy", "ddm_read_temporal_full", "ddm_analyze_motility",
           "ddm_verdict_last_error", "ddm_analyze_motility_batch", "ddm_read_mean", "ddm_fit_device_noalloc",
           "ddm_clock_create", "ddm_clock_destroy", "ddm_clock_place", "ddm_clock_drop", "ddm_clock_reset",
           "ddm_clock_state_get", "ddm_clock_snapshot", "ddm_shuffle_order", "ddm_lag_grid", "ddm_usable_bins",
           "ddm_swimmer_isf", "ddm_estimate_drift", "ddm_fit_global", "ddm_isf2d_eval")


class DdmError(RuntimeError):
    """Raised where the reference engine throws Error(message)."""


class Geometry(C.Structure):
    _fields_ = [("N", C.c_int32), ("T", C.c_int32), ("L", C.c_int32), ("NB", C.c_int32),
                ("kmin", C.c_double), ("kmax", C.c_double), ("hypot_mode", C.c_int32), ("knodes", C.c_int32)]


class Tuning(C.Structure):
    """Launch-shape tuning (0 = default): see ddm_tuning in include/ddm.h."""
    _fields_ = [("temporal_items", C.c_int32), ("chunk_mb", C.c_int32), ("wave_spatial", C.c_int32),
                ("cols_f", C.c_int32), ("fft_radix", C.c_int32),
                ("temporal_kernel", C.c_int32), ("temporal_wg_per_cu", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Sizes(C.Structure):
    _fields_ = [("N", C.c_int32), ("NH", C.c_int32), ("T", C.c_int32), ("TPAD", C.c_int32), ("L", C.c_int32),
                ("NB", C.c_int32), ("knodes", C.c_int32), ("modes", C.c_int64), ("binned_modes", C.c_int64),
                ("hist_bytes", C.c_int64), ("work_items", C.c_int32), ("chunk_frames", C.c_int32),
                ("hist_stride", C.c_int32)]


_lib = None


def load_library(path: str = LIB_PATH):
    """Load libddm.so and declare its C ABI; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise DdmError(f"libddm.so is not built ({path}); run python -m hackathon_amd.build")
    # torch's HIP runtime first: torch loads its own libamdhip64 by file name, so a copy of the runtime
    # loaded earlier through libddm's dependency would leave two runtimes in the process and torch without
    # a device; loaded after torch's, libddm's dependency resolves to torch's copy (same soname)
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(path)
    vp = C.c_void_p
    L.ddm_create.argtypes = [C.POINTER(Geometry), C.POINTER(vp)]
    L.ddm_destroy.argtypes = [vp]
    L.ddm_destroy.restype = None
    L.ddm_last_error.argtypes = [vp]
    L.ddm_last_error.restype = C.c_char_p
    L.ddm_get_sizes.argtypes = [vp, C.POINTER(Sizes)]
    L.ddm_get_bins.argtypes = [vp, vp, vp, vp, vp, vp]
    L.ddm_ingest.argtypes = [vp, vp, C.c_int32, C.c_int32, C.c_int32]
    L.ddm_ingest_device.argtypes = [vp, vp, C.c_int32, C.c_int32, C.c_int32]
    L.ddm_compute_temporal.argtypes = [vp, vp, vp]
    L.ddm_read_radial.argtypes = [vp, vp, vp, vp]
    L.ddm_read_history.argtypes = [vp, vp]
    L.ddm_read_temporal_full.argtypes = [vp, vp, vp, vp]
    L.ddm_idle.argtypes = [vp]
    L.ddm_history_device.argtypes = [vp, C.POINTER(vp)]
    L.ddm_set_profiling.argtypes = [vp, C.c_int32]
    L.ddm_kernel_times.argtypes = [vp, vp, vp, C.c_int32]
    L.ddm_create_tuned.argtypes = [C.POINTER(Geometry), C.POINTER(Tuning), C.POINTER(vp)]
    L.ddm_get_tuning.argtypes = [vp, C.POINTER(Tuning)]
    for name in EXPORTS:
        if name not in ("ddm_destroy", "ddm_last_error", "ddm_fitter_destroy", "ddm_fitter_last_error",
                        "ddm_mlp_destroy", "ddm_mlp_last_error", "ddm_synth_destroy", "ddm_synth_last_error",
                        "ddm_batch_last_error", "ddm_batch_destroy", "ddm_read_temporal_full",
                        "ddm_verdict_last_error", "ddm_clock_destroy"):
            getattr(L, name).restype = C.c_int
    _lib = L
    return L


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


@dataclass
class Radial:
    """The reference's Radial record (src/ddm_ref.js:336-348), field for field."""
    N: int
    NB: int
    L: int
    kCenter: np.ndarray
    kNodes: np.ndarray
    nModes: np.ndarray
    D: np.ndarray
    P: np.ndarray
    drift: np.ndarray = field(repr=False)


@dataclass
class Bins:
    """makeBins output (src/ddm_ref.js:84-170)."""
    N: int
    NB: int
    kmin: float
    kmax: float
    binOffsets: np.ndarray
    binModes: np.ndarray
    kCenter: np.ndarray
    nModes: np.ndarray
    kNodes: np.ndarray


_DTYPES = {np.dtype(np.uint8): DDM_U8, np.dtype(np.uint16): DDM_U16, np.dtype(np.float32): DDM_F32}


class DdmEngine:
    """One geometry, one HIP stream, engine-owned device buffers."""

    def __init__(self, handle, lib, owned: bool = True):
        self._h = handle
        self._lib = lib
        self._owned = owned
        s = Sizes()
        self._check(lib.ddm_get_sizes(self._h, C.byref(s)))
        self.sizes = s
        self.N, self.T, self.L, self.NB = s.N, s.T, s.L, s.NB
        self.NH = s.NH
        self.modes = s.modes
        self.TPAD = s.TPAD
        self.bins = self._read_bins()

    # -- lifetime ---------------------------------------------------------
    @classmethod
    def create(cls, N: int = 256, T: int = 256, L: int = 128, NB: int = 32, kmin: float = 4.0,
               kmax: float = 100.0, hypot_mode: int = 0, knodes: int = 4, tuning: dict | None = None) -> "DdmEngine":
        """DdmEngine.create (src/ddm-engine.ts:125-149); defaults = the reference DDM geometry.
        tuning: optional launch-shape choices (ddm_tuning fields), e.g. from hackathon_amd.autotune."""
        lib = load_library()
        g = Geometry(N, T, L, NB, float(kmin), float(kmax), hypot_mode, knodes)
        h = C.c_void_p()
        tn = Tuning(**(tuning or {}))
        rc = lib.ddm_create_tuned(C.byref(g), C.byref(tn), C.byref(h))
        if rc != DDM_OK:
            raise DdmError(lib.ddm_last_error(None).decode())
        return cls(h, lib)

    def destroy(self) -> None:
        if self._h:
            if self._owned:
                self._lib.ddm_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.destroy()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    def _check(self, rc: int) -> bool:
        if rc == DDM_OK:
            return True
        if rc == DDM_DROPPED:
            return False
        raise DdmError(self._lib.ddm_last_error(self._h).decode())

    def _alive(self):
        if not self._h:
            raise DdmError("engine destroyed")

    def tuning(self) -> dict:
        t = Tuning()
        self._check(self._lib.ddm_get_tuning(self._h, C.byref(t)))
        return t.as_dict()

    # -- tables -----------------------------------------------------------
    def _read_bins(self) -> Bins:
        NB, K = self.NB, self.sizes.knodes
        off = np.zeros(NB + 1, dtype=np.uint32)
        modes = np.zeros(max(1, 4 * self.sizes.binned_modes), dtype=np.int32)
        kc = np.zeros(NB)
        nm = np.zeros(NB, dtype=np.uint32)
        kn = np.zeros(max(1, NB * K))
        self._check(self._lib.ddm_get_bins(self._h, _ptr(off), _ptr(modes), _ptr(kc), _ptr(nm), _ptr(kn)))
        return Bins(self.N, NB, 0.0, 0.0, off, modes[:4 * self.sizes.binned_modes].reshape(-1, 4), kc, nm,
                    kn[:NB * K])

    # -- ingest -----------------------------------------------------------
    def ingestLuma(self, luma: np.ndarray, slot: int) -> bool:
        """ingestLuma (src/ddm-engine.ts:308-324): one N*N float luma frame into slot % T."""
        self._alive()
        luma = np.ascontiguousarray(luma, dtype=np.float32)
        if luma.size != self.N * self.N:
            raise DdmError(f"ingestLuma expects {self.N * self.N} values")
        return self._check(self._lib.ddm_ingest(self._h, _ptr(luma), DDM_F32, 1, int(slot) % self.T))

    def ingestFrames(self, frames: np.ndarray, firstSlot: int = 0) -> bool:
        """Frame stack (F, N, N) of u8/u16/f32 from host memory into slots firstSlot + i (mod T)."""
        self._alive()
        frames = np.ascontiguousarray(frames)
        if frames.ndim != 3 or frames.shape[1:] != (self.N, self.N):
            raise DdmError(f"ingestFrames expects a (F, {self.N}, {self.N}) stack")
        if frames.dtype not in _DTYPES:
            raise DdmError(f"unsupported pixel type {frames.dtype}")
        return self._check(self._lib.ddm_ingest(self._h, _ptr(frames), _DTYPES[frames.dtype], frames.shape[0],
                                                int(firstSlot)))

    def ingestDevice(self, dptr: int, dtype, nframes: int, firstSlot: int = 0) -> bool:
        """Frame stack already in GPU memory (device pointer, e.g. tensor.data_ptr())."""
        self._alive()
        code = _DTYPES[np.dtype(dtype)]
        return self._check(self._lib.ddm_ingest_device(self._h, C.c_void_p(dptr), code, int(nframes),
                                                       int(firstSlot)))

    # -- analysis ---------------------------------------------------------
    def computeTemporal(self, order, mask) -> None:
        """computeTemporal (src/ddm-engine.ts:331-365)."""
        self._alive()
        order = np.ascontiguousarray(order, dtype=np.uint32)
        mask = np.ascontiguousarray(mask, dtype=np.float32)
        if order.size != self.T or mask.size != self.T:
            raise DdmError(f"temporalMeta: order/mask must have length {self.T}")
        self._check(self._lib.ddm_compute_temporal(self._h, _ptr(order), _ptr(mask)))

    def readRadial(self, drift: bool = False) -> Radial:
        """readRadial (src/ddm-engine.ts:391-395): the Radial record, FP64 arrays of the f32 results.
        drift=True also computes the phase-drift sums (a per-mode pass over the binned modes);
        otherwise the record's drift array is zeros."""
        self._alive()
        D = np.zeros(self.NB * self.L, dtype=np.float32)
        P = np.zeros(self.NB, dtype=np.float32)
        dr = np.zeros(self.NB * self.L * 8, dtype=np.float32)
        self._check(self._lib.ddm_read_radial(self._h, _ptr(D), _ptr(P), _ptr(dr) if drift else None))
        drift = dr
        return Radial(self.N, self.NB, self.L, self.bins.kCenter.copy(), self.bins.kNodes.copy(),
                      self.bins.nModes.copy(), D.astype(np.float64), P.astype(np.float64), drift.astype(np.float64))

    def historyDevice(self) -> int:
        """Device address of the spectral history (modes rows of sizes.hist_stride complex64, mode-major;
        slots 0..T-1 of a row hold the mode's series), for callers that keep the data on the GPU
        (ddm_history_device)."""
        self._alive()
        p = C.c_void_p()
        self._check(self._lib.ddm_history_device(self._h, C.byref(p)))
        return p.value

    def readMean(self) -> float:
        """readMean (src/ddm-engine.ts:410-414): mean luma of the last frame ingested (u8/255, u16/65535)."""
        self._alive()
        m = C.c_double()
        self._check(self._lib.ddm_read_mean(self._h, C.byref(m)))
        return m.value

    def readHistory(self) -> np.ndarray:
        """readHistory (src/ddm-engine.ts:397-400): complex64 (modes, T)."""
        self._alive()
        out = np.zeros(self.modes * self.T * 2, dtype=np.float32)
        self._check(self._lib.ddm_read_history(self._h, _ptr(out)))
        return out.view(np.complex64).reshape(self.modes, self.T)

    def readTemporalFull(self):
        """readTemporalFull (src/ddm-engine.ts:403-408): per-mode D (modes, L), C complex (modes, L), P (modes,)."""
        self._alive()
        D = np.zeros(self.modes * self.L, dtype=np.float32)
        Cc = np.zeros(self.modes * self.L * 2, dtype=np.float32)
        P = np.zeros(self.modes, dtype=np.float32)
        self._check(self._lib.ddm_read_temporal_full(self._h, _ptr(D), _ptr(Cc), _ptr(P)))
        return D.reshape(self.modes, self.L), Cc.view(np.complex64).reshape(self.modes, self.L), P

    def idle(self) -> None:
        self._alive()
        self._check(self._lib.ddm_idle(self._h))

    # -- profiling --------------------------------------------------------
    def setProfiling(self, enable: bool) -> None:
        self._check(self._lib.ddm_set_profiling(self._h, 1 if enable else 0))

    def kernelTimes(self) -> dict:
        ms = np.zeros(K_COUNT)
        cnt = np.zeros(K_COUNT, dtype=np.int32)
        self._check(self._lib.ddm_kernel_times(self._h, _ptr(ms), _ptr(cnt), K_COUNT))
        return {KERNEL_NAMES[k]: (float(ms[k]), int(cnt[k])) for k in range(K_COUNT)}
