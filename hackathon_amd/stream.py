"""Streaming ring mode: the frame clock that maps live frames to history slots.

`FrameClock` is the Python face of the library's frame clock (`ddm_clock_*`,
the producer side of the reference's FrameClock, motility-ddm
src/camera.ts:306-425): frame n = round((t - t0) / dt) lives in slot n mod T
of the engine's history ring (`DdmEngine.ingestLuma(luma, slot)`); each slot
remembers which frame it holds, so a snapshot's mask leaves out dropped,
late, duplicated and overwritten frames by construction.  `snapshot` returns
the time-ordered (order, mask) pair for `computeTemporal`, oldest first, with
the oldest `guard` time indices masked (slots a producer may overwrite while
the analysis runs), optionally shuffled (the frame-order null test).  The
semantics are spelled out with `ddm_clock_create` in include/ddm.h; the Node
facade binds the same code.

T is the engine's ring length (the reference fixes it at 256).  Time stamps are
milliseconds; sources are 'captureTime', 'watchdog' (one timebase) and
'mediaTime' (a separate one).  On this engine an ingest issued after
computeTemporal is ordered behind it on the engine's stream, so a snapshot's
slots cannot change under a running analysis; the guard keeps the reference's
policy for producers that race it.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from .engine import DdmError, load_library

SOURCES = ("captureTime", "watchdog", "mediaTime")   # DDM_TIME_CAPTURE, _WATCHDOG, _MEDIA


class ClockState(C.Structure):
    _fields_ = [("newest_n", C.c_int64), ("frames", C.c_int64), ("missed", C.c_int64), ("duplicates", C.c_int64),
                ("resets", C.c_int64), ("valid", C.c_int32), ("time_source", C.c_int32),
                ("median_dt_ms", C.c_double)]


def _lib():
    L = load_library()
    if not getattr(L, "_clock_declared", False):
        vp = C.c_void_p
        L.ddm_clock_create.argtypes = [C.c_double, C.c_int32, C.POINTER(vp)]
        L.ddm_clock_destroy.argtypes = [vp]
        L.ddm_clock_destroy.restype = None
        L.ddm_clock_place.argtypes = [vp, C.c_double, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int64)]
        L.ddm_clock_drop.argtypes = [vp, C.c_int64]
        L.ddm_clock_reset.argtypes = [vp]
        L.ddm_clock_state_get.argtypes = [vp, C.POINTER(ClockState)]
        L.ddm_clock_snapshot.argtypes = [vp, C.c_int32, vp, vp, C.POINTER(C.c_int32)]
        L.ddm_shuffle_order.argtypes = [vp, vp, C.c_int32, vp]
        for f in ("ddm_clock_create", "ddm_clock_place", "ddm_clock_drop", "ddm_clock_reset", "ddm_clock_state_get",
                  "ddm_clock_snapshot", "ddm_shuffle_order"):
            getattr(L, f).restype = C.c_int
        L._clock_declared = True
    return L


@dataclass
class Snapshot:
    order: np.ndarray   # uint32 [T]: slot of time index j
    mask: np.ndarray    # float32 [T]: 1 where frame newestN - T + 1 + j is in its slot
    used: int           # valid frames in the snapshot
    newestN: int        # newest frame index at snapshot time


def shuffle_order(order, mask, rand):
    """shuffleOrder (src/ddm_ref.js:510-519): the (slot, mask) pairs permuted by the library's Fisher-Yates,
    fed with rand()'s draws (uniform in [0, 1)) in the order the reference consumes them."""
    order = np.array(order, dtype=np.uint32)
    mask = np.array(mask, dtype=np.float32)
    n = order.size
    if mask.size != n:
        raise DdmError("shuffleOrder: order and mask lengths differ")
    u = np.array([rand() for _ in range(max(n - 1, 0))], dtype=np.float64)
    if _lib().ddm_shuffle_order(order.ctypes.data, mask.ctypes.data, n, u.ctypes.data if u.size else None) != 0:
        raise DdmError("shuffleOrder: draws must lie in [0, 1)")
    return order, mask


class FrameClock:
    def __init__(self, fps: float, T: int = 256):
        self._lib = _lib()
        h = C.c_void_p()
        if self._lib.ddm_clock_create(float(fps), int(T), C.byref(h)) != 0:
            raise ValueError("FrameClock: fps must be positive and T >= 1")
        self._h = h
        self.T = int(T)
        self.dtMs = 1000.0 / fps

    def __del__(self):
        h, self._h = getattr(self, "_h", None), None
        if h:
            self._lib.ddm_clock_destroy(h)

    def _state(self) -> ClockState:
        s = ClockState()
        self._lib.ddm_clock_state_get(self._h, C.byref(s))
        return s

    frames = property(lambda self: int(self._state().frames))
    missed = property(lambda self: int(self._state().missed))
    duplicates = property(lambda self: int(self._state().duplicates))
    resets = property(lambda self: int(self._state().resets))
    newestN = property(lambda self: int(self._state().newest_n))
    validCount = property(lambda self: int(self._state().valid))
    medianDtMs = property(lambda self: float(self._state().median_dt_ms))

    @property
    def timeSource(self):
        s = self._state().time_source
        return SOURCES[s] if s >= 0 else None

    def reset(self) -> None:
        self._lib.ddm_clock_reset(self._h)

    def place(self, tMs: float, source: str = "captureTime"):
        """Slot for a frame captured at tMs: (slot, n), or None for a duplicate / late frame."""
        if source not in SOURCES:
            raise DdmError(f"FrameClock.place: unknown time source {source!r}")
        slot, n = C.c_int32(), C.c_int64()
        rc = self._lib.ddm_clock_place(self._h, float(tMs), SOURCES.index(source), C.byref(slot), C.byref(n))
        if rc < 0:
            raise DdmError("FrameClock.place: time stamps must be finite")
        return None if rc else (int(slot.value), int(n.value))

    def drop(self, n: int) -> None:
        """Frame n was not ingested (e.g. the producer's back-pressure)."""
        self._lib.ddm_clock_drop(self._h, int(n))

    def snapshot(self, guard: int = 0, shuffle: bool = False, rand=None) -> Snapshot:
        order = np.zeros(self.T, np.uint32)
        mask = np.zeros(self.T, np.float32)
        used = C.c_int32()
        if self._lib.ddm_clock_snapshot(self._h, int(guard), order.ctypes.data, mask.ctypes.data, C.byref(used)):
            raise DdmError("FrameClock.snapshot: guard must be >= 0")
        if shuffle:
            order, mask = shuffle_order(order, mask, rand if rand is not None else np.random.default_rng().random)
        return Snapshot(order, mask, int(used.value), self.newestN)
