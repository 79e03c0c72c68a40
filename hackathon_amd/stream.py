"""Streaming ring mode: the frame clock that maps live frames to history slots.

`FrameClock` keeps the reference's FrameClock semantics (motility-ddm
src/camera.ts:306-425): frame n = round((t - t0) / dt) lives in slot n mod T
of the engine's history ring (`DdmEngine.ingestLuma(luma, slot)`); the slot
remembers which frame index it holds, so a snapshot's mask excludes dropped,
late, duplicated and overwritten frames by construction.  `snapshot` returns the
time-ordered (order, mask) pair for `computeTemporal`, oldest first, with the
oldest `guard` time indices masked (slots a producer may overwrite while the
analysis runs) and optionally shuffled (the frame-order null test).

T is the engine's ring length (the reference fixes it at 256).  Time stamps are
milliseconds; time sources are 'captureTime', 'watchdog' (one timebase) and
'mediaTime' (a separate one: switching between the two restarts the clock).
On this engine an ingest issued after computeTemporal is ordered behind it on
the engine's stream, so a snapshot's slots cannot change under a running
analysis; the guard keeps the reference's policy for producers that race it.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class Snapshot:
    order: np.ndarray   # uint32 [T]: slot of time index j
    mask: np.ndarray    # float32 [T]: 1 where frame newestN - T + 1 + j is in its slot
    used: int           # valid frames in the snapshot
    newestN: int        # newest frame index at snapshot time


def shuffle_order(order: np.ndarray, mask: np.ndarray, rand):
    """shuffleOrder (src/ddm_ref.js:510-519): Fisher-Yates over the (slot, mask) pairs, from the last
    position down, j = floor(rand() (i + 1)) with rand() uniform in [0, 1)."""
    order = np.array(order, dtype=np.uint32)
    mask = np.array(mask, dtype=np.float32)
    for i in range(len(order) - 1, 0, -1):
        j = int(rand() * (i + 1))
        order[i], order[j] = order[j], order[i]
        mask[i], mask[j] = mask[j], mask[i]
    return order, mask


class FrameClock:
    def __init__(self, fps: float, T: int = 256):
        if not (fps > 0) or T < 1:
            raise ValueError("FrameClock: fps must be positive and T >= 1")
        self.T = int(T)
        self.dtMs = 1000.0 / fps
        self.frames = self.missed = self.duplicates = self.resets = 0
        self._slot_n = np.full(self.T, -1, dtype=np.int64)
        self._t0 = None
        self._last_n = -1
        self._src = None
        self._last_ts = float("nan")
        self._dts: list[float] = []

    @property
    def timeSource(self):
        return self._src

    @property
    def newestN(self) -> int:
        return self._last_n

    def reset(self) -> None:
        self._t0 = None
        self._last_n = -1
        self._slot_n.fill(-1)
        self._src = None
        self._last_ts = float("nan")
        self.resets += 1

    def place(self, tMs: float, source: str = "captureTime"):
        """Slot for a frame captured at tMs: (slot, n), or None for a duplicate / late frame."""
        if self._src is not None and (source == "mediaTime") != (self._src == "mediaTime"):
            self.reset()
        self._src = source
        if self._t0 is None:
            self._t0 = tMs
        n = _round_half_up((tMs - self._t0) / self.dtMs)
        if n <= self._last_n:
            self.duplicates += 1
            return None
        if self._last_n >= 0 and n - self._last_n > self.T:
            # a gap longer than the ring: nothing in it is usable, restart at this frame
            self.reset()
            self._src = source
            self._t0 = tMs
            n = 0
        if self._last_n >= 0:
            self.missed += n - self._last_n - 1
        if self._last_ts == self._last_ts:       # not NaN
            self._dts.append(tMs - self._last_ts)
            if len(self._dts) > 120:
                self._dts.pop(0)
        self._last_ts = tMs
        slot = n % self.T
        self._slot_n[slot] = n
        self._last_n = n
        self.frames += 1
        return slot, n

    def drop(self, n: int) -> None:
        """Frame n was not ingested (e.g. the producer's back-pressure)."""
        slot = n % self.T
        if self._slot_n[slot] == n:
            self._slot_n[slot] = -1

    @property
    def medianDtMs(self) -> float:
        if not self._dts:
            return float("nan")
        s = sorted(self._dts)
        return s[len(s) >> 1]

    def _window(self):
        n = self._last_n - self.T + 1 + np.arange(self.T, dtype=np.int64)
        slot = np.mod(n, self.T)
        return n, slot, (n >= 0) & (self._slot_n[slot] == n)

    @property
    def validCount(self) -> int:
        return int(np.count_nonzero(self._window()[2]))

    def snapshot(self, guard: int = 0, shuffle: bool = False, rand=None) -> Snapshot:
        n, slot, ok = self._window()
        ok = ok & (np.arange(self.T) >= guard)
        order = slot.astype(np.uint32)
        mask = ok.astype(np.float32)
        used = int(np.count_nonzero(ok))
        if shuffle:
            if rand is None:
                rng = np.random.default_rng()
                rand = rng.random
            order, mask = shuffle_order(order, mask, rand)
        return Snapshot(order, mask, used, self._last_n)


def _round_half_up(x: float) -> int:
    """Math.round: nearest integer, halves toward +infinity."""
    return int(np.floor(x + 0.5))
