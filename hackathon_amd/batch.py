"""Batches of uneven videos through one submission (libddm.so `ddm_batch_*`).

Each video keeps its own geometry (frame side, length, lags, bins) and engine;
the batch records every video's ingest + temporal + radial work into one HIP
graph (or enqueues it eagerly) with `chains` concurrent branches.  Per-video
results are read through `engine(i)`, a non-owning `DdmEngine` view, and the
per-q fits of all bins of all videos go to one `Fitter.fit` launch
(`fit_all`); `Batch.verdicts` runs the verdict layer of every video on host
threads.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from .engine import DdmEngine, DdmError, Geometry, Tuning, load_library
from .fit import SWIMMER_3D, Fitter, bounds, prepare

EAGER, GRAPH, QUEUE = 0, 1, 2   # QUEUE is or-ed with EAGER or GRAPH: persistent work-queue temporal stage
_DT = {np.dtype(np.uint8): 0, np.dtype(np.uint16): 1, np.dtype(np.float32): 2}


class BatchVideo(C.Structure):
    _fields_ = [("g", Geometry), ("tuning", Tuning), ("dtype", C.c_int32), ("nframes", C.c_int32),
                ("dframes", C.c_void_p), ("order", C.c_void_p), ("mask", C.c_void_p)]


def _declare(L):
    if getattr(L, "_batch_declared", False):
        return L
    vp = C.c_void_p
    L.ddm_batch_create.argtypes = [C.POINTER(BatchVideo), C.c_int32, C.c_int32, C.c_int32, C.POINTER(vp)]
    L.ddm_batch_run.argtypes = [vp, C.POINTER(C.c_double)]
    L.ddm_batch_engine.argtypes = [vp, C.c_int32, C.POINTER(vp)]
    L.ddm_batch_last_error.argtypes = [vp]
    L.ddm_batch_last_error.restype = C.c_char_p
    L.ddm_batch_destroy.argtypes = [vp]
    L.ddm_batch_destroy.restype = None
    L._batch_declared = True
    return L


class Batch:
    """videos: list of dicts {frames: device tensor (T', N, N) u8/int16(u16 bits)/f32, dtype: numpy dtype,
    T, L, NB, kmin, kmax[, tuning, order, mask, hypot_mode, knodes]}."""

    def __init__(self, videos, mode=GRAPH, chains=8):
        self._lib = _declare(load_library())
        self._keep = []
        arr = (BatchVideo * len(videos))()
        for i, v in enumerate(videos):
            fr = v["frames"]
            N = int(fr.shape[-1])
            g = Geometry(N, int(v["T"]), int(v["L"]), int(v["NB"]), float(v["kmin"]), float(v["kmax"]),
                         int(v.get("hypot_mode", 0)), int(v.get("knodes", 4)))
            order = v.get("order")
            mask = v.get("mask")
            if order is not None:
                order = np.ascontiguousarray(order, dtype=np.uint32)
                self._keep.append(order)
            if mask is not None:
                mask = np.ascontiguousarray(mask, dtype=np.float32)
                self._keep.append(mask)
            arr[i] = BatchVideo(g, Tuning(**v.get("tuning", {})), _DT[np.dtype(v["dtype"])], int(fr.shape[0]),
                                fr.data_ptr(), None if order is None else order.ctypes.data,
                                None if mask is None else mask.ctypes.data)
        self._frames = [v["frames"] for v in videos]   # the batch reads them on every run
        h = C.c_void_p()
        rc = self._lib.ddm_batch_create(arr, len(videos), mode, int(chains), C.byref(h))
        if rc != 0:
            raise DdmError(self._lib.ddm_batch_last_error(None).decode())
        self._h = h
        self.n = len(videos)
        self._views = [None] * self.n

    def run(self) -> float:
        """One pass over every video; returns the GPU time in ms."""
        ms = C.c_double()
        if self._lib.ddm_batch_run(self._h, C.byref(ms)) != 0:
            raise DdmError(self._lib.ddm_batch_last_error(self._h).decode())
        return ms.value

    def engine(self, i: int) -> DdmEngine:
        if self._views[i] is None:
            e = C.c_void_p()
            if self._lib.ddm_batch_engine(self._h, int(i), C.byref(e)) != 0:
                raise DdmError("ddm_batch_engine failed")
            self._views[i] = DdmEngine(e, self._lib, owned=False)
        return self._views[i]

    def radials(self, drift: bool = False):
        return [self.engine(i).readRadial(drift=drift) for i in range(self.n)]

    def verdicts(self, dxUm, fps, geometry="3d", exposureLocked=True, threads: int = 0):
        """analyzeMotility of every video (radial records with drift sums), on the library's host threads."""
        from .verdict import analyze_motility_batch
        return analyze_motility_batch(self.radials(drift=True), dxUm, fps, geometry, exposureLocked, threads)

    def close(self):
        if getattr(self, "_h", None):
            for v in self._views:
                if v is not None:
                    v.destroy()
            self._lib.ddm_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def fit_all(radials, dxUm, fps, fitter: Fitter | None = None, model=SWIMMER_3D, **kw):
    """All usable bins of all videos as one batched LM launch -> (per-video Prepared list, FitResult)."""
    tau, data, w, qn, bd, prs = [], [], [], [], [], []
    K = 1
    for r, f in zip(radials, np.broadcast_to(np.asarray(fps, dtype=float), (len(radials),))):
        pr = prepare(r, dxUm, float(f))
        prs.append(pr)
        if not pr.bins:
            continue
        b4 = bounds(pr)
        for b in pr.bins:
            tau.append(pr.tau)
            data.append(b.d)
            w.append(b.w)
            qn.append(b.qn)
            bd.append(b4)
            K = max(K, len(b.qn))
    qn = np.array([np.resize(q, K) for q in qn])
    res = (fitter or Fitter()).fit(tau, data, np.array(w), qn, np.array(bd), model=model, **kw)
    return prs, res
