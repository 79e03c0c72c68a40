"""Launch-shape autotuner for one geometry.

Candidates cover the engine's tunables (include/ddm.h `ddm_tuning`): the
temporal work split (target item count: modes per workgroup), the spatial
chunk (row-spectrum scratch budget, i.e. how long a chunk stays in the
last-level cache), workgroup- vs wavefront-level spatial FFTs (N = 256 .. 2048),
frames per column workgroup and the FFT radix (radix-8 passes, or radix 4 only,
in the spatial and temporal transforms).  Every candidate first has to reproduce the
default configuration's D(q, tau) and P(q) on the caller's frames (relative
1e-5 of the largest value, the FP32 bar of the path: candidates differ only in
FP32 summation order); only then is it timed.  Timing runs `inflight` analyses
at once (engines of the candidate's shape on their own streams, consecutive
analyses alternating between them, as a pipelined caller drives the engine):
small spatial chunks interleave better with the other analysis's temporal
kernel than the serial time of one engine suggests.  Results can be kept in a
JSON cache keyed by geometry and pixel type (`tuned_create`).
"""
from __future__ import annotations

import itertools
import json
import os
import time

import numpy as np

from .engine import DdmEngine

_HERE = os.path.dirname(os.path.abspath(__file__))
CACHE_PATH = os.path.join(_HERE, "data", "tuning_cache.json")
PARITY_RTOL = 1e-5


def candidates(N: int, quick: bool = False):
    items = (2048, 8192) if quick else (1024, 2048, 4096)
    chunks = (128,) if quick else (64, 96, 128, 192, 256, 512, 1024)
    waves = (1, 2) if N in (256, 512, 1024, 2048) else (1,)
    radices = (8, 4)
    out = []
    for it, ch, wv, rx in itertools.product(items, chunks, waves, radices):
        # the wavefront column pass takes 8- or 16-frame tiles (8 only at N = 2048)
        for cf in (((8,) if N == 2048 else (8, 16)) if wv == 2 else (16,)):
            out.append(dict(temporal_items=it, chunk_mb=ch, wave_spatial=wv, cols_f=cf, fft_radix=rx))
    return out


def _run(eng: DdmEngine, dptr, dtype, nframes, order, mask):
    eng.ingestDevice(dptr, dtype, nframes, 0)
    eng.computeTemporal(order, mask)
    eng.idle()


def key(N, T, L, NB, dtype) -> str:
    return f"N{N}_T{T}_L{L}_NB{NB}_{np.dtype(dtype).name}"


def _time_pipelined(engs, dptr, dtype, nframes, order, mask, reps):
    """Best of `reps` rounds of 2 x len(engs) analyses alternating over the engines: ms per analysis."""
    n = 2 * len(engs)
    best = float("inf")
    for _ in range(reps):
        t0 = time.perf_counter()
        for i in range(n):
            e = engs[i % len(engs)]
            e.ingestDevice(dptr, dtype, nframes, 0)
            e.computeTemporal(order, mask)
        for e in engs:
            e.idle()
        best = min(best, (time.perf_counter() - t0) * 1e3 / n)
    return best


def autotune(geometry: dict, dptr: int, dtype, nframes: int, reps: int = 3, quick: bool = False, log=None,
             inflight: int = 1):
    """geometry: DdmEngine.create kwargs; dptr: device frame stack.  Returns (best tuning, table)."""
    T = geometry["T"]
    order = np.arange(T, dtype=np.uint32)
    mask = (np.arange(T) < nframes).astype(np.float32)
    ref = DdmEngine.create(**geometry)
    _run(ref, dptr, dtype, nframes, order, mask)
    r0 = ref.readRadial()
    default = ref.tuning()
    ref.destroy()
    scaleD = max(np.max(np.abs(r0.D)), 1e-30)
    scaleP = max(np.max(np.abs(r0.P)), 1e-30)
    table = []
    for cand in candidates(geometry["N"], quick):
        eng = DdmEngine.create(**geometry, tuning=cand)
        _run(eng, dptr, dtype, nframes, order, mask)
        r = eng.readRadial()
        err = max(np.max(np.abs(r.D - r0.D)) / scaleD, np.max(np.abs(r.P - r0.P)) / scaleP)
        row = dict(tuning=eng.tuning(), parity_err=float(err), ok=bool(err <= PARITY_RTOL), ms=None)
        if row["ok"]:
            engs = [eng] + [DdmEngine.create(**geometry, tuning=cand) for _ in range(inflight - 1)]
            try:
                for e in engs[1:]:
                    _run(e, dptr, dtype, nframes, order, mask)
                row["ms"] = _time_pipelined(engs, dptr, dtype, nframes, order, mask, reps)
            finally:
                for e in engs[1:]:
                    e.destroy()
        eng.destroy()
        table.append(row)
        if log:
            log(f"  {row}")
    ok = [r for r in table if r["ok"]]
    if not ok:
        return default, table
    return min(ok, key=lambda r: r["ms"])["tuning"], table


def load_cache(path: str = CACHE_PATH) -> dict:
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return {}


def save_cache(entries: dict, path: str = CACHE_PATH) -> None:
    cur = load_cache(path)
    cur.update(entries)
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        json.dump(cur, f, indent=1, sort_keys=True)


def tuned_create(dtype, path: str = CACHE_PATH, **geometry) -> DdmEngine:
    """DdmEngine.create with the cached tuning of this geometry (defaults when none is cached)."""
    g = dict(geometry)
    k = key(g["N"], g["T"], g["L"], g["NB"], dtype)
    tn = load_cache(path).get(k)
    return DdmEngine.create(**g, tuning=tn)


def main(argv=None):
    """Tune one geometry on a seeded synthetic video and (optionally) cache the result."""
    import argparse
    import torch  # noqa: F401  (device memory for the frame stack)
    from .synth import GpuSynth
    ap = argparse.ArgumentParser(description="autotune the engine's launch shapes for one geometry")
    ap.add_argument("--N", type=int, default=512)
    ap.add_argument("--T", type=int, default=1024)
    ap.add_argument("--L", type=int, default=0, help="lags (default T)")
    ap.add_argument("--NB", type=int, default=256)
    ap.add_argument("--dtype", default="uint16", choices=("uint8", "uint16", "float32"))
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--quick", action="store_true")
    ap.add_argument("--inflight", type=int, default=2, help="analyses in flight while timing (bench.py's default)")
    ap.add_argument("--save", action="store_true", help=f"merge the winner into {CACHE_PATH}")
    a = ap.parse_args(argv)
    L = a.L or a.T
    g = GpuSynth()
    fr = g.video(g.params(a.T, a.N, max(200, a.N * a.N // 64), 2.0, 100.0, 200.0, D=0.3, alpha=0.6,
                          vbar_px_frame=0.15, Z=3.0, dtype=a.dtype, seed=7))
    geom = dict(N=a.N, T=a.T, L=L, NB=a.NB, kmin=0, kmax=a.N // 2)
    best, table = autotune(geom, fr.data_ptr(), np.dtype(a.dtype), a.T, reps=a.reps, quick=a.quick,
                           log=lambda m: print(m, flush=True), inflight=max(1, a.inflight))
    print(json.dumps({"best": best, "key": key(a.N, a.T, L, a.NB, a.dtype)}), flush=True)
    if a.save:
        save_cache({key(a.N, a.T, L, a.NB, a.dtype): best})
    return best


if __name__ == "__main__":
    main()
