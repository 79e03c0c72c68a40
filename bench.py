#!/usr/bin/env python3
"""Benchmark of the DDM hot path on one MI355X.

One step = one pass of the hot path over one batch: the resident frame stack
(1024 frames of 512x512 uint16) -> windowed R2C 2D FFT + corner turn into the
spectral history -> masked Wiener-Khinchin D(q, tau) for all lags 1..1023 with
radial binning into 256 unit-width |q| bins.  Inputs are in HBM before the
timed region starts.  Steps alternate between `--inflight` engines (default 2,
each with its own stream and history), so consecutive analyses overlap on the
GPU; every step is a complete pass.  Per-kernel times and the roofline come
from serial steps on one engine (no overlap), reported beside.

Prints one JSON line.  value = frame-pixels/s through to D(q, tau).
`roofline` is for the dominant kernel -- the longest launch, k_temporal_w
(the row transform takes about as long per step, in ten short launches; every
kernel's algorithmic GB/s is listed under pipeline.kernels_serial) -- timed
live with HIP events on the engine's stream, on the binned modes' spectra it
must read (its `limit` says what sets its time).  `pipeline.ring` times the
live-ring case (rotated order, guard mask) beside the identity order.
`cpu_baseline` is the reference's CPU path restated in JavaScript under Node
(one thread and over worker_threads) on config 1 and a 64-frame slice of this
workload.  `fits` is the second half of the metric: the global-parameter fit
of each of 4096 held-out simulated videos (256x256x512, 32 bins each; alpha,
vbar, Z, D shared by a video's bins), plain grid-seeded LM against
surrogate-seeded LM, with the acceptance rule RSS_seeded <= RSS_plain
(1 + 1e-6) and a plain refit of rejected videos charged to the seeded side;
the per-q comparison on the same videos beside it.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
# what sets each kernel's time, from its counters (rocprofv3 --pmc summaries under profiles/): the
# temporal kernel's HBM traffic is ~1.03x its algorithmic bytes, and its time is FP32 VALU issue
LIMIT = {"temporal_wk": "VALU issue (FP32): counter traffic ~1.03x the algorithmic bytes, duration set by "
                        "instruction issue (roofline.valu), not by HBM",
         "spatial_rows": "HBM", "spatial_cols_cornerturn": "HBM", "radial_finalize": "latency"}
METRIC = "frame-pixels/s to D(q,τ) (%HBM roofline); fits/s seeded vs plain LM @equal χ²"


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--frames", type=int, default=1024)
    p.add_argument("--size", type=int, default=512)
    p.add_argument("--bins", type=int, default=256)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-frames", type=int, default=64, help="frames in the CPU-baseline slice")
    p.add_argument("--fit-videos", type=int, default=4096, help="held-out videos for the fits/s comparison")
    p.add_argument("--no-fits", action="store_true")
    p.add_argument("--no-batch", action="store_true")
    p.add_argument("--inflight", type=int, default=2, help="analyses in flight (engines, one stream each)")
    return p.parse_args()


def fits_section(nvideos):
    """The fits half of the metric on held-out simulated videos (256x256x512, 32 bins each).  Headline: the
    global-parameter fit of every video (alpha, vbar, Z, D shared by its bins, per-bin amplitudes: the
    reference's fitSwimmer objective), plain grid-seeded LM against LM seeded by the surrogate's prediction;
    a seeded fit counts only at RSS <= plain RSS (1 + 1e-6), rejected videos are refitted plain and that
    launch is charged to the seeded side (speedup_with_fallback).  Beside it, the per-q LM comparison on the
    same videos' 32 bins each."""
    from hackathon_amd import surrogate as S
    from hackathon_amd.fit import Fitter
    if not os.path.exists(S.WEIGHTS_PATH):
        return {"skipped": "no trained surrogate weights"}
    sim = S.Simulator()
    fitter = Fitter()
    X, _, truth, probs = S.build_dataset(nvideos, S.HELDOUT_SEED0, sim, fitter, keep_problems=True,
                                         log=lambda m: print("[bench] fits: " + m, file=sys.stderr, flush=True))
    sur = S.Surrogate()
    pred, mlp_ms = S.timed_predict(sur, X)
    glob, _, _ = S.compare_global(fitter, probs.prepared, S.global_seeds(pred), truth, mlp_ms)
    perq, _, _ = S.compare(fitter, probs, S.Surrogate.seeds(pred, probs), truth, mlp_ms)
    return {
        "workload": (f"{nvideos} held-out videos of 256x256x512 u8 (seeds {S.HELDOUT_SEED0}..), 32 bins each; "
                     f"global: one fit per video (4 shared + 64 per-bin parameters), plain = the reference's "
                     f"grids (49-point ln D, 16 x 2 x 2 swimmer starts) and LM from the best 3, seeded = LM from "
                     f"the MLP's prediction; fits/s from HIP-event kernel times (+ MLP inference)"),
        "headline": "global",
        "speedup_with_fallback": glob["speedup_with_fallback"],
        "accept_rate": glob["accept_rate"],
        "global": glob,
        "per_q": perq,
    }


def batch_section(nvideos=64, seed=3, reps=3):
    """Batched uneven workload: `nvideos` seeded videos, frame side in {128, 256, 384, 512}, length uniform in
    [200, 2000], all analysed in one HIP-graph launch, then all their q-bins in one batched LM launch."""
    from hackathon_amd import batch as BA
    from hackathon_amd.synth import GpuSynth
    rng = np.random.default_rng(seed)
    g = GpuSynth()
    vids, fpx = [], 0
    for i in range(nvideos):
        N = int(rng.choice([128, 256, 384, 512]))
        T = int(rng.integers(200, 2001))
        al, vb, Z, D = rng.uniform(0, 1), rng.uniform(5, 30), rng.uniform(1, 10), rng.uniform(0.05, 1)
        fr = g.video(g.params(T, N, max(50, 200 * N * N // (256 * 256)), 1.5, 100.0, 20.0, D=D, alpha=al,
                              vbar_px_frame=vb / 100.0, Z=Z, dtype="uint8", seed=seed * 100000 + i, heading3d=True,
                              swim_diffuse=True))
        vids.append(dict(frames=fr, dtype=np.uint8, T=T, L=T, NB=N // 2, kmin=0, kmax=N // 2))
        fpx += T * N * N
    b = BA.Batch(vids, mode=BA.GRAPH, chains=8)
    b.run()                                   # warm
    ms = min(b.run() for _ in range(reps))
    rads = b.radials()
    prs, res = BA.fit_all(rads, 1.0, 100.0, nstarts=3)
    b.close()
    # the same videos with the temporal stages as persistent work-queue kernels (one per transform length)
    bq = BA.Batch(vids, mode=BA.GRAPH | BA.QUEUE, chains=8)
    bq.run()
    msq = min(bq.run() for _ in range(reps))
    bq.close()
    return {"workload": f"{nvideos} videos, N in {{128, 256, 384, 512}}, T uniform in [200, 2000], u8, lags "
                        f"1..T-1, N/2 unit-width bins; one HIP-graph launch (8 concurrent chains)",
            "frame_pixels": fpx, "ms": ms, "frame_pixels_per_s": fpx / (ms * 1e-3),
            "work_queue_ms": msq, "work_queue_frame_pixels_per_s": fpx / (msq * 1e-3),
            "fits": int(len(res.chi2)), "fit_kernel_ms": res.kernel_ms,
            "fits_per_s": len(res.chi2) / (res.kernel_ms * 1e-3)}


def copy_bandwidth(nbytes=1 << 30, reps=10):
    """Device-to-device copy rate (read + write bytes / s, GB/s): the achievable HBM figure beside the 8 TB/s peak."""
    import torch
    a = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    e1.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del a, b
    return 2 * nbytes / (ms * 1e-3) / 1e9


def cpu_baseline(N, frames, nbins):
    """The reference's CPU path (History.ingest -> analyzeRadial, restated in JavaScript:
    oracle/cpu_path.mjs) under Node on the host cores: C1 in full (64 x 128^2 u8, 64 bins) and a slice of
    this workload (the first `frames` frames of N^2 u16 with T = L = frames, `nbins` bins), each on one
    thread and sharded over worker_threads; the FP64 C oracle's fitBrownian + fitSwimmer of the C1 record
    (one thread).  value = the slice's frame-pixels/s on all threads used."""
    import platform
    import oracle as O
    from hackathon_amd import synth
    from hackathon_amd.engine import Radial
    threads = max(1, min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "16"))))
    cpu = platform.processor() or ""
    try:
        cpu = next(l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo") if l.startswith("model name"))
    except (OSError, StopIteration):
        pass
    strip = lambda r: {k: v for k, v in r.items() if k not in ("D", "P", "drift", "node", "cpu", "cpus")}
    c1 = synth.config1()
    b1 = O.make_bins(128, 64, 0, 64)
    runs = {"c1_1": strip(O.cpu_path(c1, b1, 64, threads=1)),
            f"c1_{threads}": strip(O.cpu_path(c1, b1, 64, threads=threads))}
    c2 = dict(synth.C2)
    c2.update(T=frames, N=N)
    fr = synth.video_numpy(dtype=np.uint16, seed=2, **c2)
    b2 = O.make_bins(N, nbins, 0, N // 2, with_nodes=False)
    one = O.cpu_path(fr, b2, frames, threads=1)
    many = O.cpu_path(fr, b2, frames, threads=threads)
    runs["c2_slice_1"], runs[f"c2_slice_{threads}"] = strip(one), strip(many)
    # full C2 (1024 frames, lags to 1023): ingest scales with the frame count, the temporal stage with the
    # per-mode transform length n log n (n = 2T)
    T2 = 1024
    lg = lambda t: (2 * t) * np.log2(2 * t)
    ext = {k: r["ingest_s"] * T2 / frames + r["temporal_s"] * lg(T2) / lg(frames)
           for k, r in (("1", one), (str(threads), many))}
    # the reference's fits (fitBrownian + fitSwimmer, FP64 C oracle) on the C1 record
    hist = O.history(c1, 64)
    D, P = O.analyze_radial(hist, np.arange(64), np.ones(64, np.float32), b1, 64)
    rec = dict(N=128, NB=64, L=64, D=D.ravel(), nModes=b1["nModes"], kCenter=b1["kCenter"], kNodes=b1["kNodes"])
    t0 = time.perf_counter()
    pr = O.prepare(rec, 1.0, 100.0)
    Dd, _ = O.fit_brownian(pr)
    O.fit_swimmer(pr, "3d", 0.0, Dd)
    fit_s = time.perf_counter() - t0
    return {"value": many["frame_pixels_per_s"], "unit": "frame-pixels/s", "cores": threads, "kind": "port",
            "sample": f"the reference's CPU path restated in JavaScript (oracle/cpu_path.mjs, Node "
                      f"{one['node']}, FP64 radix-2 Stockham, packed z = m + i p WK, drift sums) on {frames} "
                      f"frames of {N}x{N} u16 (same generator, T=L={frames}, {nbins} bins) over {threads} "
                      f"worker_threads; one thread: {one['frame_pixels_per_s']:.3g} frame-pixels/s",
            "cpu_model": cpu, "host_cpus": os.cpu_count(), "runs": runs,
            "c2_full_extrapolated_s": ext, "c1_fit_s": fit_s, "c1_fits_per_s": 1.0 / fit_s}


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported (the engine runs on a single "
              f"GPU; sharding one video over several GPUs is not implemented)", file=sys.stderr)
        sys.exit(2)
    import torch
    from hackathon_amd import synth
    from hackathon_amd.engine import DdmEngine, K_TEMPORAL, K_ROWS, K_COLS, K_FINALIZE, KERNEL_NAMES

    T, N, NB = a.frames, a.size, a.bins
    L = T
    dev = torch.device("cuda:0")
    c2 = dict(synth.C2)
    c2.update(T=T, N=N)
    print(f"[bench] generating {T} frames of {N}x{N} u16 on the GPU", file=sys.stderr, flush=True)
    frames = synth.video_gpu(dtype="uint16", seed=2, **c2)
    torch.cuda.synchronize()
    # `inflight` engines of the same geometry: consecutive steps go to different
    # engines (own stream, own history), so one analysis's compute-bound
    # temporal stage overlaps the next one's memory-bound spatial stage
    # launch shapes from the autotuner's cache for this geometry (defaults if none)
    from hackathon_amd.autotune import tuned_create
    engs = [tuned_create(np.uint16, N=N, T=T, L=L, NB=NB, kmin=0, kmax=N // 2) for _ in range(a.inflight)]
    order = np.arange(T, dtype=np.uint32)
    mask = np.ones(T, np.float32)

    def step(eng):
        eng.ingestDevice(frames.data_ptr(), np.uint16, T, 0)
        eng.computeTemporal(order, mask)

    print(f"[bench] warmup {a.warmup}", file=sys.stderr, flush=True)
    for i in range(max(a.warmup, a.inflight)):
        step(engs[i % a.inflight])
    for e in engs:
        e.idle()
    # per-kernel times: serial steps on one engine, HIP events on its stream
    prof_steps = max(1, min(a.steps, 5))
    eng = engs[0]
    eng.setProfiling(True)
    ts0 = time.perf_counter()
    for _ in range(prof_steps):
        step(eng)
    eng.idle()
    serial_ms = (time.perf_counter() - ts0) * 1e3 / prof_steps
    times = eng.kernelTimes()
    eng.setProfiling(False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(a.steps):
        step(engs[i % a.inflight])
    for e in engs:
        e.idle()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    r = eng.readRadial()
    assert np.all(np.isfinite(r.D)), "non-finite D"
    for e in engs[1:]:   # every in-flight engine computed the same (deterministic) result
        r2 = e.readRadial()
        assert np.array_equal(r2.D, r.D) and np.array_equal(r2.P, r.P), "in-flight engines disagree"
    ms_step = (t1 - t0) * 1e3 / a.steps
    # the live ring: the write head mid-ring (a rotated order, so the temporal kernel gathers) and a
    # guard mask over the 4 slots nearest it (masked mean, masked pair counts and tail sums)
    ring_order = ((np.arange(T) + T // 3) % T).astype(np.uint32)
    ring_mask = np.ones(T, np.float32)
    ring_mask[-4:] = 0

    def ring_step(e):
        e.ingestDevice(frames.data_ptr(), np.uint16, T, 0)
        e.computeTemporal(ring_order, ring_mask)

    for i in range(a.inflight):
        ring_step(engs[i])
    for e in engs:
        e.idle()
    tr0 = time.perf_counter()
    for i in range(a.steps):
        ring_step(engs[i % a.inflight])
    for e in engs:
        e.idle()
    ring_ms = (time.perf_counter() - tr0) * 1e3 / a.steps
    assert np.all(np.isfinite(engs[0].readRadial().D)), "non-finite D (ring)"
    fpx = T * N * N
    value = fpx / (ms_step * 1e-3)

    # Algorithmic bytes: the raw frames read once, the f32 complex half-plane
    # spectrum of every mode written once (the history holds all of them, as
    # readHistory returns), and the binned modes' spectra read once by the
    # temporal stage (modes outside every bin are never needed).  Per launch:
    # spatial rows = frames read, spatial cols = spectrum write, temporal =
    # binned spectrum read.
    NH = N // 2 + 1
    spec_bytes = T * N * NH * 8
    binned_bytes = T * int(np.sum(r.nModes)) * 8
    raw_bytes = T * N * N * 2
    per_step_algo = {"spatial_rows": raw_bytes, "spatial_cols_cornerturn": spec_bytes, "temporal_wk": binned_bytes,
                     "radial_finalize": 0}
    tun = engs[0].tuning()
    lr = 3 if tun["fft_radix"] == 8 else 2
    # the kernels' names as the profiler reports them (the traffic table below is keyed by them)
    kernel_symbol = {"spatial_rows": f"ddm::k_rows_w<9, unsigned short, {lr}, 2>",
                     "spatial_cols_cornerturn": f"ddm::k_cols_w<9, {tun['cols_f']}, {lr}>",
                     "temporal_wk": (f"ddm::k_temporal_c<11, {lr}, true, true>" if tun["temporal_kernel"] == 2
                                     else f"ddm::k_temporal_w<11, 2, {lr}>"),
                     "radial_finalize": "ddm::k_bin_sum<11>"}
    per_kernel = {}
    for name, (ms, cnt) in times.items():
        per_kernel[name] = {"ms_per_step": ms / prof_steps, "launches_per_step": cnt / prof_steps,
                            "avg_launch_ms": ms / max(cnt, 1)}
        if per_step_algo[name] and ms > 0:
            gbs = per_step_algo[name] / (ms / prof_steps * 1e-3) / 1e9
            per_kernel[name].update(algorithmic_GBs=gbs, frac=gbs / HBM_PEAK_GBS)
    # the longest launch: per step the row transform is about as long, but its
    # ten launches per step are a tenth of it each
    dom = max(per_kernel, key=lambda k: per_kernel[k]["avg_launch_ms"])
    launches = max(per_kernel[dom]["launches_per_step"], 1)
    algo_launch = per_step_algo[dom] / launches
    achieved = algo_launch / (per_kernel[dom]["avg_launch_ms"] * 1e-3) / 1e9
    traffic = None
    valu = None
    # calibrated FETCH_SIZE + WRITE_SIZE per launch of this workload and the SQ_INSTS_VALU count of the same
    # kernels (rocprofv3 passes summarised in profiles/c2_1024x512_hbm_traffic.md and
    # profiles/c2_1024x512_sq_counters.md)
    prof = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hackathon_amd", "data",
                        "hbm_traffic_1024x512_u16.json")
    if (T, N, NB) == (1024, 512, 256) and os.path.exists(prof):
        k = json.load(open(prof))["kernels"].get(kernel_symbol[dom])
        if k:
            traffic = (k["fetch_MiB"] + k["write_MiB"]) * 2 ** 20
            if "valu_insts" in k:
                # FP32 VALU issue roofline: a CU issues one wave64 VALU instruction per cycle (4 SIMDs, 4
                # cycles each); peak at the CU count and the 2.4 GHz peak engine clock
                props = torch.cuda.get_device_properties(0)
                peak = props.multi_processor_count * 2.4e9
                rate = k["valu_insts"] / (per_kernel[dom]["avg_launch_ms"] * 1e-3)
                valu = {"insts_per_launch": k["valu_insts"], "achieved_Ginst_s": rate / 1e9,
                        "peak_Ginst_s": peak / 1e9, "frac": rate / peak,
                        "note": "the kernel's binding limit: wave64 VALU instructions issued per second"}
    kern = {"temporal_wk": ("k_temporal_c (four waves per mode: gather, masked mean removal, 4 x 512-point "
                            "DIF-split TPAD FFT, |Y|^2 and p sums, direct tail lags, per-item partials)"
                            if tun["temporal_kernel"] == 2 else
                            "k_temporal_w (masked WK: gather, mean removal, TPAD FFT, |Y|^2, per-bin partials)"),
            "spatial_cols_cornerturn": "k_cols_w (column FFT + corner turn into the history)",
            "spatial_rows": "k_rows_w (decode, row-mean removal, window, R2C row FFT)",
            "radial_finalize": "k_bin_sum + k_bin_ifft + k_bin_lags(_full)"}[dom]
    pipe_bytes = raw_bytes + spec_bytes + binned_bytes
    pipeline_gbs = pipe_bytes / (ms_step * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "frame-pixels/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded spots: 60% Schulz swimmers, 40% diffusers, Poisson noise; rendered by the "
                "library's HIP generator)",
        "config": {"workload": f"plain DDM, {T} frames of {N}x{N} u16, lags 1..{L - 1}, {NB} unit-width |q| bins",
                   "frames": T, "N": N, "lags": L, "bins": NB},
        "roofline": {"bound": "hbm", "kernel": kern, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes_per_launch": algo_launch,
                     "avg_launch_ms": per_kernel[dom]["avg_launch_ms"],
                     "limit": LIMIT.get(dom), "valu": valu},
        "pipeline": {"algorithmic_bytes": pipe_bytes, "GBs": pipeline_gbs,
                     "frac_of_hbm_peak": pipeline_gbs / HBM_PEAK_GBS, "inflight": a.inflight,
                     "serial_ms_per_step": serial_ms,
                     "ring": {"ms_per_step": ring_ms, "frame_pixels_per_s": fpx / (ring_ms * 1e-3),
                              "order": f"rotated by {T // 3}", "guard": 4},
                     "serial_frame_pixels_per_s": fpx / (serial_ms * 1e-3),
                     "kernels_serial": per_kernel},
    }
    out["pipeline"]["tuning"] = engs[0].tuning()
    for e in engs:
        e.destroy()
    del frames
    out["roofline"]["achievable_copy_GBs"] = copy_bandwidth()
    if not a.no_batch:
        print("[bench] batched uneven videos", file=sys.stderr, flush=True)
        out["batch"] = batch_section()
    if not a.no_fits:
        print("[bench] fits: plain vs seeded LM", file=sys.stderr, flush=True)
        out["fits"] = fits_section(a.fit_videos)
    if not a.no_cpu_baseline:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(N, min(a.cpu_frames, T), NB)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
