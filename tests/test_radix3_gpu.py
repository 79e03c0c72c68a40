"""Frame sides with a factor 3 (N = 3 * 2^m, e.g. the 384 of the batched uneven workload) against the FP64
oracle, whose transform falls back to the direct DFT for them (the reference's own FFT takes powers of two
only, src/fft.js:35-39): the spectra (radix-3 first pass, then radix 8/4/2), D(q, tau) per element, readMean;
and a 64-video batch over N in {128, 256, 384, 512}, T in [200, 2000] against per-video engines, with the
temporal stages per video and as persistent work-queue kernels."""
import numpy as np
import pytest

import oracle as O
from hackathon_amd import batch as BA
from hackathon_amd.engine import DdmEngine
from hackathon_amd.synth import GpuSynth
from tests.test_parity_gpu import _frames, _spec_err

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N", [48, 96, 192, 384])
@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.float32])
def test_spectrum_parity_radix3(gpu, N, dtype):
    T = 4
    fr = _frames(T, N, dtype, seed=N + 3)
    print(f"spectrum N={N} dtype={np.dtype(dtype).name}", flush=True)
    with DdmEngine.create(N=N, T=T, L=2, NB=4, kmin=0, kmax=N // 2) as e:
        assert e.ingestFrames(fr, 0)
        h = e.readHistory()
        m = e.readMean()
    assert _spec_err(h, O.history(fr, T)) < 1e-5
    ref = float(np.mean(fr[-1].astype(np.float64))) * O.SCALE[np.dtype(dtype)]
    assert abs(m - ref) <= 1e-7 * abs(ref)


def test_spectrum_parity_768(gpu):
    fr = _frames(2, 768, np.uint16, seed=11)
    print("spectrum N=768", flush=True)
    with DdmEngine.create(N=768, T=2, L=2, NB=4, kmin=0, kmax=384) as e:
        assert e.ingestFrames(fr, 0)
        h = e.readHistory()
    assert _spec_err(h, O.history(fr, 2)) < 1e-5


@pytest.mark.parametrize("shuffled", [False, True])
def test_radial_parity_384(gpu, shuffled):
    N, T, L, NB = 384, 48, 48, 96
    g = GpuSynth()
    fr = g.video(g.params(T, N, 300, 1.5, 100, 20, D=0.4, alpha=0.5, vbar_px_frame=0.2, Z=3, dtype="uint8", seed=7))
    order = np.arange(T, dtype=np.uint32)
    mask = np.ones(T, np.float32)
    if shuffled:
        order = np.random.default_rng(2).permutation(T).astype(np.uint32)
        mask[[3, 17, 30]] = 0
    print(f"radial N=384 shuffled={shuffled}", flush=True)
    with DdmEngine.create(N=N, T=T, L=L, NB=NB, kmin=0, kmax=N // 2) as e:
        e.ingestDevice(fr.data_ptr(), np.uint8, T, 0)
        e.computeTemporal(order, mask)
        r = e.readRadial()
        bins = e.bins
    host = fr.cpu().numpy()
    ob = O.make_bins(N, NB, 0, N // 2)
    assert np.array_equal(bins.binOffsets, ob["binOffsets"]) and np.array_equal(bins.nModes, ob["nModes"])
    refD, refP = O.analyze_radial(O.history(host, T), order, mask, ob, L, split=True)
    D = r.D.reshape(NB, L)
    ok = ob["nModes"] > 0
    d = refD[ok][:, 1:]
    rel = np.abs(D[ok][:, 1:] - d) / np.abs(d)
    print(f"per element max {rel.max():.2e}", flush=True)
    assert rel.max() < 1e-5
    assert np.max(np.abs(r.P[ok] - refP[ok]) / refP[ok]) < 1e-5


def test_c3_batch_64_videos(gpu):
    """64 videos, frame sides from {128, 256, 384, 512}, lengths uniform in [200, 2000], u8, every lag, N/2
    bins, in one HIP graph: every record bit-identical to the same video through its own engine."""
    rng = np.random.default_rng(33)
    g = GpuSynth()
    vids = []
    for i in range(64):
        N = int(rng.choice([128, 256, 384, 512]))
        T = int(rng.integers(200, 2001))
        fr = g.video(g.params(T, N, max(50, 200 * N * N // 65536), 1.5, 100.0, 20.0, D=rng.uniform(0.05, 1),
                              alpha=rng.uniform(0, 1), vbar_px_frame=rng.uniform(5, 30) / 100, Z=rng.uniform(1, 10),
                              dtype="uint8", seed=9000 + i))
        vids.append(dict(frames=fr, dtype=np.uint8, T=T, L=T, NB=N // 2, kmin=0, kmax=N // 2))
    print("sizes", sorted({v["frames"].shape[-1] for v in vids}), flush=True)
    assert 384 in {v["frames"].shape[-1] for v in vids}
    b = BA.Batch(vids, mode=BA.GRAPH, chains=8)
    ms = b.run()
    rads = b.radials()
    bq = BA.Batch(vids, mode=BA.GRAPH | BA.QUEUE, chains=8)   # persistent work-queue temporal stage
    msq = bq.run()
    radq = bq.radials()
    print(f"batch {ms:.1f} ms, work-queue batch {msq:.1f} ms", flush=True)
    for i, v in enumerate(vids):
        T = v["T"]
        with DdmEngine.create(N=v["frames"].shape[-1], T=T, L=T, NB=v["NB"], kmin=0, kmax=v["kmax"]) as e:
            e.ingestDevice(v["frames"].data_ptr(), np.uint8, T, 0)
            e.computeTemporal(np.arange(T, dtype=np.uint32), np.ones(T, np.float32))
            r = e.readRadial()
        assert np.array_equal(rads[i].D, r.D) and np.array_equal(rads[i].P, r.P), i
        assert np.array_equal(radq[i].D, r.D) and np.array_equal(radq[i].P, r.P), i
    b.close()
    bq.close()
