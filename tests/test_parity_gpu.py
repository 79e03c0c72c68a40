"""GPU parity: libddm.so (HIP, gfx950) against the FP64 oracle.

Metrics follow the reference's GPU harness (test/gpu/harness.ts:115-166):
spectrum max|dF|/rms|F|; radial D relative to FP64.  Tolerances are written in
each test.
"""
import numpy as np
import pytest

import oracle as O
from hackathon_amd import synth
from hackathon_amd.engine import DdmEngine, DdmError

pytestmark = pytest.mark.gpu


def _frames(T, N, dtype, seed):
    rng = np.random.default_rng(seed)
    if dtype == np.float32:
        return rng.uniform(0, 1, (T, N, N)).astype(np.float32)
    hi = np.iinfo(dtype).max
    base = rng.integers(0, hi // 4, (T, N, N))
    return (base + rng.integers(0, hi // 2, (1, N, N))).astype(dtype)


def _spec_err(g, ref):
    return np.max(np.abs(g - ref)) / np.sqrt(np.mean(np.abs(ref) ** 2))


# ------------------------------------------------------------------ spatial stage (a) + corner turn (b)

@pytest.mark.parametrize("N", [16, 32, 64, 128, 256, 512])
@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.float32])
def test_spectrum_parity(gpu, N, dtype):
    T = 8
    fr = _frames(T, N, dtype, seed=N)
    print(f"spectrum N={N} dtype={np.dtype(dtype).name}", flush=True)
    with DdmEngine.create(N=N, T=T, L=4, NB=4, kmin=0, kmax=N // 2) as e:
        assert e.ingestFrames(fr, 0)
        h = e.readHistory()
    ref = O.history(fr, T)
    err = _spec_err(h, ref)
    # reference harness threshold 1e-4 (harness.ts:25-32); FP32 radix-8 Stockham reaches ~1e-6
    assert err < 1e-5, err


@pytest.mark.parametrize("N,tuning", [(1024, dict(wave_spatial=2, cols_f=8, fft_radix=4)),
                                      (512, dict(wave_spatial=2, fft_radix=4)), (256, dict(fft_radix=4)),
                                      (384, dict(fft_radix=4)),
                                      (1024, dict(wave_spatial=2, cols_f=8)), (1024, dict(wave_spatial=2, cols_f=16)),
                                      (1024, dict(wave_spatial=1)), (512, dict(wave_spatial=1)), (2048, None)])
def test_spectrum_parity_large(gpu, N, tuning):
    """The wavefront spatial kernels at N = 1024 (radix 8, 8, 4, 4; 8- and 16-frame column tiles), the
    workgroup-level kernels at 512 and 1024, and N = 2048."""
    T = 4
    fr = _frames(T, N, np.uint16, seed=N + 7)
    print(f"spectrum N={N} tuning={tuning}", flush=True)
    with DdmEngine.create(N=N, T=T, L=2, NB=4, kmin=0, kmax=N // 2, tuning=tuning) as e:
        assert e.ingestFrames(fr, 0)
        h = e.readHistory()
        if tuning:
            assert all(e.tuning()[k] == v for k, v in tuning.items())
    ref = O.history(fr, T)
    err = _spec_err(h, ref)
    assert err < 1e-5, err


@pytest.mark.parametrize("radix", [8, 4])
def test_spectrum_parity_2048_tiles(gpu, radix):
    """N = 2048 over 24 frames: three 8-frame column tiles (tiles 0 and 1 paired on one XCD, tile 2 alone)
    and 8-row row groups paired the same way; every slot of the history against the FP64 oracle."""
    N, T = 2048, 24
    fr = _frames(T, N, np.uint16, seed=99)
    print(f"spectrum N={N} T={T} radix={radix}", flush=True)
    with DdmEngine.create(N=N, T=T, L=2, NB=4, kmin=0, kmax=N // 2, tuning=dict(fft_radix=radix)) as e:
        assert e.ingestFrames(fr, 0)
        h = e.readHistory()
    ref = O.history(fr, T)
    err = _spec_err(h, ref)
    assert err < 1e-5, err


def test_spectrum_config1(gpu):
    fr = synth.config1()
    with DdmEngine.create(N=128, T=64, L=64, NB=64, kmin=0, kmax=64) as e:
        e.ingestFrames(fr, 0)
        h = e.readHistory()
    assert _spec_err(h, O.history(fr, 64)) < 1e-5


def test_history_slots_wrap_and_chunks(gpu, monkeypatch):
    """Frames go to slots (first + i) mod T; a stack longer than one spatial chunk is split."""
    N, T = 512, 40
    fr = _frames(T, N, np.uint16, seed=5)
    monkeypatch.setenv("DDM_CHUNK_MB", "8")   # 16-frame chunks
    with DdmEngine.create(N=N, T=T, L=4, NB=4, kmin=0, kmax=256) as e:
        assert e.sizes.chunk_frames < T
        assert e.sizes.hist_stride == 48   # device rows of T rounded up to 16 slots; readHistory returns T per row
        e.ingestFrames(fr[:30], 25)    # slots 25..39, 0..14
        e.ingestFrames(fr[30:], 15)    # slots 15..24
        h = e.readHistory()
    slots = [(25 + i) % T for i in range(30)] + list(range(15, 25))
    ref = O.history(fr, T)             # frame i -> slot i
    got = np.empty_like(h)
    got[:, np.arange(T)] = h[:, slots]
    assert _spec_err(got, ref) < 1e-5


def test_ingest_luma_single_frame(gpu):
    N, T = 64, 4
    fr = _frames(1, N, np.float32, seed=9)
    with DdmEngine.create(N=N, T=T, L=2, NB=4, kmin=0, kmax=32) as e:
        assert e.ingestLuma(fr[0].ravel(), 6)   # slot 6 % 4 = 2
        h = e.readHistory()
        with pytest.raises(DdmError, match="expects"):
            e.ingestLuma(np.zeros(10, np.float32), 0)
    ref = O.history(fr, T, first_slot=2)
    assert _spec_err(h[:, 2], ref[:, 2]) < 1e-5
    assert np.all(h[:, [0, 1, 3]] == 0)


@pytest.mark.parametrize("dtype", [np.uint8, np.uint16, np.float32])
@pytest.mark.parametrize("wave", [1, 2])
def test_read_mean(gpu, monkeypatch, dtype, wave):
    """readMean: the last ingested frame's mean luma (the reference's reduce_mean hi + lo), in the spectra's
    units, also when that frame is not the first of the last spatial chunk; 0 before any frame."""
    N, T = 512, 40
    fr = _frames(T, N, dtype, seed=21)
    monkeypatch.setenv("DDM_CHUNK_MB", "8")   # 16-frame chunks: the last frame is frame 7 of the third
    scale = O.SCALE[np.dtype(dtype)]
    with DdmEngine.create(N=N, T=T, L=4, NB=4, kmin=0, kmax=256, tuning=dict(wave_spatial=wave)) as e:
        assert e.readMean() == 0.0
        e.ingestFrames(fr, 0)
        m = e.readMean()
        ref = float(np.mean(fr[-1].astype(np.float64))) * scale
        # u8/u16: exact integer sums; f32: per-row (hi, lo) means
        assert abs(m - ref) <= (1e-15 if dtype != np.float32 else 1e-7) * abs(ref), (m, ref)
        e.ingestFrames(fr[3:4], 5)
        assert abs(e.readMean() - float(np.mean(fr[3].astype(np.float64))) * scale) <= 1e-7 * abs(ref)


# ------------------------------------------------------------------ bins

@pytest.mark.parametrize("N,NB,kmin,kmax", [(256, 32, 4, 100), (128, 64, 0, 64), (256, 128, 0, 128),
                                            (512, 256, 0, 256)])
def test_engine_bins_match_oracle(gpu, N, NB, kmin, kmax):
    with DdmEngine.create(N=N, T=4, L=2, NB=NB, kmin=kmin, kmax=kmax) as e:
        b = e.bins
    o = O.make_bins(N, NB, kmin, kmax)
    assert np.array_equal(b.binOffsets, o["binOffsets"])
    assert np.array_equal(b.binModes, o["binModes"])
    assert np.array_equal(b.nModes, o["nModes"])
    assert np.array_equal(b.kCenter, o["kCenter"])
    assert np.max(np.abs(b.kNodes - o["kNodes"])) < 1e-10


# ------------------------------------------------------------------ temporal + radial (c)

def _radial_errors(gD, gP, refD, refP, nmodes, mask=None):
    ok = nmodes > 0
    d = refD[ok][:, 1:]
    g = gD[ok][:, 1:]
    nz = d != 0          # lags with no valid frame pair have D = 0 on both sides
    assert np.all(g[~nz] == 0)
    rel = np.abs(g[nz] - d[nz]) / np.abs(d[nz])
    if mask is not None and rel.size:   # where the largest error sits: bin, lag, valid pairs, D / bin max
        r2 = np.zeros_like(d)
        r2[nz] = rel
        bi, li = np.unravel_index(int(np.argmax(r2)), r2.shape)
        m = np.asarray(mask, np.float64) != 0
        lag = li + 1
        npairs = int(np.sum(m[:len(m) - lag] & m[lag:]))
        print(f"   worst: bin {np.flatnonzero(ok)[bi]} lag {lag} N(tau) {npairs} "
              f"D/maxD {d[bi, li] / np.max(np.abs(refD[ok][bi])):.3e} rel {r2[bi, li]:.2e}", flush=True)
    scale = np.max(np.abs(refD[ok]), axis=1, keepdims=True)
    harness = np.max(np.abs(gD[ok] - refD[ok]) / scale)
    prel = np.max(np.abs(gP[ok] - refP[ok]) / refP[ok])
    return rel, harness, prel


def _run_radial(fr, T, L, NB, kmin, kmax, order, mask, tuning=None):
    N = fr.shape[1]
    with DdmEngine.create(N=N, T=T, L=L, NB=NB, kmin=kmin, kmax=kmax, tuning=tuning) as e:
        e.ingestFrames(fr, 0)
        e.computeTemporal(order, mask)
        r = e.readRadial()
        h = e.readHistory()
        bins = O.make_bins(N, NB, kmin, kmax, with_nodes=False)
    return r, h, bins


@pytest.mark.parametrize("case", ["full", "dropped", "shuffled"])
def test_radial_parity_config1(gpu, case):
    fr = synth.config1()
    T = L = 64
    NB, kmin, kmax = 64, 0, 64
    order = np.arange(T, dtype=np.uint32)
    mask = np.ones(T, np.float32)
    rng = np.random.default_rng(11)
    if case != "full":
        mask[rng.uniform(size=T) < 0.2] = 0
        mask[0] = 1
    if case == "shuffled":
        order, mask = O.shuffle_order(order, mask, O.mulberry32(4242))
    print(f"radial config1 case={case}", flush=True)
    r, h, bins = _run_radial(fr, T, L, NB, kmin, kmax, order, mask)
    D = r.D.reshape(NB, L)
    # (i) end to end against FP64 spectra + FP64 temporal
    refD, refP = O.analyze_radial(O.history(fr, T), order, mask, bins, L)
    rel, harness, prel = _radial_errors(D, r.P, refD, refP, bins["nModes"], mask)
    print(f"end-to-end: rel p50 {np.median(rel):.2e} p99 {np.quantile(rel, .99):.2e} max {rel.max():.2e} "
          f"harness {harness:.2e} P {prel:.2e}")
    # every element of D within 1e-5 relative of FP64
    assert harness < 1e-5 and prel < 1e-5
    assert rel.max() < 1e-5, rel.max()
    # (ii) stage (c) alone: FP64 temporal on the GPU's own f32 spectra
    refD2, refP2 = O.analyze_radial(h.astype(np.complex128), order, mask, bins, L)
    rel2, harness2, prel2 = _radial_errors(D, r.P, refD2, refP2, bins["nModes"], mask)
    print(f"stage c: rel p50 {np.median(rel2):.2e} p99 {np.quantile(rel2, .99):.2e} max {rel2.max():.2e} "
          f"harness {harness2:.2e} P {prel2:.2e}")
    assert harness2 < 1e-5 and prel2 < 1e-6
    assert rel2.max() < 1e-5, rel2.max()
    # empty bin 0 (only DC lies in [0, 1) and DC is excluded)
    assert bins["nModes"][0] == 0 and np.all(D[0] == 0) and r.P[0] == 0


@pytest.mark.parametrize("N,T,L,NB,kmin,kmax", [(64, 100, 100, 16, 2, 30), (64, 37, 20, 8, 0, 32),
                                                (256, 64, 32, 32, 4, 100)])
def test_radial_parity_geometries(gpu, N, T, L, NB, kmin, kmax):
    """Non power-of-two T (padding), L < T, non-unit bins, the reference's default binning."""
    fr = synth.video_numpy(T, N, 60, 1.5, 100.0, 20.0, D=0.5, dtype=np.uint8, seed=T)
    order = np.arange(T, dtype=np.uint32)
    mask = np.ones(T, np.float32)
    mask[::7] = 0
    mask[0] = 1
    print(f"radial N={N} T={T} L={L} NB={NB}", flush=True)
    r, h, bins = _run_radial(fr, T, L, NB, kmin, kmax, order, mask)
    refD, refP = O.analyze_radial(h.astype(np.complex128), order, mask, bins, L)
    rel, harness, prel = _radial_errors(r.D.reshape(NB, L), r.P, refD, refP, bins["nModes"], mask)
    print(f"rel max {rel.max():.2e} harness {harness:.2e} P {prel:.2e}")
    assert harness < 1e-5 and prel < 1e-6
    assert rel.max() < 1e-5, rel.max()


def _ring_meta(T, case, seed):
    """Order and mask of a live ring: the oldest slot first (rotated order), a few frames dropped and a guard
    of the newest slots invalid; or the same mask under a seeded shuffle."""
    rng = np.random.default_rng(seed)
    head = int(rng.integers(1, T))
    order = ((head + np.arange(T)) % T).astype(np.uint32)
    mask = np.ones(T, np.float32)
    mask[rng.uniform(size=T) < 0.05] = 0
    mask[-3:] = 0          # guard
    mask[0] = 1
    if case == "shuffled":
        order, mask = O.shuffle_order(order, mask, O.mulberry32(seed))
    return order, mask


def _gpu_video(T, N, seed):
    from hackathon_amd.synth import GpuSynth
    g = GpuSynth()
    fr = g.video(g.params(T, N, 120, 1.5, 100.0, 20.0, D=0.4, alpha=0.5, vbar_px_frame=0.2, Z=4.0, dtype="uint8",
                          seed=seed, heading3d=True, swim_diffuse=True))
    return fr.cpu().numpy()


@pytest.mark.parametrize("T", [300, 1000, 2000])
@pytest.mark.parametrize("case", ["ring", "shuffled"])
def test_radial_parity_ring_long(gpu, T, case):
    """Gathered series (a non-identity order) with dropped frames and a guard at lengths that run the split
    sub-transforms (TPAD 1024, 2048, 4096), against the FP64 oracle on the engine's own history.

    Per element: 1e-5 for 99.9 % of the elements.  The exceptions are the last transform lags of slow bins
    (few valid pairs, D << P); there the FP32 transform's error is ~1e-7 x (power of the samples a lag
    leaves unpaired) / (D N), and with dropped frames that power stays large.  They are held to 1e-4."""
    N, NB = 64, 32
    fr = _gpu_video(T, N, seed=T + 1)
    order, mask = _ring_meta(T, case, seed=T)
    print(f"ring parity T={T} case={case}", flush=True)
    r, h, bins = _run_radial(fr, T, T, NB, 0, N // 2, order, mask)
    refD, refP = O.analyze_radial(h.astype(np.complex128), order, mask, bins, T, split=True)
    rel, harness, prel = _radial_errors(r.D.reshape(NB, T), r.P, refD, refP, bins["nModes"], mask)
    print(f"rel p99 {np.quantile(rel, .99):.2e} p99.9 {np.quantile(rel, .999):.2e} max {rel.max():.2e} "
          f"harness {harness:.2e} P {prel:.2e}")
    assert harness < 1e-5 and prel < 1e-6
    assert np.quantile(rel, 0.999) < 1e-5 and rel.max() < 1e-4, rel.max()


@pytest.mark.parametrize("T", [64, 200, 300, 1000, 2000])
def test_radial_parity_radix4(gpu, T):
    """The autotuner's other FFT radix (radix-4 passes only, in the wavefront temporal transforms of every
    length, TPAD 256 .. 4096): the ring case against the FP64 oracle, per element as test_radial_parity_ring_long."""
    N, NB = 64, 32
    fr = _gpu_video(T, N, seed=T + 3)
    order, mask = _ring_meta(T, "ring", seed=T + 1)
    print(f"radix-4 parity T={T}", flush=True)
    r, h, bins = _run_radial(fr, T, T, NB, 0, N // 2, order, mask, tuning=dict(fft_radix=4))
    refD, refP = O.analyze_radial(h.astype(np.complex128), order, mask, bins, T, split=True)
    rel, harness, prel = _radial_errors(r.D.reshape(NB, T), r.P, refD, refP, bins["nModes"], mask)
    print(f"rel p99.9 {np.quantile(rel, .999):.2e} max {rel.max():.2e} harness {harness:.2e} P {prel:.2e}")
    assert harness < 1e-5 and prel < 1e-6
    assert np.quantile(rel, 0.999) < 1e-5 and rel.max() < 1e-4, rel.max()


@pytest.mark.parametrize("case", ["full", "shuffled"])
def test_radial_parity_block_kernel(gpu, case):
    """T = 2100 (TPAD 8192: the workgroup-per-mode temporal kernel), every bin and lag against the FP64 oracle
    on the engine's own history; identity order with every frame valid, and a shuffled order with drops."""
    N, NB, T = 64, 32, 2100
    fr = _gpu_video(T, N, seed=77)
    if case == "full":
        order, mask = np.arange(T, dtype=np.uint32), np.ones(T, np.float32)
    else:
        order, mask = _ring_meta(T, "shuffled", seed=5)
    print(f"block-kernel parity T={T} case={case}", flush=True)
    r, h, bins = _run_radial(fr, T, T, NB, 0, N // 2, order, mask)
    refD, refP = O.analyze_radial(h.astype(np.complex128), order, mask, bins, T, split=True)
    rel, harness, prel = _radial_errors(r.D.reshape(NB, T), r.P, refD, refP, bins["nModes"], mask)
    print(f"rel p99 {np.quantile(rel, .99):.2e} max {rel.max():.2e} harness {harness:.2e} P {prel:.2e}")
    assert harness < 1e-5 and prel < 1e-6
    assert rel.max() < 1e-5, rel.max()


def test_temporal_errors(gpu):
    with DdmEngine.create(N=64, T=16, L=8, NB=4, kmin=0, kmax=32) as e:
        with pytest.raises(DdmError, match="length 16"):
            e.computeTemporal(np.arange(15), np.ones(15))
        bad = np.arange(16, dtype=np.uint32)
        bad[3] = 99
        with pytest.raises(DdmError, match="slot"):
            e.computeTemporal(bad, np.ones(16))
        with pytest.raises(DdmError, match="at most T"):
            e.ingestFrames(np.zeros((17, 64, 64), np.uint8))


def test_deterministic(gpu):
    fr = synth.config1()
    outs = []
    for _ in range(2):
        with DdmEngine.create(N=128, T=64, L=64, NB=64, kmin=0, kmax=64) as e:
            e.ingestFrames(fr, 0)
            e.computeTemporal(np.arange(64), np.ones(64))
            outs.append(e.readRadial())
    assert np.array_equal(outs[0].D, outs[1].D) and np.array_equal(outs[0].P, outs[1].P)
