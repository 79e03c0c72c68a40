"""readTemporalFull: per-mode D, C, P of every mode against the FP64 oracle's direct sums."""
import numpy as np
import pytest

import oracle as O
from hackathon_amd import synth
from hackathon_amd.engine import DdmEngine, DdmError

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N,T,L,case", [(32, 100, 100, "full"), (128, 64, 64, "shuffled"), (64, 300, 250, "dropped")])
def test_temporal_full_against_direct_sums(gpu, N, T, L, case):
    fr = synth.video_numpy(T, N, 40, 1.5, 100.0, 20.0, D=0.5, dtype=np.uint8, seed=N + T)
    order = np.arange(T, dtype=np.uint32)
    mask = np.ones(T, np.float32)
    rng = np.random.default_rng(T)
    if case != "full":
        mask[rng.uniform(size=T) < 0.2] = 0
        mask[0] = 1
    if case == "shuffled":
        order, mask = O.shuffle_order(order, mask, O.mulberry32(77))
    print(f"temporal full N={N} T={T} L={L} {case}", flush=True)
    with DdmEngine.create(N=N, T=T, L=L, NB=N // 4, kmin=0, kmax=N // 2) as e:
        with pytest.raises(DdmError, match="computeTemporal has not run"):
            e.readTemporalFull()
        e.ingestFrames(fr, 0)
        e.computeTemporal(order, mask)
        D, Cc, P = e.readTemporalFull()
        h = e.readHistory().astype(np.complex128)
        r = e.readRadial()
        bins = e.bins
    modes = h.shape[0]
    assert D.shape == (modes, L) and Cc.shape == (modes, L) and P.shape == (modes,)
    sample = np.unique(np.concatenate([np.arange(min(modes, 40)), rng.integers(0, modes, 160), [modes - 1]]))
    eD = eC = eP = 0.0
    for m in sample:
        rD, rC, rP = O.temporal_mode(h, int(m), order, mask, L, direct=True)
        if np.max(np.abs(rD)) > 0:
            eD = max(eD, np.max(np.abs(D[m] - rD)) / np.max(np.abs(rD)))
        if np.max(np.abs(rC)) > 0:
            eC = max(eC, np.max(np.abs(Cc[m] - rC)) / np.max(np.abs(rC)))
        if rP > 0:
            eP = max(eP, abs(P[m] - rP) / rP)
    print(f"per-mode errors: D {eD:.2e} C {eC:.2e} P {eP:.2e}")
    assert eD < 1e-5 and eC < 1e-5 and eP < 1e-6
    # the bin averages of the per-mode D are the radial profile
    NB = N // 4
    off, bm = bins.binOffsets, bins.binModes[:, 0]
    Dr = r.D.reshape(NB, L)
    for b in range(NB):
        if off[b + 1] > off[b]:
            avg = D[bm[off[b]:off[b + 1]]].astype(np.float64).mean(axis=0)
            assert np.max(np.abs(avg - Dr[b])) <= 1e-5 * max(np.max(np.abs(Dr[b])), 1e-30), b
