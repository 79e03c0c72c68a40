"""Phase-drift sums of readRadial (opt-in) against the FP64 oracle on the engine's own spectra."""
import numpy as np
import pytest

import oracle as O
from hackathon_amd import synth
from hackathon_amd.engine import DdmEngine

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", ["full", "shuffled"])
def test_drift_sums_against_oracle(gpu, case):
    fr = synth.config1()
    N, T, L, NB = 128, 64, 48, 32
    order = np.arange(T, dtype=np.uint32)
    mask = np.ones(T, np.float32)
    if case == "shuffled":
        rng = np.random.default_rng(5)
        mask[rng.uniform(size=T) < 0.15] = 0
        mask[0] = 1
        order, mask = O.shuffle_order(order, mask, O.mulberry32(99))
    with DdmEngine.create(N=N, T=T, L=L, NB=NB, kmin=2, kmax=60) as e:
        assert np.all(e.readRadial(drift=True).drift == 0)     # nothing analysed yet
        e.ingestFrames(fr, 0)
        e.computeTemporal(order, mask)
        r = e.readRadial(drift=True)
        r0 = e.readRadial()
        h = e.readHistory().astype(np.complex128)
        bins = O.make_bins(N, NB, 2, 60, with_nodes=False)
    assert np.array_equal(r.D, r0.D) and np.all(r0.drift == 0)
    dr = r.drift.reshape(NB, L, 8)
    ref = O.radial_drift(h, order, mask, bins, L)
    ok = bins["nModes"] > 0
    for c, name in enumerate(("Sw", "Sxx", "Sxy", "Syy", "Rx", "Ry", "Rpp", "count")):
        g, rf = dr[ok][:, :, c], ref[ok][:, :, c]
        scale = np.maximum(np.max(np.abs(rf), axis=1, keepdims=True), 1e-30)
        err = np.abs(g - rf) / scale
        print(f"{case} {name}: max {err.max():.2e} p99 {np.quantile(err, 0.99):.2e}")
        if name == "count":
            assert np.array_equal(g, rf)
        elif name in ("Sw", "Sxx", "Sxy", "Syy"):
            assert err.max() < 1e-5
        else:   # phase terms: a z on the negative real axis may take +pi or -pi
            assert np.quantile(err, 0.99) < 1e-4
    assert np.all(dr[:, L - 1] == 0)
