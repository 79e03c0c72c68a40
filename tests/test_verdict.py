"""analyzeMotility (the verdict layer, libddm.so host code) against the reference's validation matrix.

CPU: scenarios of test/t_validation_matrix.mjs run through the FP64 oracle
pipeline (the reference's simulator restated, history, radial record with drift
sums), then through the library's verdict layer: required verdicts and the
reference's quantitative checks; the exposure gate; the swimmer fit equals the
oracle's restatement of fitSwimmer from the same drift speed and start; a
drift-free record's Brownian fit equals the oracle's fitBrownian; error and
insufficient-data paths.  GPU: the whole matrix through the HIP engine.
"""
import math

import numpy as np
import pytest

import oracle as O
from hackathon_amd.engine import DdmError, Radial
from hackathon_amd.verdict import analyze_motility
from tests.scenarios import SCENARIOS, check

N, NB, L = 256, 32, 128
_BINS = None
_CACHE = {}


def bins():
    global _BINS
    if _BINS is None:
        _BINS = O.make_bins(N, NB, 4, 100)
    return _BINS


def oracle_prepare(r):
    rec = dict(N=r.N, NB=r.NB, L=r.L, D=r.D, nModes=r.nModes, kCenter=r.kCenter, kNodes=r.kNodes)
    return O.prepare(rec, 1.0, 30.0)


def order_mask(sc, valid):
    order, mask = O.linear_order(valid)
    if sc.get("shuffle"):
        order, mask = O.shuffle_order(order, mask, O.mulberry32(4242))
    return order, mask


def oracle_radial(name):
    """The reference's CPU pipeline for one scenario (FP64 oracle), cached per test session."""
    if name not in _CACHE:
        sc = next(s for s in SCENARIOS if s["name"] == name)
        fr, valid, _ = O.simulate(sc["sim"])
        h = O.history(fr, sc["sim"]["T"])
        order, mask = order_mask(sc, valid)
        b = bins()
        D, P = O.analyze_radial(h, order, mask, b, L)
        dr = O.radial_drift(h, order, mask, b, L)
        _CACHE[name] = Radial(N, NB, L, b["kCenter"], b["kNodes"], b["nModes"], D.ravel(), P, dr.ravel())
    return _CACHE[name]


@pytest.mark.parametrize("name", ["motile_3d", "motile_2d", "killed_brownian", "static_debris", "flow_0deg_10ums"])
def test_matrix_on_the_oracle_pipeline(name):
    sc = next(s for s in SCENARIOS if s["name"] == name)
    res = analyze_motility(oracle_radial(name), sc["sim"]["dxUm"], sc["sim"]["fps"], sc["geometry"])
    print(name, res["verdict"], {k: res.get(k) for k in ("TSI", "dBIC", "alpha", "vbar", "Ddiff", "mu")})
    assert check(sc, res) == []
    assert [g["id"] for g in res["gates"]] == ["G1", "G2", "G3", "G4", "G5"]
    assert res["physicsVerdict"] == res["verdict"]


def test_exposure_gate():
    res = analyze_motility(oracle_radial("motile_3d"), 1.0, 30.0, "3d", exposureLocked=False)
    assert res["verdict"] == "UNLOCKED_EXPOSURE" and res["physicsVerdict"] == "MOTILE"
    assert res["gates"][-1]["id"] == "G5" and res["gates"][-1]["pass"] is False


@pytest.mark.parametrize("name,geometry", [("motile_3d", "3d"), ("motile_2d", "2d")])
def test_swimmer_fit_equals_oracle_fit(name, geometry):
    """The reported swimmer fit is fitSwimmer(pr, geometry, Vmag, Dstart) with the drift speed and start the
    analysis used: the oracle's restatement from the same inputs lands on the same optimum."""
    r = oracle_radial(name)
    res = analyze_motility(r, 1.0, 30.0, geometry)
    pr = oracle_prepare(r)
    sw = res["fit"]["swimmer"]
    ref = O.fit_swimmer(pr, geometry, sw["Vmag"], res["DdiffBrownian"])
    for k in ("vbar", "Z", "Dd", "alpha"):
        assert math.isclose(sw[k], ref[k], rel_tol=1e-6, abs_tol=1e-9), (k, sw[k], ref[k])
    assert math.isclose(sw["rss"], ref["rss"], rel_tol=1e-9)


def test_brownian_fit_equals_oracle_fit():
    """A record without drift sums above zero keeps the pass-1 fits: Brownian D and RSS equal fitBrownian's."""
    r = oracle_radial("killed_brownian")
    r0 = Radial(r.N, r.NB, r.L, r.kCenter, r.kNodes, r.nModes, r.D, r.P, np.zeros_like(r.drift))
    res = analyze_motility(r0, 1.0, 30.0, "3d")
    assert res["drift"]["samples"] == 0 and not res["drift"]["significant"]
    Dd, rss = O.fit_brownian(oracle_prepare(r0), 0.0)
    assert math.isclose(res["DdiffBrownian"], Dd, rel_tol=1e-9)
    assert math.isclose(res["fit"]["brownian"]["rss"], rss, rel_tol=1e-9)
    assert res["verdict"] == "NOT_MOTILE"


def test_insufficient_bins_and_errors():
    r = oracle_radial("killed_brownian")
    nm = np.array(r.nModes).copy()
    nm[3:] = 0                                  # 3 usable bins < 4
    res = analyze_motility(Radial(r.N, r.NB, r.L, r.kCenter, r.kNodes, nm, r.D, r.P, r.drift), 1.0, 30.0)
    assert res["verdict"] == "NO_DYNAMICS" and [g["id"] for g in res["gates"]] == ["G1", "G0", "G5"]
    assert "vbar" not in res
    with pytest.raises(DdmError, match="drift"):
        analyze_motility(Radial(r.N, r.NB, r.L, r.kCenter, r.kNodes, r.nModes, r.D, r.P, np.zeros(1)), 1.0, 30.0)
    with pytest.raises(DdmError, match="invalid"):
        analyze_motility(r, 0.0, 30.0)
    with pytest.raises(DdmError, match="geometry"):
        analyze_motility(r, 1.0, 30.0, "4d")


@pytest.mark.gpu
def test_matrix_through_the_gpu_engine(gpu):
    """Every scenario: simulated u8 frames -> HIP engine (the scenario's order and mask) -> readRadial with
    drift sums -> analyzeMotility: required verdicts and checks, plus the exposure gate on motile_3d."""
    from hackathon_amd.engine import DdmEngine
    failures = []
    with DdmEngine.create(N=N, T=256, L=L, NB=NB, kmin=4, kmax=100) as e:
        for sc in SCENARIOS:
            fr, valid, _ = O.simulate(sc["sim"])
            e.ingestFrames(fr, 0)            # dropped frames' slots are masked off
            order, mask = order_mask(sc, valid)
            e.computeTemporal(order, mask)
            r = e.readRadial(drift=True)
            res = analyze_motility(r, sc["sim"]["dxUm"], sc["sim"]["fps"], sc["geometry"])
            f = check(sc, res)
            print(f"{'PASS' if not f else 'FAIL'} {sc['name']:22s} {res['verdict']:19s} TSI={res['TSI']:.2f} "
                  f"dBIC={res.get('dBIC', float('nan')):.0f} alpha={res.get('alpha', float('nan')):.2f} "
                  f"vbar={res.get('vbar', float('nan')):.1f} |V|={res.get('drift', {}).get('mag', float('nan')):.2f} {f}",
                  flush=True)
            failures += [(sc["name"], x) for x in f]
            if sc["name"] == "motile_3d":
                u = analyze_motility(r, 1.0, 30.0, "3d", exposureLocked=False)
                if not (u["verdict"] == "UNLOCKED_EXPOSURE" and u["physicsVerdict"] == "MOTILE"):
                    failures.append(("exposure gate", u["verdict"]))
    assert failures == []


def test_batch_equals_single_records():
    """The threaded batch entry gives each record the single-record result (records, options per record)."""
    from hackathon_amd.verdict import analyze_motility_batch
    names = ["motile_3d", "killed_brownian", "flow_0deg_10ums", "static_debris"]
    rads = [oracle_radial(n) for n in names]
    geo = ["3d", "3d", "3d", "2d"]
    got = analyze_motility_batch(rads, 1.0, 30.0, geo, [True, True, False, True], threads=3)
    def same(a, b):   # equal, NaN == NaN (mu of a record without a tau_1/2 fit)
        if isinstance(a, dict):
            return a.keys() == b.keys() and all(same(a[k], b[k]) for k in a)
        if isinstance(a, list):
            return len(a) == len(b) and all(same(x, y) for x, y in zip(a, b))
        if isinstance(a, float) and math.isnan(a):
            return isinstance(b, float) and math.isnan(b)
        return a == b
    for r, g, x, out in zip(rads, geo, [True, True, False, True], got):
        assert same(out, analyze_motility(r, 1.0, 30.0, g, x))
    assert [o["verdict"] for o in got] == ["MOTILE", "NOT_MOTILE", "UNLOCKED_EXPOSURE", "NO_DYNAMICS"]
    assert analyze_motility_batch([], 1.0, 30.0) == []
    bad = Radial(rads[0].N, rads[0].NB, rads[0].L, rads[0].kCenter, rads[0].kNodes, rads[0].nModes, rads[0].D,
                 rads[0].P, rads[0].drift)
    with pytest.raises(DdmError, match="record 1"):
        analyze_motility_batch([rads[0], bad], [1.0, -1.0], 30.0)
