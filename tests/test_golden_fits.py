"""The fits, the in-plane ISF table and the movie simulator against the reference's own outputs.

tests/golden/fits_reference.json holds what the reference's code returned (make_golden_fits.py /
make_golden_fits.mjs ran its src/motility_fit.js and test/sim.mjs once in the build container) on:
  - the FP64 oracle pipeline's Radial records of six validation-matrix scenarios (noisy, with drift sums),
    rebuilt here by the same oracle code: fitBrownian, fitSwimmer, estimateDrift, analyzeMotility;
  - a grid of isf2dFast (the tabulated in-plane kernel) and isf2d (its exact Legendre form);
  - two small movies of the simulator (every population, drift, dropped frames).
The oracle's restatements and the library's verdict layer must reproduce them: fitted parameters within
1e-3 relative, the same verdicts, table values to rounding, movies byte for byte.
"""
import base64
import json
import os

import numpy as np
import pytest

import oracle as O
from hackathon_amd import drift as DR
from hackathon_amd.verdict import analyze_motility
from tests.test_verdict import oracle_prepare, oracle_radial

REF = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "fits_reference.json")))
NAMES = [r["name"] for r in REF["records"]]


def _rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def test_isf2d_table_values():
    a = np.array(REF["isf2dFast"])
    oracle = np.array([O.isf2d_fast(x, Z) for x, Z, _, _ in a])
    lib = DR.isf2d(a[:, 0].copy(), 1.0)   # shape check of the vector form
    assert lib.shape == (len(a),)
    lib = np.array([DR.isf2d(x, Z) for x, Z, _, _ in a])
    exact = np.array([O.isf2d(x, Z) for x, Z, _, _ in a])
    assert np.max(np.abs(oracle - a[:, 2])) < 1e-13
    assert np.max(np.abs(lib - a[:, 2])) < 1e-13
    assert np.max(np.abs(exact - a[:, 3])) < 1e-14
    # the table's own accuracy bound (test/t_kernels.mjs:85-102), on the reference's values
    assert np.max(np.abs(a[:, 2] - a[:, 3])) < 2e-5


@pytest.mark.parametrize("k", range(len(REF["sims"])))
def test_simulator_movies_byte_for_byte(k):
    s = REF["sims"][k]
    fr, valid, mean_speed = O.simulate(s["params"])
    want = np.frombuffer(base64.b64decode(s["frames"]), np.uint8).reshape(fr.shape)
    assert np.array_equal(fr, want)
    assert valid.astype(int).tolist() == s["valid"]
    assert mean_speed == pytest.approx(s["meanSpeed"], rel=1e-14)


@pytest.mark.parametrize("name", NAMES)
def test_oracle_fits_match_reference(name):
    ref = next(r for r in REF["records"] if r["name"] == name)
    pr = oracle_prepare(oracle_radial(name))
    assert [b["b"] for b in pr["bins"]] == ref["bins"]
    Dd, rss = O.fit_brownian(pr)
    assert _rel(Dd, ref["fitBrownian"]["Dd"]) < 1e-9 and _rel(rss, ref["fitBrownian"]["rss"]) < 1e-9
    sw = O.fit_swimmer(pr, ref["geometry"], 0, Dd)
    fs = ref["fitSwimmer"]
    for k in ("vbar", "Z", "Dd"):
        assert _rel(sw[k], fs[k]) < 1e-3, (k, sw[k], fs[k])
    assert abs(sw["alpha"] - fs["alpha"]) < 1e-3 * max(fs["alpha"], 1e-3)
    assert _rel(sw["rss"], fs["rss"]) < 1e-6


@pytest.mark.parametrize("name", NAMES)
def test_verdict_layer_matches_reference(name):
    """The library's analyzeMotility (host C++) on the same record as the reference's: the same verdict and
    gates, every fitted parameter within 1e-3 relative, the drift estimate likewise."""
    ref = next(r for r in REF["records"] if r["name"] == name)
    am = ref["analyzeMotility"]
    res = analyze_motility(oracle_radial(name), 1.0, 30.0, ref["geometry"])
    assert res["verdict"] == am["verdict"] and res["physicsVerdict"] == am["physicsVerdict"]
    assert [(g["id"], g["pass"]) for g in res["gates"]] == [(g["id"], g["pass"]) for g in am["gates"]]
    assert _rel(res["TSI"], am["TSI"]) < 1e-9
    for k in ("vbar", "Z", "Ddiff", "DdiffBrownian"):
        assert _rel(res[k], am[k]) < 1e-3, (k, res[k], am[k])
    assert abs(res["alpha"] - am["alpha"]) < 1e-3 * max(am["alpha"], 1e-3)
    assert abs(res["dBIC"] - am["dBIC"]) < 1e-3 * max(abs(am["dBIC"]), 1.0)
    assert res["drift"]["significant"] == am["drift"]["significant"]
    assert abs(res["drift"]["mag"] - am["drift"]["mag"]) < 1e-3 * max(am["drift"]["mag"], 1e-2)
    for part in ("swimmer", "brownian"):
        for k, v in am[part].items():
            got = res["fit"][part]["Dd" if k == "Dd" else k]
            assert abs(got - v) <= 1e-3 * max(abs(v), 1e-3), (part, k, got, v)


@pytest.mark.parametrize("name", NAMES)
def test_drift_estimate_matches_reference(name):
    """estimateDrift (the library routine the facades call) with the reference's swimmer gating choice."""
    ref = next(r for r in REF["records"] if r["name"] == name)
    r = oracle_radial(name)
    from hackathon_amd.fit import prepare
    fs = ref["fitSwimmer"]
    d = DR.estimate_drift(r, prepare(r, 1.0, 30.0), 1.0, ref["geometry"],
                          swimmer=(fs["vbar"], fs["Z"]) if fs["alpha"] > 0.01 else None)
    want = ref["estimateDrift"]
    assert d.pairs == want["pairs"] and d.samples == want["samples"] and d.significant == want["significant"]
    for k in ("Vx", "Vy", "mag", "sigma"):
        assert abs(getattr(d, k) - want[k]) <= 1e-6 * max(abs(want[k]), 1e-6), (k, getattr(d, k), want[k])
