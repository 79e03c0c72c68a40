"""Fitting stage oracle (CPU): the reference's VarPro recovery KAT, the FP64 per-q LM,
and the product's pure-Python preparation against the oracle's."""
import numpy as np
import pytest

import oracle as O
from hackathon_amd import fit as F
from hackathon_amd.engine import Radial


def truth_radial(geometry, truth=dict(vbar=30, Z=4, Dd=0.3, alpha=0.6), NB=32, L=128, N=256, fps=30, kn=None):
    """Noise-free model data of test/t_kernels.mjs:283-298 (A = 1000 exp(-q^2), B = 2)."""
    kC = 4 + 3 * (np.arange(NB) + 0.5)
    D = np.zeros((NB, L))
    kern = O.isf3d if geometry == "3d" else O.isf2d_fast
    for b in range(NB):
        q = 2 * np.pi * kC[b] / N
        A, B = 1000 * np.exp(-q * q), 2
        for t in range(L):
            tau = t / fps
            E = np.exp(-truth["Dd"] * q * q * tau)
            f = E * (1 - truth["alpha"] + truth["alpha"] * kern(q * truth["vbar"] * tau, truth["Z"]))
            D[b, t] = A * (1 - f) + (B if t else 0)
    return dict(N=N, NB=NB, L=L, kCenter=kC, nModes=np.full(NB, 50), D=D.ravel(), kNodes=kn)


@pytest.mark.parametrize("geometry", ["3d", "2d"])
def test_reference_varpro_recovery(geometry):
    """t_kernels.mjs:281-308: vbar, alpha to 1e-3; Z, D to 1e-2; swimmer beats Brownian 1000x."""
    pr = O.prepare(truth_radial(geometry), 1, 30)
    Dd, rb = O.fit_brownian(pr)
    fit = O.fit_swimmer(pr, geometry, 0, Dd)
    assert abs(fit["vbar"] / 30 - 1) < 1e-3 and abs(fit["alpha"] - 0.6) < 1e-3
    assert abs(fit["Z"] / 4 - 1) < 1e-2 and abs(fit["Dd"] / 0.3 - 1) < 1e-2
    assert fit["rss"] < 1e-3 * rb


def test_lm_brownian_closed_form():
    rng = np.random.default_rng(0)
    tau = O.lag_grid(256) / 30.0
    for _ in range(20):
        q = rng.uniform(0.1, 2)
        Dt = rng.uniform(0.05, 2)
        A, B = rng.uniform(10, 1000), rng.uniform(0, 5)
        d = A * (1 - np.exp(-Dt * q * q * tau)) + B
        bd = np.array([2, 3 / (q / 30), 0.01 / (q * q * tau[-1]), 10 / (q * q / 30)])
        th, chi, it = O.lm_plain(tau, d, 1 / d.max(), [q], bd, swim=False)
        assert abs(np.exp(th[5]) / Dt - 1) < 1e-6 and abs(th[0] / A - 1) < 1e-6 and abs(th[1] - B) < 1e-6 * A


def test_lm_recovers_truth_on_conditioned_bins():
    """Per-q LM: exact recovery (chi^2 -> 0) wherever the bin's curve pins the parameters."""
    pr = O.prepare(truth_radial("3d"), 1, 30)
    b4 = O.bounds(pr)[:4]
    good = 0
    for bn in pr["bins"]:
        th, chi, it = O.lm_plain(pr["tau"], bn["d"], bn["w"], bn["qn"], b4, True, "3d")
        if chi < 1e-20:
            good += 1
            assert abs(th[2] - 0.6) < 1e-6 and abs(np.exp(th[3]) / 30 - 1) < 1e-6
            assert abs(np.exp(th[4]) / 4 - 1) < 1e-5 and abs(np.exp(th[5]) / 0.3 - 1) < 1e-5
    assert good >= 24


def drift_problem(q, swim, geometry, V=4.0, truth=dict(vbar=30, Z=4, Dd=0.3, alpha=0.6), L=128, fps=30):
    """One bin's noise-free curve with a uniform drift |V|, written by the oracle's model column
    (modelColumn, src/motility_fit.js:417-432: E = exp(-D q^2 tau) J0(q |V| tau)); A = 1000, B = 2."""
    tau = O.lag_grid(L) / fps
    c = O.model_column(tau, [q], truth["vbar"], truth["Z"], truth["Dd"], V, truth["alpha"] if swim else 0.0, swim,
                       O.GEOMETRY[geometry])
    d = 1000 * c + 2
    bd = np.array([max(2, 0.3 / (q * tau[-1])), 3 / (q / fps), 0.01 / (q * q * tau[-1]), 10 / (q * q / fps)])
    return tau, d, 1 / d.max(), bd


@pytest.mark.parametrize("swim,geometry", [(False, "3d"), (True, "3d"), (True, "2d")])
def test_lm_with_drift_recovers_truth(swim, geometry):
    """Given the drift speed, the per-q LM recovers the model column's parameters exactly (its own
    column is a separate evaluation of the same formula); without it the same data cannot be fitted."""
    for q in (0.5, 1.0, 1.6):
        tau, d, w, bd = drift_problem(q, swim, geometry)
        th, chi, it = O.lm_plain(tau, d, w, [q], bd, swim, geometry, vmag=4.0)
        assert chi < 1e-20, (q, chi)
        assert abs(np.exp(th[5]) / 0.3 - 1) < 1e-6 and abs(th[0] / 1000 - 1) < 1e-6 and abs(th[1] - 2) < 1e-5
        if swim:
            assert abs(th[2] - 0.6) < 1e-6 and abs(np.exp(th[3]) / 30 - 1) < 1e-6 and abs(np.exp(th[4]) / 4 - 1) < 1e-5
        th0, chi0, _ = O.lm_plain(tau, d, w, [q], bd, swim, geometry)
        assert chi0 > 1e6 * max(chi, 1e-30)


def test_prepare_and_lag_grid_match_oracle():
    r = truth_radial("3d", kn=np.repeat(4 + 3 * (np.arange(32) + 0.5), 4))
    po = O.prepare(r, 1, 30)
    pp = F.prepare(Radial(r["N"], r["NB"], r["L"], r["kCenter"], r["kNodes"], r["nModes"], r["D"], np.zeros(8),
                          np.zeros(1)), 1, 30)
    assert np.array_equal(po["lags"], pp.lags) and np.allclose(po["tau"], pp.tau)
    assert [b["b"] for b in po["bins"]] == [b.b for b in pp.bins]
    for a, b in zip(po["bins"], pp.bins):
        assert np.array_equal(a["d"], b.d) and a["w"] == b.w and np.allclose(a["qn"], b.qn)
    assert np.allclose(O.bounds(po)[:4], F.bounds(pp), rtol=1e-14)
    for L in (2, 11, 64, 128, 1024, 4096):
        assert np.array_equal(O.lag_grid(L), F.lag_grid(L))
