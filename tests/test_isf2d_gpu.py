"""The in-plane swimmer ISF on the GPU in FP64 (ddm_isf2d_eval): isf2d by the Legendre quadrature of Laplace's
integral (src/motility_fit.js:120-159) and isf2dFast, the table the in-plane fits interpolate, built on the
device from the same quadrature (:161-220).

Held to: the reference's own isf2d and isf2dFast values (tests/golden/fits_reference.json, written by its
code); P_2 and P_5 in closed form (integer Z, test/t_kernels.mjs:46-51); scipy's Legendre function of real
degree (lpmv) to the quadrature's own accuracy; the small-L series and its continuity; the host table the
verdict layer uses."""
import json
import os

import numpy as np
import pytest

from hackathon_amd import drift as DR
from hackathon_amd import fit as F

pytestmark = pytest.mark.gpu
REF = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "fits_reference.json")))


def test_against_reference_values(gpu):
    a = np.array(REF["isf2dFast"])
    f = F.Fitter()
    exact = f.isf2d(a[:, 0], a[:, 1], exact=True)
    fast = f.isf2d(a[:, 0], a[:, 1], exact=False)
    print("max |exact - ref isf2d|", np.max(np.abs(exact - a[:, 3])), "max |table - ref isf2dFast|",
          np.max(np.abs(fast - a[:, 2])))
    # the device's exp/log/cos/atan2 differ from V8's in the last ulp, over up to ~170 quadrature terms
    assert np.max(np.abs(exact - a[:, 3])) < 5e-13
    assert np.max(np.abs(fast - a[:, 2])) < 5e-13


def test_integer_degree_closed_forms(gpu):
    x = np.concatenate([np.linspace(1e-3, 2, 40), np.geomspace(2, 60, 40)])
    f = F.Fitter()
    for Z, P in ((2.0, lambda c: (3 * c ** 2 - 1) / 2), (5.0, lambda c: (63 * c ** 5 - 70 * c ** 3 + 15 * c) / 8)):
        L = x / (Z + 1)
        u = 1 + L * L
        ok = -0.5 * (Z + 1) * np.log(u) >= -25
        want = P(1 / np.sqrt(u)) * u ** (-(Z + 1) / 2)
        got = f.isf2d(x, np.full_like(x, Z))
        print(Z, np.max(np.abs(got[ok] - want[ok])))
        assert np.max(np.abs(got[ok] - want[ok])) < 1e-12
        assert np.all(got[~ok] == 0)


def test_real_degree_against_scipy(gpu):
    from scipy.special import lpmv
    f = F.Fitter()
    for Z in (0.7, 1.3, 3.5, 7.2, 20.5, 61.0):
        L = np.geomspace(1e-3, 3, 30)
        x = L * (Z + 1)
        u = 1 + L * L
        ok = -0.5 * (Z + 1) * np.log(u) >= -25
        pz = f.isf2d(x, np.full_like(x, Z)) * u ** ((Z + 1) / 2)
        want = lpmv(0, Z, 1 / np.sqrt(u))
        err = np.max(np.abs(pz[ok] - want[ok]))
        print(Z, err)
        assert err < 5e-6   # the trapezoid rule's own accuracy (legendreP's comment, :113-117)


def test_small_lambda_series_and_continuity(gpu):
    f = F.Fitter()
    Z = np.array([0.5, 2.0, 9.0, 80.0])
    below = 0.99e-4 * (Z + 1)
    above = 1.01e-4 * (Z + 1)
    gb, ga = f.isf2d(below, Z), f.isf2d(above, Z)
    assert np.allclose(gb, 1 - below ** 2 * (Z + 2) / (4 * (Z + 1)), rtol=0, atol=1e-16)
    # the series and the quadrature meet within the quadrature's own accuracy (the reference's junction)
    assert np.max(np.abs(gb - ga)) < 5e-6


def test_device_table_equals_host_table(gpu):
    rng = np.random.default_rng(3)
    Z = np.exp(rng.uniform(np.log(0.3), np.log(120), 400))
    x = np.exp(rng.uniform(np.log(1e-5), np.log(80), 400))
    dev = F.Fitter().isf2d(x, Z, exact=False)
    host = np.array([DR.isf2d(xi, zi) for xi, zi in zip(x, Z)])
    assert np.max(np.abs(dev - host)) < 5e-13
