"""Global-parameter fits on the GPU (ddm_fit_global): alpha, vbar, Z, D shared by a video's bins, per-bin
amplitudes by exact NNLS -- the reference's fitSwimmer / fitBrownian objective (src/motility_fit.js:450-611).

Held to: the noise-free truth of test/t_kernels.mjs:136-163 (3D and in-plane kernels); the reference's own
fitSwimmer / fitBrownian outputs on six noisy validation-matrix records (tests/golden/fits_reference.json):
the GPU's RSS is never above the reference's (1 + 1e-6), and where the swimmer share is resolved (alpha >
0.05) every parameter agrees to 1e-3; from the reference's optimum as the seed it stays there; a video's
result does not depend on the batch around it or on how its starts are scheduled."""
import json
import os

import numpy as np
import pytest

from hackathon_amd import fit as F
from hackathon_amd.engine import Radial
from tests.test_fit_oracle import truth_radial
from tests.test_verdict import oracle_radial

pytestmark = pytest.mark.gpu
REF = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "fits_reference.json")))
MODEL = {"3d": F.SWIMMER_3D, "2d": F.SWIMMER_2D}


def _prep(name):
    return F.prepare(oracle_radial(name), 1.0, 30.0)


@pytest.mark.parametrize("geometry", ["3d", "2d"])
def test_noise_free_truth(gpu, geometry):
    r = truth_radial(geometry)
    rad = Radial(r["N"], r["NB"], r["L"], r["kCenter"], np.zeros(0), r["nModes"], r["D"], np.zeros(r["NB"]), None)
    res = F.fit_global(F.Fitter(), [F.prepare(rad, 1, 30)], model=MODEL[geometry])
    print(geometry, res.alpha, res.vbar, res.Z, res.D, res.rss, res.iters, f"{res.kernel_ms:.2f} ms")
    assert res.rss[0] < 1e-9
    assert abs(res.alpha[0] - 0.6) < 1e-4 and abs(res.vbar[0] / 30 - 1) < 1e-4
    assert abs(res.Z[0] / 4 - 1) < 1e-4 and abs(res.D[0] / 0.3 - 1) < 1e-4
    A, B = res.amps[0][:, 0], res.amps[0][:, 1]
    q = 2 * np.pi * r["kCenter"] / r["N"]
    assert np.allclose(A, 1000 * np.exp(-q * q), rtol=1e-4) and np.allclose(B, 2, rtol=1e-3)


@pytest.mark.parametrize("geometry", ["3d", "2d"])
def test_against_reference_fits(gpu, geometry):
    recs = [r for r in REF["records"] if r["geometry"] == geometry]
    prs = [_prep(r["name"]) for r in recs]
    f = F.Fitter()
    res = F.fit_global(f, prs, model=MODEL[geometry])
    for i, r in enumerate(recs):
        fs, fb = r["fitSwimmer"], r["fitBrownian"]
        print(r["name"], f"alpha {res.alpha[i]:.6f}/{fs['alpha']:.6f} vbar {res.vbar[i]:.5f}/{fs['vbar']:.5f} "
              f"Z {res.Z[i]:.5f}/{fs['Z']:.5f} D {res.D[i]:.6f}/{fs['Dd']:.6f} rss {res.rss[i]:.9g}/{fs['rss']:.9g} "
              f"| Brownian D {res.brown[i, 0]:.6f}/{fb['Dd']:.6f} rss {res.brown[i, 1]:.9g}/{fb['rss']:.9g}")
        # never a worse optimum than the reference's Nelder-Mead
        assert res.rss[i] <= fs["rss"] * (1 + 1e-6), (r["name"], res.rss[i], fs["rss"])
        assert res.brown[i, 1] <= fb["rss"] * (1 + 1e-6)
        assert abs(res.brown[i, 0] / fb["Dd"] - 1) < 1e-3
        if fs["alpha"] > 0.05:   # the swimmer share is resolved: a unique optimum
            assert abs(res.alpha[i] - fs["alpha"]) < 1e-3
            for got, want in ((res.vbar[i], fs["vbar"]), (res.Z[i], fs["Z"]), (res.D[i], fs["Dd"])):
                assert abs(got / want - 1) < 1e-3, (r["name"], got, want)
        # seeded from the reference's optimum (passed to the device in FP32): stays at it
        seed = np.array([[fs["alpha"], np.log(fs["vbar"]), np.log(fs["Z"]), np.log(fs["Dd"])]])
        s1 = F.fit_global(f, [prs[i]], model=MODEL[geometry], seeds=seed)
        assert s1.rss[0] <= fs["rss"] * (1 + 1e-6)


def test_batch_and_schedule_independence(gpu):
    """A video's fit is the same alone, in a small batch (one workgroup per start) and in a large one (its
    starts in sequence in one workgroup)."""
    names = [r["name"] for r in REF["records"] if r["geometry"] == "3d"]
    prs = [_prep(n) for n in names]
    f = F.Fitter()
    small = F.fit_global(f, prs)
    big = F.fit_global(f, [prs[0]] * 180 + prs[1:])          # 184 videos x 3 starts > 512: sequential starts
    alone = F.fit_global(f, [prs[2]])
    assert np.array_equal(small.theta[0], big.theta[0]) and small.rss[0] == big.rss[0]
    assert np.array_equal(small.theta[1:], big.theta[180:]) and np.array_equal(small.rss[1:], big.rss[180:])
    assert np.array_equal(alone.theta[0], small.theta[2]) and np.array_equal(alone.amps[0], small.amps[2])
    assert np.all(big.theta[:180] == big.theta[0])


def test_global_fit_errors(gpu):
    import ctypes as C
    f = F.Fitter()
    pr = _prep("motile_3d")
    bad = _prep("motile_3d")
    bad.bins = []
    with pytest.raises(Exception, match="usable bin"):
        F.fit_global(f, [pr, bad])
    # the library's own checks: a video without bins, bad bounds
    z = np.zeros(256)
    off = np.array([0, 0], np.int32)
    nt = np.array([4], np.int32)
    bd = np.array([1.0, 2.0, 1.0, 2.0])
    b = F.GlobalFitBatch(1, 4, 1, F.SWIMMER_3D, F.SEED_GRID, 3, 10, off.ctypes.data, nt.ctypes.data, z.ctypes.data,
                         z.ctypes.data, z.ctypes.data, z.ctypes.data, bd.ctypes.data, None, None)
    out = np.zeros(16)
    it = np.zeros(4, np.int32)
    assert f._lib.ddm_fit_global(f._h, C.byref(b), out.ctypes.data, out.ctypes.data, it.ctypes.data, None, None,
                                 None) == -1
    assert b"bins" in f._lib.ddm_fitter_last_error(f._h)
    off[1] = 1
    bd[1] = 0.5
    assert f._lib.ddm_fit_global(f._h, C.byref(b), out.ctypes.data, out.ctypes.data, it.ctypes.data, None, None,
                                 None) == -1
    assert b"bounds" in f._lib.ddm_fitter_last_error(f._h)
