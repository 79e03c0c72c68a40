"""ddm_fit_device_noalloc (the allocation-free variant of ddm_fit) against ddm_fit, through the C ABI.

Device-resident inputs and outputs; two calls in a row on the same GPU scratch, the second with fewer fits
(and a different model), each bit-identical to ddm_fit on the same problems.  A fit that fails ddm_fit's
checks comes back as NaN / -1 without disturbing its neighbours; an oversized batch is refused."""
import ctypes as C

import numpy as np
import pytest

import oracle as O
from hackathon_amd import fit as F
from hackathon_amd.engine import load_library
from tests.test_fit_oracle import truth_radial

pytestmark = pytest.mark.gpu


def _lib():
    L = F._declare(load_library())
    L.ddm_fit_device_noalloc.argtypes = [C.c_void_p, C.POINTER(F.FitBatch), C.c_void_p, C.c_void_p, C.c_void_p,
                                         C.c_int32]
    L.ddm_fit_device_noalloc.restype = C.c_int
    return L


def _problems(geometry, noise, seed):
    r = truth_radial(geometry)
    rng = np.random.default_rng(seed)
    r["D"] = r["D"] * (1 + noise * rng.standard_normal(r["D"].shape))
    pr = O.prepare(r, 1, 30)
    b4 = O.bounds(pr)[:4]
    n = len(pr["bins"])
    tau = np.tile(pr["tau"], (n, 1))
    data = np.array([b["d"] for b in pr["bins"]])
    w = np.array([b["w"] for b in pr["bins"]])
    qn = np.array([b["qn"] for b in pr["bins"]])
    return tau, data, w, qn, np.tile(b4, (n, 1))


def _device_fit(L, fitter, tau, data, w, qn, bd, model, ntau=None):
    import torch
    n, m = tau.shape
    dv = lambda a, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(a)).to(dtype=dt, device="cuda")
    nt = dv(np.full(n, m) if ntau is None else ntau, torch.int32)
    arrs = [dv(tau), dv(data), dv(w), dv(qn), dv(bd)]
    th = torch.full((n, 6), -7.0, dtype=torch.float64, device="cuda")
    chi = torch.full((n,), -7.0, dtype=torch.float64, device="cuda")
    it = torch.full((n,), -7, dtype=torch.int32, device="cuda")
    b = F.FitBatch(n, m, qn.shape[1], model, F.SEED_GRID, 3, 400, nt.data_ptr(), *[a.data_ptr() for a in arrs],
                   None, None)
    rc = L.ddm_fit_device_noalloc(fitter._h, C.byref(b), th.data_ptr(), chi.data_ptr(), it.data_ptr(), 1)
    assert rc == 0, L.ddm_fitter_last_error(fitter._h).decode()
    return th.cpu().numpy(), chi.cpu().numpy(), it.cpu().numpy()


def _same(a, ref):
    th, chi, it = a
    assert np.array_equal(th, ref.theta) and np.array_equal(chi, ref.chi2) and np.array_equal(it, ref.iters)


def test_noalloc_matches_ddm_fit_twice(gpu):
    L = _lib()
    fitter = F.Fitter()
    # call 1: 3D swimmer, noisy, 29 fits
    tau, data, w, qn, bd = _problems("3d", 0.003, 1)
    print(f"no-alloc call 1: {len(data)} fits, 3D swimmer", flush=True)
    ref = F.Fitter().fit(list(tau), list(data), w, qn, bd, model=F.SWIMMER_3D)
    _same(_device_fit(L, fitter, tau, data, w, qn, bd, F.SWIMMER_3D), ref)
    # call 2 on the same scratch: fewer fits, the in-plane table kernel
    tau2, data2, w2, qn2, bd2 = [a[:11] for a in _problems("2d", 0.003, 2)]
    print(f"no-alloc call 2: {len(data2)} fits, in-plane swimmer", flush=True)
    ref2 = F.Fitter().fit(list(tau2), list(data2), w2, qn2, bd2, model=F.SWIMMER_2D)
    _same(_device_fit(L, fitter, tau2, data2, w2, qn2, bd2, F.SWIMMER_2D), ref2)
    fitter.close()


def test_noalloc_checks(gpu):
    L = _lib()
    fitter = F.Fitter()
    tau, data, w, qn, bd = [a[:6] for a in _problems("3d", 0.0, 3)]
    ntau = np.full(6, tau.shape[1])
    ntau[2] = 0                      # ddm_fit refuses this fit; here it alone comes back NaN / -1
    bd = bd.copy()
    bd[4, 1] = bd[4, 0] / 2          # vmax < vmin
    print("no-alloc checks: invalid fits 2 and 4", flush=True)
    th, chi, it = _device_fit(L, fitter, tau, data, w, qn, bd, F.BROWNIAN, ntau=ntau)
    ok = [0, 1, 3, 5]
    ref = F.Fitter().fit([tau[i] for i in ok], [data[i] for i in ok], w[ok], qn[ok], bd[ok], model=F.BROWNIAN)
    assert np.array_equal(th[ok], ref.theta) and np.array_equal(chi[ok], ref.chi2) and np.array_equal(it[ok], ref.iters)
    for i in (2, 4):
        assert np.all(np.isnan(th[i])) and np.isnan(chi[i]) and it[i] == -1
    b = F.FitBatch(16385, 4, 1, F.BROWNIAN, F.SEED_GRID, 1, 10, *([1] * 6), None, None)
    assert L.ddm_fit_device_noalloc(fitter._h, C.byref(b), 1, 1, 1, 1) == -1
    assert b"nfits" in L.ddm_fitter_last_error(fitter._h)
    fitter.close()
