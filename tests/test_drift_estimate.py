"""estimate_drift: the scenarios of the reference's test/t_drift_phase.mjs.

1. V recovered from synthetic per-mode C(q, tau) with noise, at several speeds
   and angles; 2. incoherent phases give no flow; 3. swimmer-ISF gating uses
   exactly the lag pairs with |g| < 0.2 and removes the early-lag swimmer bias
   (1-3 on the CPU: per-mode C -> drift sums by the oracle's radialFromFull);
4. a translating image through the GPU pipeline (frames -> spectra -> temporal
   -> radial drift sums) returns +V (sign convention end to end).
The host ISF kernels of the gating are checked against the oracle's.
"""
import math

import numpy as np
import pytest

import oracle as O
from hackathon_amd import drift as DR
from hackathon_amd.engine import DdmEngine, Radial
from hackathon_amd.fit import prepare

N, NB, L, DX, FPS = 64, 16, 64, 1.0, 30.0
DT = 1 / FPS


def q_of(k):
    return 2 * math.pi * np.asarray(k, dtype=np.float64) / (N * DX)


def synth_radial(gen):
    """Per-mode C from gen(kx, ky, t) (arrays), D = 2 (1 - Re C), P = 1, reduced as radialFromFull."""
    bins = O.make_bins(N, NB, 2, 30)
    M = N * (N // 2 + 1)
    D = np.zeros((M, L))
    Cm = np.zeros((M, L), dtype=np.complex128)
    P = np.ones(M)
    bm = bins["binModes"]
    m, kx, ky = bm[:, 0], bm[:, 1].astype(np.float64), bm[:, 2].astype(np.float64)
    t = np.arange(L, dtype=np.float64)
    c = gen(kx[:, None], ky[:, None], t[None, :])
    Cm[m] = c
    D[m] = 2 * (1 - c.real)
    r = O.radial_from_full(D, Cm, P, bins, L)
    return Radial(r["N"], r["NB"], r["L"], r["kCenter"], r["kNodes"], r["nModes"], r["D"], r["P"], r["drift"])


def test_isf_kernels_match_oracle():
    for Z in (0.7, 1.5, 4.0, 12.0, 60.0):
        for x in (0.0, 1e-5, 0.03, 0.4, 1.0, 3.7, 11.0, 40.0):
            assert DR.isf3d(x, Z) == pytest.approx(O.isf3d(x, Z), rel=1e-12, abs=1e-15)
            # the in-plane kernel is the tabulated one of the fits (isf2dFast): its FP64 oracle to rounding,
            # the exact Legendre form to the table's own 2e-5
            assert DR.isf2d(x, Z) == pytest.approx(O.isf2d_fast(x, Z), rel=1e-10, abs=1e-13)
            assert abs(DR.isf2d(x, Z) - O.isf2d(x, Z)) < 2e-5


@pytest.mark.parametrize("speed,deg", [(6, -34), (2.5, 90), (12, 180), (4, 225)])
def test_recovery(speed, deg):
    rng = np.random.default_rng(int(speed * 100 + deg))
    Vx, Vy = speed * math.cos(math.radians(deg)), speed * math.sin(math.radians(deg))

    def gen(kx, ky, t):
        qx, qy = q_of(kx), q_of(ky)
        ph = -(qx * Vx + qy * Vy) * t * DT
        amp = np.exp(-0.2 * (qx ** 2 + qy ** 2) * t * DT)
        shape = np.broadcast(ph, amp).shape
        return (amp * np.cos(ph) + 0.02 * rng.standard_normal(shape)) + 1j * (
            amp * np.sin(ph) + 0.02 * rng.standard_normal(shape))

    r = synth_radial(gen)
    d = DR.estimate_drift(r, prepare(r, DX, FPS), DX)
    ang = abs(d.angleDeg - deg) % 360
    ang = 360 - ang if ang > 180 else ang
    assert abs(d.mag - speed) / speed < 0.01, d
    assert ang < 0.5, d
    assert d.significant and d.sigma < 0.1 * speed, d


def test_no_drift():
    rng = np.random.default_rng(77)

    def gen(kx, ky, t):
        shape = np.broadcast(kx, t).shape
        return 0.3 * rng.standard_normal(shape) + 0.3j * rng.standard_normal(shape)

    r = synth_radial(gen)
    d = DR.estimate_drift(r, prepare(r, DX, FPS), DX)
    assert d.mag < DR.FLOW_MIN_UM_S, d


def test_swimmer_gating():
    vbar, Z = 40.0, 4.0
    V, U = (5.0, 0.0), (-25.0, 0.0)        # passive drift, coherent swimmer contaminant
    isf = np.vectorize(O.isf3d)

    def gen(kx, ky, t):
        qx, qy = q_of(kx), q_of(ky)
        g = isf(np.hypot(qx, qy) * vbar * t * DT, Z)
        pv = -(qx * V[0] + qy * V[1]) * t * DT
        pu = -(qx * U[0] + qy * U[1]) * t * DT
        return (np.cos(pv) + 2 * g * np.cos(pu)) + 1j * (np.sin(pv) + 2 * g * np.sin(pu))

    r = synth_radial(gen)
    pr = prepare(r, DX, FPS)
    expect = sum(1 for b in pr.bins for t in range(1, L - 1)
                 if abs(O.isf3d(b.q * vbar * t * DT, Z)) < DR.DRIFT_ISF_MAX)
    gated = DR.estimate_drift(r, pr, DX, swimmer=(vbar, Z))
    open_ = DR.estimate_drift(r, pr, DX)
    assert gated.pairs == expect
    eg = math.hypot(gated.Vx - V[0], gated.Vy - V[1])
    eo = math.hypot(open_.Vx - V[0], open_.Vy - V[1])
    assert eg < 0.15 * math.hypot(*V) and eg < 0.5 * eo, (eg, eo)


def test_requires_drift_sums():
    r = synth_radial(lambda kx, ky, t: np.ones(np.broadcast(kx, t).shape, dtype=np.complex128))
    r.drift = None
    with pytest.raises(ValueError):
        DR.estimate_drift(r, prepare(r, DX, FPS), DX)


@pytest.mark.gpu
@pytest.mark.parametrize("sx,sy", [(1, 0), (0, 1), (-1, 1)])
def test_translation_end_to_end(gpu, sx, sy):
    T = 64
    rng = np.random.default_rng(5)
    base = rng.uniform(size=(N, N))
    sm = sum(np.roll(np.roll(base, j, axis=0), i, axis=1) for j in (-1, 0, 1) for i in (-1, 0, 1)) / 9
    yy, xx = np.mgrid[0:N, 0:N]
    frames = np.stack([sm[(yy - sy * t) % N, (xx - sx * t) % N] for t in range(T)]).astype(np.float32)
    with DdmEngine.create(N=N, T=T, L=32, NB=NB, kmin=2, kmax=30) as e:
        e.ingestFrames(frames, 0)
        e.computeTemporal(np.arange(T, dtype=np.uint32), np.ones(T, np.float32))
        r = e.readRadial(drift=True)
    d = DR.estimate_drift(r, prepare(r, DX, FPS), DX)
    Vx, Vy = sx * DX * FPS, sy * DX * FPS
    err = math.hypot(d.Vx - Vx, d.Vy - Vy) / math.hypot(Vx, Vy)
    print(f"translation ({sx},{sy}): V = ({d.Vx:.3f}, {d.Vy:.3f}) um/s, sigma {d.sigma:.3g}, rel err {err:.2e}")
    assert err < 0.1
