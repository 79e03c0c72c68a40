"""The FP64 oracle against the reference's own known-answer tests and
independent mathematics (numpy/scipy, brute force, closed forms).

Each test names the reference check it restates (motility-ddm/test/...).
"""
import json
import math
import os
import struct

import numpy as np
import pytest

import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")


# ---------------------------------------------------------------- FFT (t_fft.mjs:23-40)

@pytest.mark.parametrize("n", [2 ** k for k in range(1, 10)])
def test_fft_vs_dft(n):
    rng = np.random.default_rng(1234 + n)
    x = rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)
    for inv in (False, True):
        ref = O.dft(x, inv)
        got = O.fft(x, inv)
        assert np.max(np.abs(got - ref)) / np.max(np.abs(ref)) < 1e-12
    # numpy's convention: forward e^{-i}, inverse unnormalised = n * ifft
    assert np.allclose(O.fft(x), np.fft.fft(x), rtol=0, atol=1e-12 * n)
    assert np.allclose(O.fft(x, True), n * np.fft.ifft(x), rtol=0, atol=1e-12 * n)
    rt = O.fft(O.fft(x), True) / n
    assert np.max(np.abs(rt - x)) < 1e-13


def test_fft_rejects_non_pow2():
    with pytest.raises(ValueError):
        O.fft(np.ones(12))


# ---------------------------------------------------------------- window (t_fft.mjs:71-77)

def test_blackman_harris_kats():
    w = O.blackman_harris(256)
    assert abs(w[0] - 6e-5) < 1e-12
    assert abs(w[128] - 1) < 1e-12
    assert abs(w[10] - w[246]) < 1e-14
    # the 4-term periodic window's DFT is nonzero only at |k| <= 3
    W = np.fft.fft(w)
    assert np.max(np.abs(W[4:-3])) < 1e-10
    a = [0.35875, 0.48829, 0.14128, 0.01168]
    assert np.allclose(W[:4].real, [256 * a[0], -128 * a[1], 128 * a[2], -128 * a[3]], atol=1e-9)


# ---------------------------------------------------------------- 2D spectrum (t_fft.mjs:42-69)

def _direct_2d(img):
    N = img.shape[0]
    NH = N // 2 + 1
    w = O.blackman_harris(N)
    v = (img - img.mean()) * w[None, :] * w[:, None]
    y = np.arange(N)
    out = np.zeros((N, NH), complex)
    for ky in range(N):
        for kx in range(NH):
            ph = np.exp(-2j * np.pi * (kx * y[None, :] + ky * y[:, None]) / N)
            out[ky, kx] = (v * ph).sum() / N
    return out


def test_spectrum2d_vs_direct_dft():
    rng = np.random.default_rng(7)
    img = rng.uniform(size=(16, 16))
    got = O.spectrum2d_half(img)
    ref = _direct_2d(img)
    assert np.max(np.abs(got - ref)) / np.max(np.abs(ref)) < 1e-12


@pytest.mark.parametrize("N", [32, 64, 384])
def test_spectrum2d_vs_numpy_rfft2(N):
    # mode m = ky*NH + kx: numpy's rfft2 with the last axis (x) as the half axis
    rng = np.random.default_rng(N)
    img = rng.uniform(size=(N, N))
    w = O.blackman_harris(N)
    ref = np.fft.rfft2((img - img.mean()) * w[None, :] * w[:, None]) / N
    got = O.spectrum2d_half(img)
    # N=384 goes through the O(n^2) direct DFT (no radix-3 in the reference), whose
    # summation error grows with n
    tol = 1e-12 if (N & (N - 1)) == 0 else 1e-11
    assert np.max(np.abs(got - ref)) / np.max(np.abs(ref)) < tol


# ---------------------------------------------------------------- WK (t_fft.mjs:79-137)

@pytest.mark.parametrize("label,drop,shuffle", [("full", 0, False), ("20% dropped", 0.2, False),
                                                ("shuffled + 15% dropped", 0.15, True)])
@pytest.mark.parametrize("T,L", [(256, 128), (100, 100), (37, 20)])
def test_wk_vs_direct(label, drop, shuffle, T, L):
    rng = np.random.default_rng(99 + T)
    hist = rng.uniform(-1, 1, (3, T)) + 1j * rng.uniform(-1, 1, (3, T))
    t = np.arange(T)
    hist[1] += 3 * np.exp(1j * 0.07 * t)
    order = np.arange(T, dtype=np.uint32)
    if shuffle:
        order = rng.permutation(T).astype(np.uint32)
    mask = (rng.uniform(size=T) >= drop).astype(np.float32)
    mask[0] = 1
    for m in range(3):
        D1, C1, P1 = O.temporal_mode(hist, m, order, mask, L)
        D2, C2, P2 = O.temporal_mode(hist, m, order, mask, L, direct=True)
        assert abs(P1 - P2) / P2 < 1e-12
        assert abs(D1[0]) / P2 < 1e-12
        assert np.max(np.abs(D1 - D2)) / P2 < 1e-11
        assert np.max(np.abs(C1 - C2)) / P2 < 1e-11
        # brute force from the definition: mean |x(t+tau) - x(t)|^2 over valid pairs
        x = hist[m][order] * mask
        for tau in (0, 1, L - 1):
            v = (mask[:T - tau] * mask[tau:]).astype(bool)
            d = np.abs(x[tau:] - x[:T - tau])[v] ** 2
            assert abs(D2[tau] - (d.mean() if v.any() else 0)) < 1e-12 * P2


def test_phase_rotation_closed_form():
    T, L, w = 256, 128, 0.05
    hist = np.exp(-1j * w * np.arange(T))[None, :]
    D, Cc, _ = O.temporal_mode(hist, 0, np.arange(T), np.ones(T), L)
    tau = np.arange(L)
    assert np.max(np.abs(D - (2 - 2 * np.cos(w * tau)))) < 1e-12
    assert np.max(np.abs(Cc - np.exp(-1j * w * tau))) < 1e-12


def test_temporal_meta_pairs():
    rng = np.random.default_rng(3)
    T, L = 50, 30
    mask = (rng.uniform(size=T) > 0.3).astype(np.float32)
    npairs, mspec = O.temporal_meta(T, L, np.arange(T), mask)
    for tau in range(L):
        assert npairs[tau] == sum(1 for t in range(T - tau) if mask[t] and mask[t + tau])
    z = np.zeros(128)
    z[:T] = mask
    assert np.allclose(mspec, np.fft.fft(z), atol=1e-5)


# ---------------------------------------------------------------- bins (src/ddm_ref.js:84-170)

def test_v8_hypot_matches_node():
    d = json.load(open(os.path.join(GOLD, "v8_hypot.json")))
    for x, y, h in d["half"]:
        assert format(struct.unpack(">Q", struct.pack(">d", O.v8_hypot(x, y)))[0], "016x") == h
    below = set(map(tuple, d["below"]))
    n = 0
    for a in range(0, 1025):
        for b in range(a % 7, 1025, 7):  # a deterministic 1/7 sample of the grid
            s = a * a + b * b
            r = math.isqrt(s)
            if r * r != s:
                continue
            n += 1
            assert (O.v8_hypot(a, b) < r) == ((a, b) in below)
    assert n > 100


GEOMS = [(256, 32, 4, 100), (128, 64, 0, 64), (256, 128, 0, 128), (512, 256, 0, 256), (64, 16, 2, 30),
         (16, 8, 0, 8)]


@pytest.mark.parametrize("N,NB,kmin,kmax", GEOMS)
def test_bins_vs_reference_vectors(N, NB, kmin, kmax):
    d = json.load(open(os.path.join(GOLD, f"bins_{N}_{NB}_{kmin}_{kmax}.json")))
    b = O.make_bins(N, NB, kmin, kmax)
    assert np.array_equal(b["nModes"], d["nModes"])
    assert np.array_equal(b["binOffsets"], d["binOffsets"])
    assert np.array_equal(b["kCenter"], d["kCenter"])
    # kNodes: window-leakage sums use libm cos/sin vs V8's; agree to ~1e-14
    assert np.max(np.abs(b["kNodes"] - np.array(d["kNodes"])) / np.maximum(1, np.abs(d["kNodes"]))) < 1e-12


@pytest.mark.parametrize("N,NB,kmin,kmax", GEOMS)
def test_bins_brute_force(N, NB, kmin, kmax):
    """Indices and counts by integer enumeration; exact hypot except where V8 rounds low."""
    b = O.make_bins(N, NB, kmin, kmax, with_nodes=False)
    NH = N // 2 + 1
    width = (kmax - kmin) / NB
    expect = [[] for _ in range(NB)]
    for m in range(N * NH):
        kx = m % NH
        kyi = m // NH
        ky = kyi if kyi < N // 2 else kyi - N
        if (kx == 0 and ky <= 0) or kx == N // 2 or ky == -N // 2:
            continue
        kb = math.floor((O.v8_hypot(kx, ky) - kmin) / width)
        if 0 <= kb < NB:
            expect[kb].append(m)
        # exact rule differs only at Pythagorean edges where V8 rounds below the integer
        ke = math.floor((math.sqrt(kx * kx + ky * ky) - kmin) / width)
        if ke != kb:
            r = math.isqrt(kx * kx + ky * ky)
            assert r * r == kx * kx + ky * ky and O.v8_hypot(kx, ky) < r
    got = [list(b["binModes"][b["binOffsets"][i]:b["binOffsets"][i + 1], 0]) for i in range(NB)]
    assert got == expect
    assert np.all(b["binModes"][:, 3] == 0)


def test_bins_unit_width_flips():
    """Unit-width N=256: V8 hypot moves some Pythagorean edge modes one bin down."""
    v8 = O.make_bins(256, 128, 0, 128, hypot_mode=0, with_nodes=False)
    ex = O.make_bins(256, 128, 0, 128, hypot_mode=1, with_nodes=False)
    assert not np.array_equal(v8["nModes"], ex["nModes"])
    assert v8["nModes"].sum() == ex["nModes"].sum()


# ---------------------------------------------------------------- special functions (t_kernels.mjs:39-105)

def test_special_function_kats():
    for x, ref in [(0, 1), (1, 0.7651976865579666), (2.404825557695773, 0), (5, -0.1775967713143383),
                   (10, -0.2459357644513483), (30, -0.0863679835810403)]:
        assert abs(O.bessel_j0(x) - ref) < 2e-8
    for x, ref in [(1, 0), (5, math.log(24)), (0.5, math.log(math.sqrt(math.pi))), (10.5, 13.940625219403763)]:
        assert abs(O.lngamma(x) - ref) < 1e-12
    for c in (1, 0.9, 0.5, 0.1):
        assert abs(O.legendre_p(2, c) - (3 * c * c - 1) / 2) < 1e-12
        assert abs(O.legendre_p(5, c) - (63 * c ** 5 - 70 * c ** 3 + 15 * c) / 8) < 1e-12


def test_special_functions_vs_scipy():
    from scipy import special
    xs = np.linspace(0, 40, 97)
    assert max(abs(O.bessel_j0(x) - special.j0(x)) for x in xs) < 2e-8
    for x in (0.3, 1.7, 4.2, 25.0, 101.5):
        assert abs(O.lngamma(x) - special.gammaln(x)) < 1e-12
    for nu in (0.5, 1.3, 7.7, 30.0):
        for c in (0.95, 0.6, 0.2):
            assert abs(O.legendre_p(nu, c) - special.lpmv(0, nu, c)) < 1e-5


def _schulz_average(f, vbar, Z):
    n = 60000
    vmax = vbar * 30
    h = vmax / n
    v = np.arange(n + 1) * h
    w = np.where((np.arange(n + 1) % 2) == 1, 4.0, 2.0)
    w[0] = w[-1] = 1
    pdf = np.array([O.schulz_pdf(x, vbar, Z) for x in v])
    return float((w * pdf * f(v)).sum() * h / 3)


@pytest.mark.parametrize("Z", [0.8, 3, 40])
def test_schulz_norm_mean(Z):
    assert abs(_schulz_average(lambda v: np.ones_like(v), 10, Z) - 1) < 1e-4
    assert abs(_schulz_average(lambda v: v, 10, Z) - 10) < 1e-3


@pytest.mark.parametrize("Z", [0.8, 2, 10, 60])
def test_isf_closed_forms_vs_quadrature(Z):
    from scipy import special
    e3 = e2 = 0
    for x in (0.01, 0.3, 1, 2.5, 5, 12, 30):
        q3 = _schulz_average(lambda v: np.where(v > 0, np.sin(x * v) / np.maximum(x * v, 1e-300), 1.0), 1, Z)
        q2 = _schulz_average(lambda v: special.j0(x * v), 1, Z)
        e3 = max(e3, abs(O.isf3d(x, Z) - q3))
        e2 = max(e2, abs(O.isf2d(x, Z) - q2))
    assert e3 < 2e-5 and e2 < 2e-5


def test_isf_series_switch():
    for Z in (1, 20):
        L = 1.0001e-4
        x = L * (Z + 1)
        assert abs(O.isf3d(x * 0.9999, Z) - O.isf3d(x, Z)) < 1e-9
        assert abs(O.isf2d(x * 0.9999, Z) - O.isf2d(x, Z)) < 1e-9
    assert abs(O.isf3d(1, 100) - math.sin(1)) < 0.005
    assert abs(O.isf2d(1, 100) - O.bessel_j0(1)) < 0.005


def test_isf2d_table_vs_exact():
    worst = 0
    for lz in np.arange(math.log(0.5), math.log(100) + 1e-9, 0.0371 * 3):
        Z = math.exp(lz)
        for lx in np.arange(-3, 3.5, 0.0137 * 5):
            x = 10 ** lx
            worst = max(worst, abs(O.isf2d_fast(x, Z) - O.isf2d(x, Z)))
    assert worst < 2e-5


def test_lag_grid():
    g = O.lag_grid(128)
    assert list(g[:10]) == list(range(1, 11)) and len(g) == 43
    assert len(O.lag_grid(1024)) == 70 and len(O.lag_grid(64)) == 34
    assert np.all(np.diff(g) > 0) and g[-1] <= 127


def test_nnls3_kkt():
    rng = np.random.default_rng(99)
    for _ in range(300):
        Cm = rng.uniform(-1, 1, (3, 12))
        d = rng.uniform(-1, 1, 12)
        G = Cm @ Cm.T
        h = Cm @ d
        a, rss = O.nnls3(G, h, d @ d)
        r = d - a @ Cm
        assert abs(r @ r - rss) < 1e-9
        g = G @ a - h
        assert np.all(a >= 0)
        assert np.all(np.where(a > 1e-12, np.abs(g) < 1e-9, g > -1e-9))


def test_split_transform_matches_direct_sums():
    """The split FP64 WK (mask and p transformed separately) against the direct
    O(T L) sums, with masks and a shuffled order, including the last lags."""
    rng = np.random.default_rng(21)
    T, L, N = 300, 300, 32
    fr = rng.integers(0, 255, (T, N, N)).astype(np.uint8)
    h = O.history(fr, T)
    bins = O.make_bins(N, 8, 0, 16, with_nodes=False)
    order = rng.permutation(T).astype(np.uint32)
    mask = (rng.uniform(size=T) > 0.2).astype(np.float32)
    Ds, Ps = O.analyze_radial(h, order, mask, bins, L, split=True)
    Dd, Pd = O.analyze_radial(h, order, mask, bins, L, direct=True)
    ok = bins["nModes"] > 0
    scale = np.max(np.abs(Dd[ok]), axis=1, keepdims=True)
    assert np.max(np.abs(Ds[ok] - Dd[ok]) / scale) < 1e-12
    assert np.max(np.abs(Ps[ok] - Pd[ok]) / Pd[ok]) < 1e-12


def test_packed_transform_loses_p_precision_at_small_power():
    """The reference's packed z = m + i p transform: with |x|^2 << 1 the last
    lags' A, B carry ~ulp x (mask energy) errors; the split form does not."""
    rng = np.random.default_rng(3)
    T, L = 1024, 1024
    x = (rng.standard_normal((4, T)) + 1j * rng.standard_normal((4, T))) * 3e-5
    h = x.astype(np.complex128)
    bins = dict(NB=1, binOffsets=np.array([0, 4], np.uint32),
                binModes=np.array([[m, 0, 0, 0] for m in range(4)], np.int32), nModes=np.array([4], np.uint32))
    order, mask = np.arange(T, dtype=np.uint32), np.ones(T, np.float32)
    Dp, _ = O.analyze_radial(h, order, mask, bins, L)
    Ds, _ = O.analyze_radial(h, order, mask, bins, L, split=True)
    Dd, _ = O.analyze_radial(h, order, mask, bins, L, direct=True)
    err_split = np.abs(Ds[0, -8:] - Dd[0, -8:]) / np.max(Dd[0])
    err_packed = np.abs(Dp[0, -8:] - Dd[0, -8:]) / np.max(Dd[0])
    assert err_split.max() < 1e-11 and err_packed.max() > 100 * err_split.max()


def test_drift_sums_pure_phase_ramp():
    """A mode whose series is exp(i (k.V) t) drifts by the same phase every step: z has phase k.V,
    so Rx / Sxx recovers Vx when only kx modes are present (closed form of the least-squares sums)."""
    T, L = 64, 16
    Vx = 0.05
    modes = [(0, kx, 0) for kx in (1, 2, 3, 4)]
    h = np.array([np.exp(1j * kx * Vx * np.arange(T)) for _, kx, _ in modes])
    bins = dict(NB=1, binOffsets=np.array([0, 4], np.uint32),
                binModes=np.array([[i, kx, ky, 0] for i, (_, kx, ky) in enumerate(modes)], np.int32))
    dr = O.radial_drift(h, np.arange(T, dtype=np.uint32), np.ones(T, np.float32), bins, L)
    Sw, Sxx, Rx, count = dr[0, :L - 1, 0], dr[0, :L - 1, 1], dr[0, :L - 1, 4], dr[0, :L - 1, 7]
    assert np.allclose(Rx / Sxx, Vx, rtol=1e-10) and np.all(count == 4) and np.all(Sw > 0)
    assert np.all(dr[0, L - 1] == 0)
