"""CPU checks of the generator restatement: Philox known answers, draw moments."""
import math

import numpy as np

from oracle import synth_oracle as SO


def test_philox_known_answers():
    # Random123 kat_vectors, philox4x32 with 10 rounds
    assert SO.philox4x32_10([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    f = 0xFFFFFFFF
    assert SO.philox4x32_10([f, f, f, f], [f, f]) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert SO.philox4x32_10([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_uniform_range_and_moments():
    g = SO.Stream(7, 3, 9, 12345)
    u = np.array([g.uniform() for _ in range(20000)])
    assert u.min() > 0 and u.max() < 1
    assert abs(u.mean() - 0.5) < 4 * math.sqrt(1 / 12 / u.size)
    assert abs(u.var() - 1 / 12) < 0.004


def test_poisson_moments_both_branches():
    for lam in (0.7, 9.5, 10.0, 55.0, 800.0):
        k = np.array([SO.poisson(SO.Stream(i, 0, SO.S_NOISE, 99), lam) for i in range(6000)])
        assert np.all(k == np.floor(k)) and k.min() >= 0
        assert abs(k.mean() - lam) < 5 * math.sqrt(lam / k.size), (lam, k.mean())
        assert abs(k.var() / lam - 1) < 5 * math.sqrt((lam + 2 * lam * lam) / k.size) / lam + 0.01, (lam, k.var())
    k = np.array([SO.poisson(SO.Stream(i, 1, SO.S_NOISE, 5), 0.7) for i in range(6000)])
    assert abs(np.mean(k == 0) - math.exp(-0.7)) < 0.03


def test_gamma_moments():
    Z, vbar = 3.0, 2.0
    s = np.array([SO.gamma_mt(SO.Stream(i, 0, SO.S_INIT, 1), Z + 1, vbar / (Z + 1)) for i in range(8000)])
    assert abs(s.mean() / vbar - 1) < 0.02
    assert abs(s.var() / (vbar * vbar / (Z + 1)) - 1) < 0.06


def test_intensity_single_spot_closed_form():
    lam = SO.intensity(32, [(10.25, 20.5)], 1.5, 100.0, 20.0)
    assert abs(lam[20, 10] - (20 + 100 * math.exp(-0.0625 / 4.5 - 0.25 / 4.5))) < 1e-12
    assert lam[0, 0] == 20.0
    # periodic wrap: a spot at the corner lights all four corners
    lam = SO.intensity(32, [(0.0, 0.0)], 1.5, 100.0, 0.0)
    assert lam[0, 0] == 100.0 and abs(lam[31, 31] - 100 * math.exp(-2 / 4.5)) < 1e-12
