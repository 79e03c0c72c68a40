"""CPU checks of the FP64 MLP oracle: BF16 rounding, back-propagation against
central finite differences, Adam's first-step closed form."""
import numpy as np

from oracle import mlp_oracle as MO
from hackathon_amd import mlp as MLP


def _small_problem(seed=0, M=3):
    rng = np.random.default_rng(seed)
    p = MLP.init_params(seed).astype(np.float64)
    p[MLP.NPARAMS - MO.OUT:] = rng.standard_normal(MO.OUT) * 0.1   # non-zero output bias
    X = rng.standard_normal((M, MO.IN))
    Tg = rng.standard_normal((M, MO.OUT))
    wt = rng.uniform(0.5, 2, MO.OUT)
    return p, X, Tg, wt


def test_layout_constants_agree():
    assert MO.NPARAMS == MLP.NPARAMS == 542532
    assert MO.LAYERS == MLP.LAYERS


def test_bf16_round_known_values():
    x = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 2 ** -8 + 2 ** -20, -2.5, 0.0, 3.0e38, 1e-40],
                 dtype=np.float32)
    got = MO.bf16_round(x)
    # 1+2^-8 is the tie between 1 and 1+2^-7: even mantissa -> 1; 1+3*2^-8 ties to 1+2^-6 (even)
    want = [1.0, 1.0, 1 + 2 ** -6, 1 + 2 ** -7, -2.5, 0.0, float(np.float32(3.0e38)) , None]
    for g, w in zip(got[:6], want[:6]):
        assert g == w, (g, w)
    assert abs(got[6] / 3.0e38 - 1) < 2 ** -8
    # every result is representable in 8 significant bits
    u = got.astype(np.float32).view(np.uint32)
    assert np.all((u & 0xFFFF) == 0)


def test_gelu_derivative_matches_difference():
    z = np.linspace(-6, 6, 61)
    h = 1e-6
    fd = (MO.gelu(z + h) - MO.gelu(z - h)) / (2 * h)
    assert np.max(np.abs(fd - MO.gelu_grad(z))) < 1e-8


def test_backprop_matches_finite_differences():
    p, X, Tg, wt = _small_problem()
    loss, g = MO.loss_grad(p, X, Tg, wt)
    rng = np.random.default_rng(5)
    # sample coordinates in every weight and bias block
    o, idx = 0, []
    for i, n in MO.LAYERS:
        idx += list(o + rng.integers(0, i * n, 6)) + list(o + i * n + rng.integers(0, n, 3))
        o += i * n + n
    h = 1e-5
    for k in idx:
        pp, pm = p.copy(), p.copy()
        pp[k] += h
        pm[k] -= h
        fd = (MO.loss_grad(pp, X, Tg, wt)[0] - MO.loss_grad(pm, X, Tg, wt)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g[k])) + 1e-8, (k, fd, g[k])


def test_backprop_directional_derivative():
    p, X, Tg, wt = _small_problem(1, M=5)
    _, g = MO.loss_grad(p, X, Tg, wt)
    d = np.random.default_rng(2).standard_normal(p.size)
    h = 1e-6
    fd = (MO.loss_grad(p + h * d, X, Tg, wt)[0] - MO.loss_grad(p - h * d, X, Tg, wt)[0]) / (2 * h)
    assert abs(fd - g @ d) < 1e-6 * abs(g @ d)


def test_bf16_mode_is_close_to_exact():
    p, X, Tg, wt = _small_problem(3, M=4)
    y0, _ = MO.forward(p, X)
    y1, _ = MO.forward(p, X, bf16=True)
    rel = np.max(np.abs(y1 - y0)) / np.max(np.abs(y0))
    assert 1e-5 < rel < 5e-2


def test_adam_first_step_is_signed_lr():
    g = np.array([1e-3, -2.0, 5e2])
    p, m, v = MO.adam_step(np.zeros(3), g, np.zeros(3), np.zeros(3), 1, 0.01)
    assert np.allclose(p, -0.01 * np.sign(g), rtol=1e-4)
