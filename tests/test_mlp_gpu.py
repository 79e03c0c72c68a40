"""MLP kernels (BF16 MFMA GEMMs, GELU, MSE, backward, Adam) against the FP64 oracle."""
import numpy as np
import pytest

from oracle import mlp_oracle as MO
from hackathon_amd import mlp as MLP
from hackathon_amd.engine import DdmError

pytestmark = pytest.mark.gpu


def _problem(M, seed=0):
    rng = np.random.default_rng(seed)
    p = MLP.init_params(seed)
    p[MLP.NPARAMS - MLP.OUT:] = (rng.standard_normal(MLP.OUT) * 0.1).astype(np.float32)
    X = rng.standard_normal((M, MLP.IN)).astype(np.float32)
    Tg = rng.standard_normal((M, MLP.OUT)).astype(np.float32)
    wt = rng.uniform(0.5, 2, MLP.OUT).astype(np.float32)
    return p, X, Tg, wt


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("M", [1, 37, 64, 200])
def test_forward_matches_oracle(gpu, M):
    p, X, _, _ = _problem(M)
    net = MLP.Mlp(max_batch=256, params=p)
    y = net.forward(X)
    yb, _ = MO.forward(p, X, bf16=True)
    y64, _ = MO.forward(p, X)
    eb, e64 = _rel(y, yb), _rel(y, y64)
    print(f"M={M} rel err vs bf16-rounded oracle {eb:.2e}, vs exact FP64 {e64:.2e}")
    # Same roundings, FP32 vs FP64 sums: rows agree to FP32 accuracy, except the
    # rare row where an FP32 hidden value lands on the other side of a BF16
    # rounding tie than its FP64 value (an FP32 CPU restatement of the forward
    # pass shows the same few rows at M=200).  The BF16 format itself: loose.
    row = np.linalg.norm(y - yb, axis=1) / np.linalg.norm(yb, axis=1)
    assert np.median(row) < 1e-5
    assert np.mean(row > 1e-5) < 0.2 and np.max(row) < 5e-3
    assert eb < 1e-3
    assert np.max(np.abs(y - yb)) < 3e-3 * np.max(np.abs(yb))
    assert e64 < 3e-2


def test_gradients_match_oracle(gpu):
    M = 96
    p, X, Tg, wt = _problem(M, 1)
    net = MLP.Mlp(max_batch=128, params=p)
    loss, g = net.loss_grad(X, Tg, wt)
    lo, go = MO.loss_grad(p, X, Tg, wt, bf16=True)
    assert abs(loss / lo - 1) < 1e-4
    o = 0
    for li, (i, n) in enumerate(MO.LAYERS):
        gw, gwo = g[o:o + i * n], go[o:o + i * n]
        gb, gbo = g[o + i * n:o + i * n + n], go[o + i * n:o + i * n + n]
        print(f"layer {li}: dW rel {_rel(gw, gwo):.2e}  db rel {_rel(gb, gbo):.2e}")
        assert _rel(gw, gwo) < 4e-3 and _rel(gb, gbo) < 4e-3
        o += i * n + n
    # and against the exact FP64 network: BF16 level
    _, g64 = MO.loss_grad(p, X, Tg, wt)
    assert _rel(g, g64) < 5e-2


def test_adam_step_matches_oracle(gpu):
    M = 64
    p, X, Tg, wt = _problem(M, 2)
    net = MLP.Mlp(max_batch=64, params=p)
    _, g1 = net.loss_grad(X, Tg, wt)
    lr = 1e-3
    m = np.zeros(p.size)
    v = np.zeros(p.size)
    pe = p.astype(np.float64)
    net.train_step(X, Tg, wt, lr)
    pe, m, v = MO.adam_step(pe, g1.astype(np.float64), m, v, 1, lr)
    p1 = net.get_params()
    assert np.max(np.abs(p1 - pe) - (1e-4 * lr + 2e-7 * np.abs(pe))) <= 0
    # second step on a smaller batch: gradient at the new parameters, moments carried
    _, g2 = net.loss_grad(X[:17], Tg[:17], wt)
    net.train_step(X[:17], Tg[:17], wt, lr)
    pe, m, v = MO.adam_step(p1.astype(np.float64), g2.astype(np.float64), m, v, 2, lr)
    p2 = net.get_params()
    assert np.max(np.abs(p2 - pe) - (1e-3 * lr + 2e-7 * np.abs(pe))) <= 0
    # the BF16 compute copy follows the master weights
    yb, _ = MO.forward(p2, X[:8], bf16=True)
    assert _rel(net.forward(X[:8]), yb) < 2e-4


def test_training_fits_a_linear_teacher(gpu):
    rng = np.random.default_rng(4)
    Wt = rng.standard_normal((MLP.IN, MLP.OUT)) / np.sqrt(MLP.IN)
    X = rng.standard_normal((2048, MLP.IN)).astype(np.float32)
    Y = (X @ Wt).astype(np.float32)
    wt = np.ones(MLP.OUT, np.float32)
    net = MLP.Mlp(max_batch=256)
    first = None
    for ep in range(40):
        for s in range(0, 2048, 256):
            l = net.train_step(X[s:s + 256], Y[s:s + 256], wt, 1e-3)
            first = first if first is not None else l
    final = 0.5 * np.mean(np.sum((net.forward(X) - Y) ** 2, axis=1))
    print("loss first", first, "final", final)
    assert final < 0.1 * first


def test_batch_limits(gpu):
    net = MLP.Mlp(max_batch=16)
    X = np.zeros((17, MLP.IN), np.float32)
    with pytest.raises(DdmError):
        net.train_step(X, np.zeros((17, MLP.OUT), np.float32), np.ones(MLP.OUT, np.float32), 1e-3)
    with pytest.raises(DdmError):
        net.loss_grad(X[:4], np.zeros((5, MLP.OUT), np.float32), np.ones(MLP.OUT, np.float32))
    assert net.forward(X).shape == (17, MLP.OUT)     # inference chunks by max_batch
