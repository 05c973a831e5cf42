"""Pins the NumPy MLP oracle of the surrogate (oracle/mlp_oracle.py):
its backward against central finite differences of its own forward, and its
bfloat16 rounding against torch's conversion."""
import numpy as np
import pytest

from oracle import mlp_oracle as M


def params(seed=0, scale=0.2):
    rng = np.random.default_rng(seed)
    lay, n = M.layout()
    return rng.normal(0.0, scale, n)


def test_layout_size():
    lay, n = M.layout()
    assert n == 14 * 256 + 256 + (M.NL - 1) * (256 * 256 + 256) + 256 * 4 + 4   # 136452 for NL = 3
    assert lay["W5"][0] == 14 * 256 + 256 + (M.NL - 1) * (256 * 256 + 256)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = params(1)
    X = rng.uniform(-1, 1, (7, 16))
    T = rng.normal(0, 1, (7, 4))
    loss, g = M.loss_grad(p, X, T, emulate_bf16=False)
    lay, _ = M.layout()
    h = 1e-6
    for name, (off, shape) in lay.items():
        for _ in range(4):
            i = off + int(rng.integers(int(np.prod(shape))))
            pp, pm = p.copy(), p.copy()
            pp[i] += h
            pm[i] -= h
            fd = (M.loss_grad(pp, X, T, False)[0] - M.loss_grad(pm, X, T, False)[0]) / (2 * h)
            assert g[i] == pytest.approx(fd, rel=1e-5, abs=1e-10), (name, i)


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(2)
    x = np.concatenate([rng.normal(0, 10, 10000), rng.uniform(-1e-3, 1e-3, 1000)]).astype(np.float32)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(M.bf16(x), ref)


def test_loss_is_mean_square():
    rng = np.random.default_rng(3)
    p = params(3)
    X = rng.uniform(-1, 1, (5, 16))
    Y, _ = M.forward(p, X, False)
    T = Y + 0.5
    loss, _ = M.loss_grad(p, X, T, False)
    assert loss == pytest.approx(0.25, rel=1e-12)
