"""Pins the MLP restatement (oracle/mlp_ref.py) that the GPU training kernels are checked against: its analytic
gradient against central finite differences, its loss against a direct evaluation, and its shapes."""
import numpy as np

from oracle import mlp_ref


def _case(seed, B=24):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.3, mlp_ref.N_PARAMS)
    X = rng.normal(0, 1, (B, 16))
    t = rng.normal(0, 1, B)
    c = rng.integers(0, 3, B)
    return p, X, t, c


def test_parameter_count():
    assert mlp_ref.N_PARAMS == 16 * 64 + 64 + 64 * 64 + 64 + 64 * 4 + 4 == 5508


def test_gradient_matches_finite_differences():
    p, X, t, c = _case(1)
    g = mlp_ref.grad(p, X, t, c)
    rng = np.random.default_rng(2)
    idx = rng.choice(mlp_ref.N_PARAMS, 60, replace=False)
    idx = np.concatenate([idx, [0, 1023, 1024, 1087, 1088, 5183, 5184, 5247, 5248, 5503, 5504, 5507]])
    fd = mlp_ref.finite_difference(p, X, t, c, idx)
    np.testing.assert_allclose(g[idx], fd, rtol=1e-5, atol=1e-8)


def test_loss_is_mse_plus_cross_entropy():
    p, X, t, c = _case(3, B=5)
    y, _ = mlp_ref.forward(p, X)
    mse = np.mean((y[:, 0] - t) ** 2)
    z = y[:, 1:]
    ce = np.mean([np.log(np.exp(z[i]).sum()) - z[i, c[i]] for i in range(5)])
    assert abs(mlp_ref.loss(p, X, t, c) - (mse + ce)) < 1e-12
