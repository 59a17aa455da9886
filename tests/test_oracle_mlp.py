"""CPU tests pinning the MLP oracle (oracle/mlp_ref.py) that the GPU training kernels are checked against: the
analytic gradient equals central finite differences of the loss, Adam's first step has the closed-form size lr."""
import numpy as np

import mlp_ref


def _setup(n=40, seed=0):
    rng = np.random.default_rng(seed)
    npar = mlp_ref.NF * mlp_ref.NH + mlp_ref.NH + mlp_ref.NH * mlp_ref.NH + 2 * mlp_ref.NH + 1
    p = rng.standard_normal(npar) * 0.3
    X = rng.standard_normal((n, mlp_ref.NF))
    y = rng.standard_normal(n)
    fmean, fstd = rng.standard_normal(mlp_ref.NF) * 0.1, 0.5 + rng.random(mlp_ref.NF)
    return p, fmean, fstd, X, y


def test_gradient_matches_finite_differences():
    p, fm, fs, X, y = _setup()
    g = mlp_ref.grad(p, fm, fs, X, y)
    rng = np.random.default_rng(1)
    idx = np.concatenate([rng.choice(len(p), 60, replace=False), [len(p) - 1]])   # every group, and b3
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p); e[i] = h
        fd = (mlp_ref.loss(p + e, fm, fs, X, y) - mlp_ref.loss(p - e, fm, fs, X, y)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), i


def test_adam_first_step():
    """With bias correction the first step moves every parameter by lr * sign(g) (|g| >> eps)."""
    p, fm, fs, X, y = _setup()
    g = mlp_ref.grad(p, fm, fs, X, y)
    p1, m, v = mlp_ref.adam(p, g, np.zeros_like(p), np.zeros_like(p), 1, 1e-3)
    big = np.abs(g) > 1e-4
    assert np.allclose((p - p1)[big], 1e-3 * np.sign(g[big]), rtol=1e-3)


def test_shift_block_is_block_jacobi_at_zero():
    D = np.diag(np.arange(1.0, 7.0))
    assert np.array_equal(mlp_ref.shift_block(D, 0.0), D)
    assert np.allclose(mlp_ref.shift_block(D, 0.1) - D, 0.1 * 3.5 * np.eye(6))
