"""GPU parity of the dense reduced solve (bls_chol_*, the :cholesky branch of solve_reduced_dense!,
src/schur_linear_systems.jl:312-319) against oracle/dense_ref.py: U within 1e-12 of the oracle's factor (relative,
Frobenius), x within 1e-11 cond(S) of the oracle's solve, the failing column exact, bitwise reproducible; the lower
triangle of S is never read (it holds NaN here)."""
import numpy as np
import pytest

import dense_ref as D
from test_conditioning import _spd

pytestmark = pytest.mark.gpu


def _dev_S(S):
    """S column-major with NaN below the diagonal, as a flat device array."""
    import torch
    Sn = np.where(np.triu(np.ones_like(S)) > 0, S, np.nan)
    return torch.as_tensor(np.ascontiguousarray(Sn.T).reshape(-1)).cuda()


@pytest.mark.parametrize("n", [1, 63, 64, 65, 130, 300])
def test_dense_cholesky_matches_oracle(blsamd, n):
    import torch
    dev = blsamd.device
    rng = np.random.default_rng(n)
    cond = 1e4
    S = _spd(rng, n, cond)
    g = rng.standard_normal(n)
    ch = dev.DenseCholesky(n)
    ok, fc = ch.factor(_dev_S(S))
    assert ok and fc == 0
    U = torch.empty(n * n, dtype=torch.float64, device="cuda")
    ch.factor_copy(U)
    Ud = np.triu(U.cpu().numpy().reshape(n, n).T)
    Uo, info = D.cholesky_upper(S)
    assert info == 0
    assert np.linalg.norm(Ud - Uo) <= 1e-12 * np.linalg.norm(Uo)
    x = torch.full((n,), np.nan, dtype=torch.float64, device="cuda")
    ch.solve(torch.as_tensor(g).cuda(), x)
    xo = D.cholesky_solve(Uo, g)
    assert np.linalg.norm(x.cpu().numpy() - xo) <= 1e-11 * cond * np.linalg.norm(xo)
    # the same factor and solution again, bit for bit
    ch2 = dev.DenseCholesky(n)
    assert ch2.factor(_dev_S(S))[0]
    U2 = torch.empty(n * n, dtype=torch.float64, device="cuda")
    ch2.factor_copy(U2)
    x2 = torch.empty(n, dtype=torch.float64, device="cuda")
    ch2.solve(torch.as_tensor(g).cuda(), x2)
    assert np.array_equal(np.triu(U2.cpu().numpy().reshape(n, n).T), Ud) and torch.equal(x, x2)


def test_dense_cholesky_failing_column(blsamd):
    """A negative pivot at column 151 (the third panel), then NaN at column 70: BLS_FAILED with that column, x left as
    g by the solve; a positive definite matrix afterwards factors again."""
    import torch
    dev = blsamd.device
    rng = np.random.default_rng(2)
    n = 200
    ch = dev.DenseCholesky(n)
    S = np.diag(rng.uniform(1.0, 2.0, n))
    S[150, 150] = -1.0
    assert ch.factor(_dev_S(S)) == (False, 151) and D.cholesky_upper(S)[1] == 151
    g = torch.as_tensor(rng.standard_normal(n)).cuda()
    x = torch.empty(n, dtype=torch.float64, device="cuda")
    ch.solve(g, x)
    assert torch.equal(x, g)
    S = _spd(rng, n)
    S[69, 69] = np.nan
    assert ch.factor(_dev_S(S)) == (False, 70) and D.cholesky_upper(S)[1] == 70
    S = _spd(rng, n)
    assert ch.factor(_dev_S(S)) == (True, 0)


def test_dense_cholesky_float32_input(blsamd):
    """A Float32 S is widened and factored in Float64: the same factor and solution, bit for bit, as the Float64 path
    on the widened values."""
    import torch
    dev = blsamd.device
    rng = np.random.default_rng(7)
    n = 150
    S32 = _spd(rng, n).astype(np.float32)
    g = torch.as_tensor(rng.standard_normal(n)).cuda()
    a, b = dev.DenseCholesky(n), dev.DenseCholesky(n)
    Sd = _dev_S(S32.astype(np.float64))
    assert a.factor(Sd.to(torch.float32)) == (True, 0) and b.factor(Sd) == (True, 0)
    xa, xb = torch.empty(n, dtype=torch.float64, device="cuda"), torch.empty(n, dtype=torch.float64, device="cuda")
    a.solve(g, xa)
    b.solve(g, xb)
    assert torch.equal(xa, xb)
