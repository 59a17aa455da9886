"""GPU factor_diagonal_blocks! / solve_diagonal_blocks! (src/diagonal_blocks.jl:6-61) against the oracle's loops:
Float64 factors and solves within 1e-13 relative per block, Float32 factors within 4 eps(Float32), the lowest
non-positive-definite block exact."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _blocks(rng, dims, dtype):
    out = []
    for d in dims:
        A = rng.standard_normal((d, d))
        out.append((A @ A.T + d * np.eye(d)).T.reshape(-1))
    return np.concatenate(out).astype(dtype)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_factor_and_solve(blsamd, dtype):
    import torch
    dev = blsamd.device
    rng = np.random.default_rng(7)
    dims = rng.choice([3, 6, 9, 16], size=300)
    D = _blocks(rng, dims, dtype)
    blocks = dev.DiagonalBlocks(dims)
    Dd = torch.as_tensor(D).cuda()
    ok, fb = dev.factor_diagonal_blocks_(Dd, blocks)
    Do = D.copy()
    assert ok and fb == 0 and oracle.factor_diagonal_blocks(dims, Do)
    Df = Dd.cpu().numpy()
    o = 0
    tol = 1e-13 if dtype == np.float64 else 4 * float(np.finfo(np.float32).eps)
    for d in dims:
        L, Lo = np.tril(Df[o:o + d * d].reshape(d, d).T), np.tril(Do[o:o + d * d].reshape(d, d).T)
        assert np.linalg.norm(L - Lo) <= tol * np.linalg.norm(Lo)
        o += d * d
    x = rng.standard_normal(int(dims.sum()))
    y = torch.empty(len(x), dtype=torch.float64, device="cuda")
    dev.solve_diagonal_blocks_(y, Dd, blocks, torch.as_tensor(x).cuda())
    yo = np.zeros(len(x))
    oracle.solve_diagonal_blocks(dims, yo, Df, x)    # the oracle's solve on the device's factors
    assert np.abs(y.cpu().numpy() - yo).max() <= 1e-12 * np.abs(yo).max()
    xs = torch.as_tensor(x).cuda()                   # y may be x
    dev.solve_diagonal_blocks_(xs, Dd, blocks, xs)
    assert np.array_equal(xs.cpu().numpy(), y.cpu().numpy())


def test_lowest_non_positive_definite_block(blsamd):
    import torch
    dev = blsamd.device
    rng = np.random.default_rng(8)
    dims = np.array([6] * 40)
    D = _blocks(rng, dims, np.float64)
    for b in (21, 6):   # 0-based blocks made indefinite
        D[b * 36: b * 36 + 36] = -np.eye(6).reshape(-1)
    ok, fb = dev.factor_diagonal_blocks_(torch.as_tensor(D).cuda(), dev.DiagonalBlocks(dims))
    assert not ok and fb == 7
    assert not oracle.factor_diagonal_blocks(dims, D.copy())
