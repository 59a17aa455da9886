"""GPU parity of the dense reduced solve's Bunch-Kaufman fallback (bls_ldl_*: the :ldl branch of
solve_reduced_dense!, src/schur_linear_systems.jl:312-319; LAPACK dsytf2 / dsytrs on the lower triangle) against the
oracle (oracle/dense_ref.py, itself pinned against scipy's LAPACK dsytrf): ipiv equal and the factor within 1e-12
relative for matrices of up to 60 rows; for larger ones the inertia against numpy's eigenvalues and the solve
against numpy.linalg.solve (relative residual within 1e-12 of ||S|| ||x||); a zero pivot reported at the oracle's
column."""
import numpy as np
import pytest

import dense_ref as D

pytestmark = pytest.mark.gpu


def _sym(rng, n, kind):
    A = rng.standard_normal((n, n))
    A = A + A.T
    if kind == "smalldiag":      # small diagonal: 2 x 2 pivots and interchanges
        A[np.diag_indices(n)] *= 1e-3
    elif kind == "zeroblocks" and n > 3:
        A[:2, :2] = 0.0
        A[n // 2:n // 2 + 2, n // 2:n // 2 + 2] = 0.0
    elif kind == "spd":
        A = A @ A.T / n + np.eye(n)
    return A


def _device(blsamd, A):
    import torch
    dev = blsamd.device
    n = A.shape[0]
    ldl = dev.DenseLDL(n)
    S = torch.as_tensor(np.asfortranarray(np.triu(A)).reshape(-1, order="F")).cuda()
    ok, v = ldl.factor(S)
    return ldl, ok, v


@pytest.mark.parametrize("n", [1, 2, 3, 17, 60])
@pytest.mark.parametrize("kind", ["random", "smalldiag", "zeroblocks", "spd"])
def test_ldl_matches_oracle(blsamd, n, kind):
    import torch
    rng = np.random.default_rng(100 * n + len(kind))
    A = _sym(rng, n, kind)
    ldl, ok, neg = _device(blsamd, A)
    F, ipiv, info = D.bunch_kaufman_lower(np.triu(A))
    assert ok and info == 0
    Fd = torch.empty(n * n, dtype=torch.float64, device="cuda")
    ipd = torch.empty(n, dtype=torch.int64, device="cuda")
    ldl.factor_copy(Fd, ipd)
    torch.cuda.synchronize()
    assert np.array_equal(ipd.cpu().numpy(), ipiv)
    Fdh = np.tril(Fd.cpu().numpy().reshape(n, n, order="F"))
    assert np.abs(Fdh - F).max() <= 1e-12 * max(1.0, np.abs(F).max())
    assert neg == D.bk_inertia(F, ipiv) == int((np.linalg.eigvalsh(A) < 0).sum())
    g = rng.standard_normal(n)
    x = torch.full((n,), float("nan"), dtype=torch.float64, device="cuda")
    ldl.solve(torch.as_tensor(g).cuda(), x)
    xo = D.ldl_solve(F, ipiv, g)
    assert np.abs(x.cpu().numpy() - xo).max() <= 1e-10 * max(1.0, np.abs(xo).max())


@pytest.mark.parametrize("n,kind", [(200, "random"), (700, "smalldiag"), (1500, "zeroblocks")])
def test_ldl_large_solves(blsamd, n, kind):
    """Beyond the unblocked range: the inertia and the solve (x = g in place too)."""
    import torch
    rng = np.random.default_rng(n)
    A = _sym(rng, n, kind)
    ldl, ok, neg = _device(blsamd, A)
    assert ok and neg == int((np.linalg.eigvalsh(A) < 0).sum())
    g = rng.standard_normal(n)
    x = torch.as_tensor(g).cuda()
    ldl.solve(x, x)
    xh = x.cpu().numpy()
    assert np.linalg.norm(A @ xh - g) <= 1e-12 * np.linalg.norm(A, 2) * np.linalg.norm(xh) * n ** 0.5
    assert np.linalg.norm(xh - np.linalg.solve(A, g)) <= max(1e-10, 1e-14 * np.linalg.cond(A) * n ** 0.5) * np.linalg.norm(xh)


def test_ldl_zero_pivot(blsamd):
    rng = np.random.default_rng(7)
    A = _sym(rng, 9, "random")
    A[4, :] = A[:, 4] = 0.0
    _, ipiv, info = D.bunch_kaufman_lower(np.triu(A))
    assert info > 0
    ldl, ok, col = _device(blsamd, A)
    assert not ok and col == info
    import torch
    with pytest.raises(blsamd.device.BLSError):   # no factor to solve with
        ldl.solve(torch.zeros(9, dtype=torch.float64, device="cuda"), torch.zeros(9, dtype=torch.float64, device="cuda"))
