"""Cholesky of the dense reduced system (TEST INFRASTRUCTURE ONLY): the checker of the device's bls_chol_*.

The reference's dense reduced solve (solve_reduced_dense!, src/schur_linear_systems.jl:312-319) factors a copy of S
with Cholesky through DampedLinearSolve 0.1.1, a dependency absent here, whose Cholesky is LAPACK's potrf (upper):
S = U'U, U upper triangular with a positive diagonal, computed row by row of U (dpotf2: each row scaled by the
reciprocal of its pivot); a leading minor that is not positive
definite stops it (info = its order). Restated here unblocked, in numpy Float64.

Parity pinning (tests/test_conditioning.py): U'U == S and U equal to numpy.linalg.cholesky's (LAPACK) transposed on
random SPD matrices; the failing column on matrices built so that leading minor j is the first that is not positive
definite (a negative pivot, an exactly singular minor, a NaN); the solve against numpy.linalg.solve.
"""
import numpy as np


def cholesky_upper(S):
    """(U, info): U'U = S from the upper triangle of S; info 0, or the first column (1-based) whose pivot is not
    positive (U is then partial)."""
    n = S.shape[0]
    U = np.triu(np.array(S, dtype=np.float64))
    for j in range(n):
        piv = U[j, j]
        if not piv > 0.0:
            return U, j + 1
        U[j, j] = np.sqrt(piv)
        U[j, j + 1:] *= 1.0 / U[j, j]          # dpotf2 scales the row by the reciprocal
        for r in range(j + 1, n):
            U[r, r:] -= U[j, r] * U[j, r:]
    return U, 0


def cholesky_solve(U, g):
    """x = (U'U)^-1 g by U'y = g, then U x = y."""
    n = U.shape[0]
    y = np.array(g, dtype=np.float64)
    for i in range(n):
        y[i] = (y[i] - U[:i, i] @ y[:i]) / U[i, i]
    x = y
    for i in range(n - 1, -1, -1):
        x[i] = (x[i] - U[i, i + 1:] @ x[i + 1:]) / U[i, i]
    return x
