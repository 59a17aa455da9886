"""Cholesky of the dense reduced system (TEST INFRASTRUCTURE ONLY): the checker of the device's bls_chol_*.

The reference's dense reduced solve (solve_reduced_dense!, src/schur_linear_systems.jl:312-319) factors a copy of S
with Cholesky through DampedLinearSolve 0.1.1, a dependency absent here, whose Cholesky is LAPACK's potrf (upper):
S = U'U, U upper triangular with a positive diagonal, computed row by row of U (dpotf2: each row scaled by the
reciprocal of its pivot); a leading minor that is not positive
definite stops it (info = its order). Restated here unblocked, in numpy Float64.

Parity pinning (tests/test_conditioning.py): U'U == S and U equal to numpy.linalg.cholesky's (LAPACK) transposed on
random SPD matrices; the failing column on matrices built so that leading minor j is the first that is not positive
definite (a negative pivot, an exactly singular minor, a NaN); the solve against numpy.linalg.solve.

Its fallback when Cholesky fails (`:ldl`, src/schur_linear_systems.jl:312-319) is a Bunch-Kaufman factorization:
Julia's bunchkaufman, LAPACK's sytrf. bunch_kaufman_lower restates the unblocked LAPACK dsytf2 on the lower
triangle (alpha = (1 + sqrt 17) / 8, 1 x 1 and 2 x 2 pivots with symmetric interchanges, LAPACK's ipiv), and
ldl_solve its dsytrs. Pinned (tests/test_conditioning.py) against scipy's LAPACK dsytrf on matrices of up to 60
rows, where dsytrf runs dsytf2 itself (ipiv equal, factor within 1e-13), the inertia against numpy's eigenvalues
(Sylvester's law: the negative eigenvalues of D are those of S) and the solve against numpy.linalg.solve.
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


BK_ALPHA = (1.0 + np.sqrt(17.0)) / 8.0


def bunch_kaufman_lower(S):
    """(A, ipiv, info): LAPACK dsytf2 ('L') on the lower triangle of S (mirrored from its upper triangle): A holds
    L below the diagonal and D on it (2 x 2 blocks: the diagonal and A[k+1, k]); ipiv is LAPACK's, 1-based (k: a
    1 x 1 pivot after interchanging k and ipiv[k]; -p on both rows of a 2 x 2 pivot after interchanging k + 1 and
    p); info 0 or the first zero pivot column (1-based; the factorization continues past it, as dsytf2 does)."""
    A = np.triu(np.array(S, dtype=np.float64))
    A = A + np.triu(A, 1).T
    n = A.shape[0]
    ipiv = np.zeros(n, dtype=np.int64)
    info = 0
    k = 0
    while k < n:
        kstep = 1
        absakk = abs(A[k, k])
        if k < n - 1:
            imax = k + 1 + int(np.argmax(np.abs(A[k + 1:, k])))
            colmax = abs(A[imax, k])
        else:
            imax, colmax = k, 0.0
        if max(absakk, colmax) == 0.0 or np.isnan(absakk):
            if info == 0:
                info = k + 1
            kp = k
        else:
            if absakk >= BK_ALPHA * colmax:
                kp = k
            else:
                row = np.concatenate([np.abs(A[imax, k:imax]), np.abs(A[imax + 1:, imax])])
                rowmax = float(row.max())   # the largest off-diagonal entry of row / column imax
                if absakk >= BK_ALPHA * colmax * (colmax / rowmax):
                    kp = k
                elif abs(A[imax, imax]) >= BK_ALPHA * rowmax:
                    kp = imax
                else:
                    kp = imax
                    kstep = 2
            kk = k + kstep - 1
            if kp != kk:   # interchange rows and columns kk and kp of the trailing lower triangle
                A[kp + 1:, [kk, kp]] = A[kp + 1:, [kp, kk]]
                for j in range(kk + 1, kp):
                    A[j, kk], A[kp, j] = A[kp, j], A[j, kk]
                A[kk, kk], A[kp, kp] = A[kp, kp], A[kk, kk]
                if kstep == 2:
                    A[k + 1, k], A[kp, k] = A[kp, k], A[k + 1, k]
            if kstep == 1:
                if k < n - 1:
                    d11 = 1.0 / A[k, k]
                    x = A[k + 1:, k].copy()
                    for j in range(k + 1, n):   # dsyr: A(i, j) += x(i) (-d11 x(j)), then the column scaled
                        A[j:, j] += x[j - k - 1:] * (-d11 * x[j - k - 1])
                    A[k + 1:, k] *= d11
            elif k < n - 2:
                d21 = A[k + 1, k]
                d11 = A[k + 1, k + 1] / d21
                d22 = A[k, k] / d21
                t = 1.0 / (d11 * d22 - 1.0)
                d21 = t / d21
                for j in range(k + 2, n):
                    wk = d21 * (d11 * A[j, k] - A[j, k + 1])
                    wkp1 = d21 * (d22 * A[j, k + 1] - A[j, k])
                    A[j:, j] = A[j:, j] - A[j:, k] * wk - A[j:, k + 1] * wkp1
                    A[j, k] = wk
                    A[j, k + 1] = wkp1
        if kstep == 1:
            ipiv[k] = kp + 1
        else:
            ipiv[k] = ipiv[k + 1] = -(kp + 1)
        k += kstep
    return np.tril(A), ipiv, info


def bk_inertia(A, ipiv):
    """Number of negative eigenvalues of D (= of S, by Sylvester's law of inertia)."""
    n, neg, k = len(ipiv), 0, 0
    while k < n:
        if ipiv[k] > 0:
            neg += A[k, k] < 0
            k += 1
        else:
            a, b, c = A[k, k], A[k + 1, k], A[k + 1, k + 1]
            det, tr = a * c - b * b, a + c
            neg += 1 if det < 0 else (2 if tr < 0 else 0)
            k += 2
    return int(neg)


def ldl_solve(A, ipiv, g):
    """x = S^-1 g from bunch_kaufman_lower's factor: LAPACK dsytrs ('L'), L D y = P g then L' x = y."""
    n = len(ipiv)
    B = np.array(g, dtype=np.float64)
    k = 0
    while k < n:
        if ipiv[k] > 0:
            kp = ipiv[k] - 1
            B[k], B[kp] = B[kp], B[k]
            if k < n - 1:
                B[k + 1:] -= A[k + 1:, k] * B[k]
            B[k] *= 1.0 / A[k, k]
            k += 1
        else:
            kp = -ipiv[k] - 1
            B[k + 1], B[kp] = B[kp], B[k + 1]
            if k < n - 2:
                B[k + 2:] -= A[k + 2:, k] * B[k]
                B[k + 2:] -= A[k + 2:, k + 1] * B[k + 1]
            akm1k = A[k + 1, k]
            akm1 = A[k, k] / akm1k
            ak = A[k + 1, k + 1] / akm1k
            denom = akm1 * ak - 1.0
            bkm1 = B[k] / akm1k
            bk = B[k + 1] / akm1k
            B[k] = (ak * bkm1 - bk) / denom
            B[k + 1] = (akm1 * bk - bkm1) / denom
            k += 2
    k = n - 1
    while k >= 0:
        if ipiv[k] > 0:
            if k < n - 1:
                B[k] -= A[k + 1:, k] @ B[k + 1:]
            kp = ipiv[k] - 1
            B[k], B[kp] = B[kp], B[k]
            k -= 1
        else:
            if k < n - 1:
                B[k] -= A[k + 1:, k] @ B[k + 1:]
                B[k - 1] -= A[k + 1:, k - 1] @ B[k + 1:]
            kp = -ipiv[k] - 1
            B[k], B[kp] = B[kp], B[k]
            k -= 2
    return B
