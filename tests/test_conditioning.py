"""The restatements of the conditioning scale (oracle/conditioning_ref.py, src/conditioning.jl:24-87), of getdiag!
(oracle/assembly_ref.py) and of the table quantities jtr!, jtj_diag!, jtj_quadform (oracle/jacobian_ref.py,
src/qr_reduced.jl:529-645), pinned by what the mathematics fixes and by the dense J of the same tables."""
import math

import numpy as np
import pytest

import assembly_ref as A
import conditioning_ref as C
import jacobian_ref as JR
from helpers import assembly_problem
from test_assembly import _dense_J as _dense_J_rows
from test_oracle_qr import _dense_J, _jac_with_free_rows_cpu, _small_jacobian


def _scale_cases(rng, n=20000):
    """Damped diagonals over 1e-40 .. 1e40, both signs, zeros, exact powers of two and their neighbours."""
    d = np.exp(rng.uniform(np.log(1e-40), np.log(1e40), n)) * rng.choice([-1.0, 1.0], n)
    p2 = np.ldexp(1.0, rng.integers(-120, 120, 2000)).astype(np.float64)
    return np.concatenate([d, [0.0, -0.0, 1.0, 4.0, 0.5, 1e-24, 1e-30, 1e300, 2.0 ** -1074], p2,
                           np.nextafter(p2, 0.0), np.nextafter(p2, np.inf)])


def test_scale_is_the_nearest_power_of_two():
    """s is a power of two and q / s lies in [sqrt(1/2), sqrt(2)) for q = 1 / (1e-12 + sqrt|d|): s is the power of
    two nearest q on a log scale (src/conditioning.jl:29-35). Checked with math.frexp per value, independently."""
    rng = np.random.default_rng(0)
    d = _scale_cases(rng)
    s = C.conditioning_scale(d)
    for dj, sj in zip(d, s):
        q = 1.0 / (1e-12 + math.sqrt(abs(dj)))
        f, _ = math.frexp(sj)
        assert f == 0.5, (dj, sj)                      # a power of two
        assert math.sqrt(0.5) <= q / sj < math.sqrt(2.0), (dj, sj, q)


def test_scale_special_cases():
    """d = 0: 1 / 1e-12 ~ 2^39.86, so 2^40; non-finite: 1; the floor floatmin(T) (2^-126 for Float32, so 1e300 is
    floored there but not in Float64); the sign of d is ignored."""
    assert C.conditioning_scale([0.0])[0] == 2.0 ** 40
    assert list(C.conditioning_scale([np.inf, -np.inf, np.nan])) == [1.0, 1.0, 1.0]
    assert C.conditioning_scale([1e300], np.float32)[0] == 2.0 ** -126
    s64 = C.conditioning_scale([1e300])[0]
    assert s64 < 2.0 ** -126 and math.frexp(s64)[0] == 0.5
    assert C.conditioning_scale([-4.0])[0] == C.conditioning_scale([4.0])[0] == 0.5
    assert C.conditioning_scale([1.0])[0] == 1.0


def _drift_sequence(prep, d, c):
    """test/conditioning.jl:72-80: lambda = 0 computes s0; 1e-2 max(d) keeps it (within the band); 1e16 max(d)
    recomputes it; 0 again recomputes it back to s0."""
    md = float(np.max(d))
    s0 = prep(0.0).copy()
    assert c_nrescale(c) == 0
    assert np.array_equal(prep(1e-2 * md), s0) and c_nrescale(c) == 0
    s1 = prep(1e16 * md).copy()
    assert c_nrescale(c) == 1 and not np.array_equal(s1, s0)
    assert np.array_equal(prep(0.0), s0) and c_nrescale(c) == 2
    return s0


def c_nrescale(c):
    return c.nrescale if isinstance(c.nrescale, int) else c.nrescale()


def test_prepare_scale_band_and_recomputation():
    rng = np.random.default_rng(1)
    n = 500
    d = rng.uniform(0.1, 10.0, n) * 10.0 ** rng.integers(-2, 3, n)   # 1e-3 .. 1e3: 1e-2 max(d) stays in the band
    sc = rng.uniform(0.5, 2.0, n)
    c = C.EliminationConditioning(n)
    calls = []
    prep = lambda lam: C.prepare_scale(c, lambda dd: (calls.append(1), dd.__setitem__(slice(None), d)), lam, sc)
    s0 = _drift_sequence(prep, d, c)
    assert len(calls) == 1                               # diag! only at the first solve of the fill
    assert np.array_equal(s0, C.conditioning_scale(d))
    c.stale = True                                       # zero! / markstale!
    prep(0.5)
    assert len(calls) == 2 and np.array_equal(c.s, C.conditioning_scale(d + 0.5 * sc))


@pytest.mark.parametrize("sparse", [False, True])
def test_getdiag_is_the_diagonal_of_JtJ(blsamd, sparse):
    """getdiag! of the assembled storage == diag(J'J) of the stacked row blocks."""
    st, rowblocks, rb_rows, J, r, R_layout = assembly_problem(blsamd, 3, nland=25, sparse=sparse)
    H = A.assemble(A.HessianStorage(st, *(R_layout or ())), rowblocks, rb_rows, J, r)
    Jd, _ = _dense_J_rows(st, rowblocks, rb_rows, J, r)
    ref = np.sum(Jd * Jd, axis=0)
    assert np.abs(A.getdiag(H) - ref).max() <= 1e-12 * ref.max()


def test_getdiag_missing_diagonal_block_reads_zero(blsamd):
    """diag_retained! (src/schur_linear_systems.jl:259-270): a retained block whose diagonal block R does not store
    has a zero diagonal."""
    st, _, _, _, _, (cp, rv, ro) = assembly_problem(blsamd, 4, nland=10, sparse=True)
    H = A.HessianStorage(st, cp, rv, ro)
    H.R[:] = 1.0 + np.arange(len(H.R))
    keep = [k for k in range(len(rv)) if not (rv[k] == 1 and k < cp[1] - 1)]   # drop block (1, 1)
    cp2 = np.array([1] + [1 + sum(1 for k in keep if k < cp[q + 1] - 1) for q in range(len(cp) - 1)])
    H2 = A.HessianStorage(st, cp2, np.asarray(rv)[keep], np.asarray(ro)[keep])
    H2.R = H.R
    d = A.getdiag(H2)
    assert np.all(d[:int(st.r_dims[0])] == 0.0) and np.all(d[int(st.r_dims[0]):H.nr] > 0.0)


@pytest.mark.parametrize("kind", ["eliminated", "free"])
def test_table_quantities_match_dense_J(blsamd, kind):
    """jtr!, jtj_diag!, jtj_quadform from the tables == J'r, diag(J'J), ||J v||^2 of the dense J of the same tables
    (free tables too)."""
    p = _small_jacobian(blsamd, 2, nland=30, ncam=6) if kind == "eliminated" else _jac_with_free_rows_cpu(blsamd)
    Jd, rd = _dense_J(p)
    v = np.random.default_rng(3).standard_normal(Jd.shape[1])
    g = JR.jtr(p.jst, p.data, p.r)
    d = JR.jtj_diag(p.jst, p.data)
    q = JR.jtj_quadform(p.jst, p.data, v)
    gr, dr, qr = Jd.T @ rd, np.sum(Jd * Jd, axis=0), float(np.sum((Jd @ v) ** 2))
    assert np.abs(g - gr).max() <= 1e-12 * np.abs(gr).max()
    assert np.abs(d - dr).max() <= 1e-12 * dr.max()
    assert abs(q - qr) <= 1e-12 * qr


def _spd(rng, n, cond=1e3):
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    return (Q * np.geomspace(1.0, cond, n)) @ Q.T


@pytest.mark.parametrize("n", [1, 5, 40])
def test_dense_cholesky_oracle(n):
    """U'U == S, U == numpy's (LAPACK potrf) Cholesky factor transposed, and the solve == numpy.linalg.solve."""
    import dense_ref as D
    rng = np.random.default_rng(n)
    S = _spd(rng, n)
    U, info = D.cholesky_upper(S)
    assert info == 0
    assert np.abs(U.T @ U - S).max() <= 1e-13 * np.abs(S).max()
    assert np.abs(U - np.linalg.cholesky(S).T).max() <= 1e-12 * np.abs(U).max()
    g = rng.standard_normal(n)
    x = D.cholesky_solve(U, g)
    assert np.abs(x - np.linalg.solve(S, g)).max() <= 1e-10 * np.abs(x).max()


def test_dense_cholesky_oracle_failing_column():
    """The first leading minor that is not positive definite: a negative pivot at column 4, an exactly singular
    2 x 2 minor ([[1, 1], [1, 1]]), a NaN at column 3, and numpy agreeing that the matrix is not SPD."""
    import dense_ref as D
    S = np.diag([1.0, 2.0, 3.0, -1.0, 5.0])
    assert D.cholesky_upper(S)[1] == 4
    S = np.array([[1.0, 1.0, 0.0], [1.0, 1.0, 0.0], [0.0, 0.0, 1.0]])
    assert D.cholesky_upper(S)[1] == 2
    S = np.eye(4)
    S[2, 2] = np.nan
    assert D.cholesky_upper(S)[1] == 3
    with pytest.raises(np.linalg.LinAlgError):
        np.linalg.cholesky(np.diag([1.0, 2.0, 3.0, -1.0, 5.0]))


def test_bunch_kaufman_oracle_equals_lapack_dsytrf():
    """dense_ref.bunch_kaufman_lower restates LAPACK's dsytf2 (the reference's :ldl fallback is Julia's bunchkaufman,
    LAPACK sytrf): on matrices of up to 60 rows scipy's dsytrf runs dsytf2 itself, so ipiv and info are equal and the
    factor agrees to rounding; 2 x 2 pivots are forced with zero leading blocks and small diagonals. The inertia of D
    equals that of S (Sylvester's law) and ldl_solve solves."""
    import scipy.linalg.lapack as L
    import dense_ref as D
    rng = np.random.default_rng(0)
    n2 = 0
    for trial in range(200):
        n = int(rng.integers(1, 61))
        A = rng.standard_normal((n, n))
        A = A + A.T
        if trial % 3 == 0:
            A[np.diag_indices(n)] *= 1e-3
        if trial % 5 == 0 and n > 3:
            A[:2, :2] = 0.0
        F, ipiv, info = D.bunch_kaufman_lower(np.triu(A))
        lu, ipiv_l, info_l = L.dsytrf(A, lower=1)
        assert np.array_equal(ipiv, ipiv_l) and info == info_l
        assert np.abs(np.tril(lu) - F).max() <= 1e-12 * max(1.0, np.abs(F).max())
        n2 += int((ipiv < 0).sum())
        assert D.bk_inertia(F, ipiv) == int((np.linalg.eigvalsh(A) < 0).sum())
        g = rng.standard_normal(n)
        x = D.ldl_solve(F, ipiv, g)
        assert np.linalg.norm(A @ x - g) <= 1e-10 * np.linalg.norm(A, 2) * np.linalg.norm(x) + 1e-12
    assert n2 > 0   # 2 x 2 pivots were exercised
    # a zero column: info is its index and the factorization continues (dsytf2)
    A = rng.standard_normal((5, 5))
    A = A + A.T
    A[2, :] = A[:, 2] = 0.0
    _, ipiv, info = D.bunch_kaufman_lower(np.triu(A))
    _, ipiv_l, info_l = L.dsytrf(A, lower=1)
    assert info == info_l > 0 and np.array_equal(ipiv, ipiv_l)
