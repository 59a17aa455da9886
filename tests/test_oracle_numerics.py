"""CPU tests that pin the oracle's Schur numerics to what the mathematics fixes (dense brute force, full solves,
invariants) and to the reference's own test cases. The GPU tests then compare the HIP kernels with this oracle."""
import json
import os

import numpy as np
import pytest

import oracle
from helpers import problem_from_dense, reference_schur_problem, schur_test_problem, standard_problem_dims

HERE = os.path.dirname(os.path.abspath(__file__))
LIT = json.load(open(os.path.join(HERE, "golden", "reference_literals.json")))


def _dense_parts(p):
    st = p.structure
    nr, ne = st.nr, st.ne
    E = np.zeros((ne, ne))
    C = np.zeros((ne, nr))
    for e in range(st.neblocks):
        d = st.e_dims[e]
        o = st.boffsets_e[e] - 1
        eo = st.eoffsets[e] - 1
        E[o:o + d, o:o + d] = p.E_data[eo:eo + d * d].reshape(d, d).T
        for c in range(st.eliminated_ptr[e] - 1, st.eliminated_ptr[e + 1] - 1):
            r = st.entry_retained[c] - 1
            dr = st.r_dims[r]
            ro = st.boffsets_r[r] - 1
            co = st.coffsets[c] - 1
            C[o:o + d, ro:ro + dr] = p.C_data[co:co + dr * d].reshape(dr, d).T
    return E, C


def _S_dense(p, S):
    st = p.structure
    if p.dense:
        return S.reshape(st.nr, st.nr).T
    M = np.zeros((st.nr, st.nr))
    pat = p.pattern
    for j in range(st.nr):
        for k in range(pat.S_colptr[j] - 1, pat.S_colptr[j + 1] - 1):
            M[pat.S_rowval[k] - 1, j] = S[k]
    return M


def _R_dense(p):
    st = p.structure
    if p.dense:
        return p.R
    return p.H[:st.nr, :st.nr] if hasattr(p, "H") else _R_from_blocks(p)


def _R_from_blocks(p):
    st = p.structure
    R = np.zeros((st.nr, st.nr))
    for r in range(st.nrblocks):
        o, d = st.boffsets_r[r] - 1, st.r_dims[r]
        R[o:o + d, o:o + d] = p.Rblocks[r, :d, :d]
    return R


@pytest.mark.parametrize("which", ["cfg1", "cfg2sparse", "refdense", "refsparse", "standard"])
def test_reduction_equals_dense_brute_force(blsamd, which):
    """S = R + lambda D_r - C'(E + lambda D_e)^-1 C and g = b_r - C'(E + lambda D_e)^-1 b_e, upper triangle."""
    rng = np.random.default_rng(1)
    if which == "cfg1":
        p = blsamd.generators.config1()
    elif which == "cfg2sparse":
        p = blsamd.generators.config2(scale=0.002)
    elif which in ("refdense", "refsparse"):
        r_dims, e_dims, cpl = reference_schur_problem()
        H, b = schur_test_problem(rng, r_dims, e_dims, cpl, rr_pairs=[(1, 2), (2, 3), (3, 4)])
        p = problem_from_dense(blsamd, H, b, r_dims, e_dims, cpl, 0.3, which == "refdense", rr_pairs=[(1, 2), (2, 3), (3, 4)])
    else:
        r_dims, e_dims = standard_problem_dims()
        cpl = [sorted(rng.choice(np.arange(1, 5), size=rng.integers(1, 5), replace=False).tolist()) for _ in e_dims]
        H, b = schur_test_problem(rng, r_dims, e_dims, cpl)
        p = problem_from_dense(blsamd, H, b, r_dims, e_dims, cpl, 1e-3, False)
    res = oracle.schur_reduce_problem(p)
    assert res["ok"]
    st = p.structure
    E, C = _dense_parts(p)
    lam = p.lam
    R = _R_dense(p)
    R = np.triu(R) + np.triu(R, 1).T
    Ed = E + lam * np.diag(p.scaling_e)
    Sx = R + lam * np.diag(p.scaling_r) - C.T @ np.linalg.solve(Ed, C)
    gx = p.b_r - C.T @ np.linalg.solve(Ed, p.b_e)
    S = _S_dense(p, res["S"])
    iu = np.triu_indices(st.nr)
    pattern_mask = np.ones(len(iu[0]), bool) if p.dense else (np.abs(S[iu]) > 0) | (np.abs(Sx[iu]) > 0)
    assert np.abs(S[iu] - Sx[iu])[pattern_mask].max() <= 1e-12 * np.abs(Sx).max()
    assert np.abs(res["g"] - gx).max() <= 1e-12 * np.abs(gx).max()


def test_sparse_equals_dense(blsamd):
    """test/schur.jl:369-374: the sparse and dense reductions give the same S entries."""
    rng = np.random.default_rng(4)
    r_dims, e_dims, cpl = reference_schur_problem()
    H, b = schur_test_problem(rng, r_dims, e_dims, cpl)
    pd = problem_from_dense(blsamd, H, b, r_dims, e_dims, cpl, 0.1, True)
    ps = problem_from_dense(blsamd, H, b, r_dims, e_dims, cpl, 0.1, False)
    rd, rs = oracle.schur_reduce_problem(pd), oracle.schur_reduce_problem(ps)
    Sd, Ss = np.triu(_S_dense(pd, rd["S"])), np.triu(_S_dense(ps, rs["S"]))
    mask = Ss != 0
    assert np.abs(Sd[mask] - Ss[mask]).max() <= 1e-12
    assert np.array_equal(rd["g"], rs["g"])


def test_chunking_invariance_bitwise(blsamd):
    """test/schur.jl:255-256 and :399-400: K = 1 and K = 4 chunkings give identical results (here bitwise)."""
    p1 = blsamd.generators.config2(scale=0.003, nchunks=1)
    p4 = blsamd.generators.config2(scale=0.003, nchunks=4)
    a = oracle.schur_reduce_problem(p1, nthreads=1)
    b = oracle.schur_reduce_problem(p4, nthreads=4)
    assert np.array_equal(a["S"], b["S"]) and np.array_equal(a["g"], b["g"])


def test_rank_one_fails_at_lowest_block(blsamd):
    """test/schur.jl:405-425: a rank-1 E at lambda = 0 makes the reduction return false with the lowest block."""
    p = blsamd.generators.config2(scale=0.002, nchunks=3)
    st = p.structure
    for e in (9, 5, 14):
        eo = st.eoffsets[e - 1] - 1
        v = np.array([1.0, -2.0, 0.5])
        p.E_data[eo:eo + 9] = np.outer(v, v).T.reshape(-1)
    for nt in (1, 3):
        res = oracle.schur_reduce_problem(p, lam=0.0, nthreads=nt)
        assert res["ok"] is False and res["failed_block"] == 5
    res = oracle.schur_reduce_problem(p, lam=1e-2)   # damping makes the blocks invertible again
    assert res["ok"]


def test_full_solve_equals_dense_cholesky(blsamd):
    """reference_solve (test/fixtures/SchurFixtures.jl:114-117): reduce, solve S, back-substitute == solve of H."""
    rng = np.random.default_rng(9)
    r_dims, e_dims, cpl = reference_schur_problem()
    H, b = schur_test_problem(rng, r_dims, e_dims, cpl)
    for dense in (True, False):
        p = problem_from_dense(blsamd, H, b, r_dims, e_dims, cpl, 0.3, dense)
        res = oracle.schur_reduce_problem(p)
        S = _S_dense(p, res["S"])
        S = np.triu(S) + np.triu(S, 1).T
        x_r = np.linalg.solve(S, res["g"])
        x_e = np.zeros(p.structure.ne)
        oracle.backsubstitute_schur(res["args"], x_e, x_r, p.C_data, res["E_inv"], res["z"])
        x = np.concatenate([x_r, x_e])
        x_ref = np.linalg.solve(H + 0.3 * np.eye(len(b)), b)
        assert np.linalg.norm(x - x_ref) <= 1e-10 * np.linalg.norm(x_ref)


@pytest.mark.parametrize("de", [1, 2, 3, 4, 5, 7, 8, 9, 12, 300])
def test_eliminated_sizes_full_solve(blsamd, de):
    """test/schur_problems.jl:13-41: d_e in {1,3,4,5,7,8,9,...}, static and dynamic (d_e > 8) paths."""
    rng = np.random.default_rng(de)
    cpl = [sorted(rng.choice(np.arange(1, 5), size=rng.integers(1, 4), replace=False).tolist()) for _ in range(6)]
    H, b = schur_test_problem(rng, [6, 6, 6, 6], [de] * 6, cpl)
    p = problem_from_dense(blsamd, H, b, [6, 6, 6, 6], [de] * 6, cpl, 0.0, True)
    res = oracle.schur_reduce_problem(p)
    assert res["ok"]
    S = _S_dense(p, res["S"])
    S = np.triu(S) + np.triu(S, 1).T
    x_r = np.linalg.solve(S, res["g"])
    x_e = np.zeros(p.structure.ne)
    oracle.backsubstitute_schur(res["args"], x_e, x_r, p.C_data, res["E_inv"], res["z"])
    x_ref = np.linalg.solve(H, b)
    assert np.linalg.norm(np.concatenate([x_r, x_e]) - x_ref) <= 1e-9 * np.linalg.norm(x_ref)


@pytest.mark.parametrize("dense", [True, False])
def test_wide_retained_and_eliminated_full_solve(blsamd, dense):
    """test/schur_problems.jl:36-41 (R9): a retained block of 260 and an eliminated block of 300; the oracle's
    reduction + backsubstitute! equal the dense solve of (H + 0.2 I) x = b."""
    rng = np.random.default_rng(256)
    cpl = [[1, 2], [2]]
    H, b = schur_test_problem(rng, [6, 260], [300, 4], cpl)
    p = problem_from_dense(blsamd, H, b, [6, 260], [300, 4], cpl, 0.2, dense)
    res = oracle.schur_reduce_problem(p)
    assert res["ok"]
    S = _S_dense(p, res["S"])
    S = np.triu(S) + np.triu(S, 1).T
    x_r = np.linalg.solve(S, res["g"])
    x_e = np.zeros(p.structure.ne)
    oracle.backsubstitute_schur(res["args"], x_e, x_r, p.C_data, res["E_inv"], res["z"])
    x_ref = np.linalg.solve(H + 0.2 * np.eye(len(b)), b)
    assert np.linalg.norm(np.concatenate([x_r, x_e]) - x_ref) <= 1e-10 * np.linalg.norm(x_ref)


def test_inv_sym_block_rejects_not_pd():
    """test/kernels.jl:19-25: not positive definite yet with a positive determinant or 1x1 entry test."""
    for A in LIT["inv_sym_block_not_pd"]["matrices"]:
        A = np.array(A, dtype=float)
        d = A.shape[0]
        assert not oracle.inv_sym_block(A, np.zeros(d))[0]
        assert not oracle.inv_sym_block(A, np.zeros(d), chol=True)[0]


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_closed_forms_agree_with_cholesky_path(dtype):
    """test/kernels.jl:27-48: for sizes 1-3 the closed forms accept exactly what the Cholesky path accepts, and
    their inverses agree; both agree with numpy's inverse."""
    rng = np.random.default_rng(1)
    for D in (1, 2, 3):
        for kind in ("pd", "indefinite", "nd", "nearsingular"):
            for damped in (False, True):
                for _ in range(50):
                    Q, _r = np.linalg.qr(rng.standard_normal((D, D)))
                    if kind == "pd":
                        ev = rng.random(D) + 0.1
                    elif kind == "nd":
                        ev = -(rng.random(D) + 0.1)
                    elif kind == "indefinite":
                        ev = rng.standard_normal(D)
                        ev[0] = abs(ev[0]) + 0.1
                        ev[-1] = -abs(ev[-1]) - 0.1
                    else:
                        ev = rng.random(D) + 0.1
                        ev[0] = rng.choice([1e-3, 1e-8, 1e-16, 0.0, -1e-16])
                    A = Q @ np.diag(ev) @ Q.T
                    H = ((A + A.T) / 2).astype(dtype)
                    lam = dtype(rng.choice([1e-6, 1e-2, 1.0])) if damped else dtype(0)
                    damp = (lam * np.abs(np.diag(H))).astype(dtype) if damped else np.zeros(D, dtype)
                    ok1, inv1 = oracle.inv_sym_block(H, damp)
                    ok2, inv2 = oracle.inv_sym_block(H, damp, chol=True)
                    assert ok1 == ok2
                    if ok1:
                        rtol = np.sqrt(np.finfo(dtype).eps) * 1e3
                        assert np.allclose(inv1, inv2, rtol=rtol, atol=0)
                        Hd = H.astype(np.float64) + np.diag(damp.astype(np.float64))
                        assert np.allclose(inv1, np.linalg.inv(Hd), rtol=rtol * 10, atol=rtol * np.abs(np.linalg.inv(Hd)).max())


def test_diagonal_blocks_factor_and_solve():
    """test/diagonal_blocks.jl:16-40: per-block Cholesky equals numpy's, the solve equals the block-diagonal solve."""
    rng = np.random.default_rng(7)
    dims = [1, 3, 6, 2, 9]
    blocks = []
    for d in dims:
        X = rng.standard_normal((d, d))
        blocks.append(X @ X.T + d * np.eye(d))
    D = np.concatenate([B.T.reshape(-1) for B in blocks])
    F = D.copy()
    assert oracle.factor_diagonal_blocks(np.array(dims), F)
    o = 0
    for B, d in zip(blocks, dims):
        L = np.tril(F[o:o + d * d].reshape(d, d).T)
        assert np.allclose(L, np.linalg.cholesky(B), rtol=1e-12, atol=1e-12)
        o += d * d
    x = rng.standard_normal(sum(dims))
    y = np.zeros_like(x)
    oracle.solve_diagonal_blocks(np.array(dims), y, F, x)
    import scipy.linalg as sl
    ref = np.linalg.solve(sl.block_diag(*blocks), x)
    assert np.allclose(y, ref, rtol=1e-12, atol=1e-12)
    # the strict upper triangle is not read
    G = F.copy()
    o = 0
    for d in dims:
        for c in range(d):
            for r in range(c):
                G[o + c * d + r] = np.nan
        o += d * d
    y2 = np.zeros_like(x)
    oracle.solve_diagonal_blocks(np.array(dims), y2, G, x)
    assert np.array_equal(y, y2)
    # an indefinite block fails
    bad = D.copy()
    bad[0] = -1.0
    assert not oracle.factor_diagonal_blocks(np.array(dims), bad)


def test_pcg_reaches_tolerance_and_matches_direct_solve(blsamd):
    """The oracle PCG (no reference counterpart) drives the true relative residual below 1e-8 and agrees with a
    direct solve of the reduced system."""
    p = blsamd.generators.config2(scale=0.01)
    res = oracle.schur_reduce_problem(p)
    st = p.structure
    x, it, rel = oracle.pcg(st, res["args"], res["S"], res["g"], tol=1e-8)
    assert it > 0 and rel <= 1e-8
    y = oracle.sym_spmv(st, res["args"], res["S"], x)
    assert np.linalg.norm(y - res["g"]) <= 1e-8 * np.linalg.norm(res["g"]) * 1.0000001
    S = _S_dense(p, res["S"])
    S = np.triu(S) + np.triu(S, 1).T
    xd = np.linalg.solve(S, res["g"])
    assert np.linalg.norm(x - xd) <= 1e-5 * np.linalg.norm(xd)
    # the symmetric product equals the dense symmetric matrix times x
    assert np.allclose(S @ x, y, rtol=1e-12, atol=1e-12 * np.abs(y).max())
