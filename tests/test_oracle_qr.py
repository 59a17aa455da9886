"""CPU tests pinning the oracle's QR elimination (reduce_qr!) and its reduced-problem products to what the
mathematics fixes: U'U = E + lambda diag(scaling), U'B = C, U'q = b_e (test/qr_reduction.jl:20-58), reduced normal
equations == the Schur reduction (test/schur_qr.jl:13-57), dense least squares (test/fixtures/QRFixtures.jl:164-169)."""
import numpy as np
import pytest

import oracle


def _small_jacobian(blsamd, seed=0, nland=40, ncam=6, mixed=True, kmax=4):
    rng = np.random.default_rng(seed)
    cam_dims = np.where(rng.random(ncam) < 0.5, 6, 9) if mixed else np.full(ncam, 6)
    k = rng.integers(2, kmax + 1, size=nland)
    ptr = np.concatenate([[0], np.cumsum(k)])
    cams = np.concatenate([rng.choice(ncam, size=kk, replace=False) for kk in k])
    return blsamd.generators.jacobian_from_tracks(rng, ncam, cam_dims.astype(np.int64), nland, k, ptr, cams, 0.1)


def _dense_J(p):
    """Dense J (rows of all tables' data rows, columns [x_r; x_e]) and r, from the raw tables."""
    jst, st = p.jst, p.structure
    nr, ne = st.nr, st.ne
    rows, res = [], []
    for t in range(jst.ntables):
        M = jst.table_nrows[t]
        top = jst.table_top[t]
        me = M - top if t < st.neblocks else M
        off = jst.raw_offsets[t] - 1
        C = jst.table_ncols[t] - 1
        T = p.data[off:off + M * C].reshape(C, M).T
        for i in range(me):
            row = np.zeros(nr + ne)
            for u in range(jst.table_ptr[t] - 1, jst.table_ptr[t + 1] - 1):
                r = jst.entry_retained[u] - 1
                ro, dr = st.boffsets_r[r] - 1, st.r_dims[r]
                c0 = jst.entry_col[u] - 1
                row[ro:ro + dr] = T[i, c0:c0 + dr]
            if t < st.neblocks:
                de = st.e_dims[t]
                eo = nr + st.boffsets_e[t] - 1
                row[eo:eo + de] = T[i, C - de:C]
            rows.append(row)
            res.append(p.r[jst.residual_offsets[t] - 1 + i])
    return np.array(rows), np.array(res)


def _qr(p, lam, nthreads=1):
    jst, st = p.jst, p.structure
    from blsload import load
    ws = load().structure.SchurReductionWorkspaceIndex(st, 1)
    sargs = oracle.SchurArgs(st, ws)
    jargs = oracle.JacobianArgs(jst)
    F = np.zeros(jst.raw_length())
    Fr = np.zeros(jst.residual_length())
    ok, fb = oracle.reduce_qr(sargs, jargs, F, Fr, p.data, p.r, p.scaling_e, lam, nthreads=nthreads)
    return ok, fb, F, Fr, sargs, jargs


def test_qr_factors_match_normal_equations(blsamd):
    """U'U = E + lambda diag(scaling_e), U'B = C (the coupling block), U'q = b_e, per eliminated block."""
    p = _small_jacobian(blsamd, 1)
    lam = 0.1
    ok, fb, F, Fr, _, _ = _qr(p, lam)
    assert ok and fb == 0
    jst, st = p.jst, p.structure
    J, r = _dense_J(p)
    nr = st.nr
    for e in range(st.neblocks):
        de = st.e_dims[e]
        M = jst.table_nrows[e]
        C = jst.table_ncols[e] - 1
        off = jst.raw_offsets[e] - 1
        T = F[off:off + M * C].reshape(C, M).T
        U = T[:de, C - de:C]
        assert np.allclose(np.tril(U, -1), 0)
        eo = nr + st.boffsets_e[e] - 1
        Je = J[:, eo:eo + de]
        E = Je.T @ Je + lam * np.diag(p.scaling_e[st.boffsets_e[e] - 1: st.boffsets_e[e] - 1 + de])
        assert np.allclose(U.T @ U, E, rtol=1e-11, atol=1e-11 * np.abs(E).max())
        B = T[:de, :C - de]
        Cr = np.zeros((de, C - de))
        for u in range(jst.table_ptr[e] - 1, jst.table_ptr[e + 1] - 1):
            rr = jst.entry_retained[u] - 1
            ro, dr = st.boffsets_r[rr] - 1, st.r_dims[rr]
            c0 = jst.entry_col[u] - 1
            Cr[:, c0:c0 + dr] = Je.T @ J[:, ro:ro + dr]
        assert np.allclose(U.T @ B, Cr, rtol=1e-11, atol=1e-11 * max(1, np.abs(Cr).max()))
        q = Fr[jst.residual_offsets[e] - 1: jst.residual_offsets[e] - 1 + de]
        be = Je.T @ r
        assert np.allclose(U.T @ q, be, rtol=1e-11, atol=1e-11 * max(1, np.abs(be).max()))


def test_reduced_normal_equals_schur_and_solves_least_squares(blsamd):
    """J~'J~ + lambda D_r and J~'r~ equal the Schur complement of the damped J'J (test/schur_qr.jl:13-57); the
    full step equals the dense damped least-squares solution."""
    p = _small_jacobian(blsamd, 2)
    lam = 0.05
    ok, fb, F, Fr, sargs, jargs = _qr(p, lam)
    assert ok
    jst, st = p.jst, p.structure
    nr = st.nr
    S = np.zeros(nr * nr)
    g = np.zeros(nr)
    oracle.reduced_normal(False, sargs, jargs, S, g, F, Fr)
    S = S.reshape(nr, nr).T
    S = np.triu(S) + np.triu(S, 1).T
    J, r = _dense_J(p)
    H = J.T @ J
    b = J.T @ r
    n = H.shape[0]
    D = np.zeros(n)
    D[nr:] = lam * p.scaling_e
    Hd = H + np.diag(D)
    Sch = Hd[:nr, :nr] - Hd[:nr, nr:] @ np.linalg.solve(Hd[nr:, nr:], Hd[nr:, :nr])
    gs = b[:nr] - Hd[:nr, nr:] @ np.linalg.solve(Hd[nr:, nr:], b[nr:])
    assert np.allclose(S, Sch, rtol=1e-10, atol=1e-10 * np.abs(Sch).max())
    assert np.allclose(g, gs, rtol=1e-10, atol=1e-10 * np.abs(gs).max())
    # full solve with retained damping lambda too, vs dense damped least squares
    Sd = S + lam * np.eye(nr)
    x_r = np.linalg.solve(Sd, g)
    x_e = np.zeros(st.ne)
    oracle.qr_backsubstitute(sargs, jargs, x_e, x_r, F, Fr)
    x = np.concatenate([x_r, x_e])
    x_ref = np.linalg.solve(H + lam * np.eye(n), b)
    assert np.linalg.norm(x - x_ref) <= 1e-9 * np.linalg.norm(x_ref)
    # the diagonal blocks and the products
    Dblk = np.zeros(int((st.r_dims ** 2).sum()))
    oracle.reduced_diagonal_blocks(sargs, jargs, Dblk, F)
    o = 0
    for rr in range(st.nrblocks):
        ro, d = st.boffsets_r[rr] - 1, st.r_dims[rr]
        assert np.allclose(Dblk[o:o + d * d].reshape(d, d).T, S[ro:ro + d, ro:ro + d], rtol=1e-10, atol=1e-10 * np.abs(S).max())
        o += d * d
    v = np.random.default_rng(0).standard_normal(nr)
    y = np.zeros(jst.reduced_nrows())
    oracle.reduced_mul(sargs, jargs, y, F, v)
    gt = np.zeros(nr)
    oracle.reduced_tmul(sargs, jargs, gt, F, y)
    assert np.allclose(gt, S @ v, rtol=1e-10, atol=1e-10 * np.abs(S @ v).max())
    rt = np.zeros(jst.reduced_nrows())
    oracle.reduced_residual(jargs, rt, Fr)
    gt2 = np.zeros(nr)
    oracle.reduced_tmul(sargs, jargs, gt2, F, rt)
    assert np.allclose(gt2, g, rtol=1e-12, atol=1e-12 * np.abs(g).max())


def test_qr_failures_lowest_block(blsamd):
    """test/qr_reduction.jl:60-92: an all-zero eliminated block at lambda = 0 is singular, a non-finite entry
    fails the table; the lowest failing block is reported, whatever the threading."""
    p = _small_jacobian(blsamd, 3)
    jst, st = p.jst, p.structure
    for e in (7, 11):   # zero the eliminated columns of tables 7 and 11 (1-based)
        M = jst.table_nrows[e - 1]
        C = jst.table_ncols[e - 1] - 1
        off = jst.raw_offsets[e - 1] - 1
        de = st.e_dims[e - 1]
        p.data[off + (C - de) * M: off + C * M] = 0.0
    off = jst.raw_offsets[20] - 1
    p.data[off] = np.nan    # table 21: a NaN in a retained column
    for nt in (1, 4):
        ok, fb, *_ = _qr(p, 0.0, nthreads=nt)
        assert not ok and fb == 7
    ok, fb, *_ = _qr(p, 0.5)      # damping makes the zero blocks regular; the NaN still fails table 21
    assert not ok and fb == 21


def test_qr_re_reduction_exact(blsamd):
    """test/qr_reduction.jl:116-142: raw tables are read only, so reducing twice gives identical results."""
    p = _small_jacobian(blsamd, 4)
    raw = p.data.copy()
    a = _qr(p, 0.2)
    b = _qr(p, 0.2)
    assert np.array_equal(a[2], b[2]) and np.array_equal(a[3], b[3])
    assert np.array_equal(p.data, raw)


def _column_scaled(p, s):
    """The raw tables with every column multiplied by its variable's s (the conditioned Jacobian J S), built from
    the table layout: entry u of table t holds retained block entry_retained[u] at columns entry_col[u] ..;
    an eliminated table's last d_e columns hold its eliminated variables."""
    jst, st = p.jst, p.structure
    data = p.data.copy()
    for t in range(jst.ntables):
        M, C, off = jst.table_nrows[t], jst.table_ncols[t] - 1, jst.raw_offsets[t] - 1
        T = data[off:off + M * C].reshape(C, M)
        for u in range(jst.table_ptr[t] - 1, jst.table_ptr[t + 1] - 1):
            r = jst.entry_retained[u] - 1
            for k in range(st.r_dims[r]):
                T[jst.entry_col[u] - 1 + k] *= s[st.boffsets_r[r] - 1 + k]
        if t < st.neblocks:
            de = st.e_dims[t]
            for k in range(de):
                T[C - de + k] *= s[st.nr + st.boffsets_e[t] - 1 + k]
    return data


def _pow2_scale(rng, n):
    return 2.0 ** rng.integers(-6, 7, n).astype(np.float64)


def test_qr_conditioning_scale_equals_prescaled_columns(blsamd):
    """reduce_qr!(ws, J, sc, lambda, s) factors J S with damping rows sqrt(lambda sc) s_e (src/qr_reduction.jl:
    162-255): with powers of two for s, the same tables as the unconditioned QR of the column-scaled raw tables with
    scaling_e multiplied by s_e^2, bit for bit."""
    import copy
    rng = np.random.default_rng(9)
    for p in (_small_jacobian(blsamd, 4), _jac_with_free_rows_cpu(blsamd)):
        st = p.structure
        s = _pow2_scale(rng, st.nr + st.ne)
        jst = p.jst
        from blsload import load
        sargs = oracle.SchurArgs(st, load().structure.SchurReductionWorkspaceIndex(st, 1))
        jargs = oracle.JacobianArgs(jst)
        F = np.zeros(jst.raw_length()); Fr = np.zeros(jst.residual_length())
        ok, fb = oracle.reduce_qr(sargs, jargs, F, Fr, p.data, p.r, p.scaling_e, 0.2, cs=s)
        q = copy.copy(p)
        q.data = _column_scaled(p, s)
        q.scaling_e = p.scaling_e * s[st.nr:] ** 2
        ok2, fb2, F2, Fr2, _, _ = _qr(q, 0.2)
        assert ok and ok2
        assert np.array_equal(F, F2) and np.array_equal(Fr, Fr2)


def _jac_with_free_rows_cpu(blsamd):
    from test_gpu_qr import _jac_with_free_rows
    return _jac_with_free_rows(blsamd)


def _qr_pattern(blsamd, p):
    """build_schur_pattern with the free tables' block pairs as extra blocks (free_block_pairs,
    src/qr_reduced.jl:7-18)."""
    jst, st = p.jst, p.structure
    extra = []
    for t in range(st.neblocks, jst.ntables):
        u = list(range(jst.table_ptr[t] - 1, jst.table_ptr[t + 1] - 1))
        for b in range(len(u)):
            for a in range(b):
                extra.append((jst.entry_retained[u[a]], jst.entry_retained[u[b]]))
    return blsamd.structure.build_schur_pattern(st, extra_blocks=extra)


def test_reduced_normal_sparse_equals_dense(blsamd):
    """reduced_normal_sparse! writes the same upper blocks as reduced_normal_dense! (same loops, sparse addressing),
    bit for bit, including free tables."""
    from helpers import dense_upper_blocks, sparse_upper_blocks
    for p in (_small_jacobian(blsamd, 3), _jac_with_free_rows_cpu(blsamd)):
        ok, fb, F, Fr, _, jargs = _qr(p, 0.1)
        assert ok
        st = p.structure
        pat = _qr_pattern(blsamd, p)
        sargs = oracle.SchurArgs(st, blsamd.structure.SchurReductionWorkspaceIndex(st, 1), pat)
        Sd, gd = np.zeros(st.nr * st.nr), np.zeros(st.nr)
        oracle.reduced_normal(False, sargs, jargs, Sd, gd, F, Fr)
        Ss, gs = np.zeros(pat.nnz), np.zeros(st.nr)
        oracle.reduced_normal(True, sargs, jargs, Ss, gs, F, Fr)
        A, B = dense_upper_blocks(st, Sd), sparse_upper_blocks(st, pat, Ss)
        for k, blk in B.items():
            assert np.array_equal(A[k], blk), k
        assert np.array_equal(gd, gs)


@pytest.mark.parametrize("free", [False, True])
def test_reduced_jacobian_pattern_and_values(blsamd, free):
    """build_reduced_jacobian_pattern / reduced_jacobian! (src/qr_reduced.jl:247-370), restated in
    oracle/jacobian_ref.py: the CSC invariants (strictly ascending rows per column, nnz = sum over entries of
    d_r (M - top)), J~ v and J~'w equal to the C oracle's reduced_mul / reduced_tmul on the same factored tables, and
    the dense form equal to the sparse one entry for entry (test/schur_qr.jl:180-270)."""
    import scipy.sparse as sp
    import jacobian_ref as JR
    p = _jac_with_free_rows_cpu(blsamd) if free else _small_jacobian(blsamd, 4, nland=50, ncam=7)
    ok, _, F, Fr, sargs, jargs = _qr(p, 0.3)
    assert ok
    jst, st = p.jst, p.structure
    m = int(jst.reduced_nrows())
    colptr, rowval = JR.build_reduced_jacobian_pattern(jst)
    assert len(colptr) == st.nr + 1 and colptr[-1] - 1 == len(rowval)
    assert len(rowval) == sum(int(st.r_dims[jst.entry_retained[u] - 1]) *
                              (jst.table_nrows[jst.entry_table[u] - 1] - jst.table_top[jst.entry_table[u] - 1])
                              for u in range(len(jst.entry_retained)))
    for c in range(st.nr):
        rows = rowval[colptr[c] - 1:colptr[c + 1] - 1]
        assert np.all(np.diff(rows) > 0) and (len(rows) == 0 or (rows[0] >= 1 and rows[-1] <= m))
    nz = JR.reduced_jacobian_sparse(jst, F, colptr)
    Jt = sp.csc_matrix((nz, rowval - 1, colptr - 1), shape=(m, st.nr))
    rng = np.random.default_rng(8)
    v, w = rng.standard_normal(st.nr), rng.standard_normal(m)
    y, g = np.zeros(m), np.zeros(st.nr)
    oracle.reduced_mul(sargs, jargs, y, F, v)
    oracle.reduced_tmul(sargs, jargs, g, F, w)
    assert np.abs(Jt @ v - y).max() <= 1e-12 * np.abs(y).max()
    assert np.abs(Jt.T @ w - g).max() <= 1e-12 * np.abs(g).max()
    assert np.array_equal(JR.reduced_jacobian_dense(jst, F), Jt.toarray())


def _rowblock_values(jst, rng):
    """Random blocks of every row block (vals: row block after row block, blocks in ascending layout column block,
    each m x d column-major), residual rows, and the dense J / r they make (rows in row-block order, columns
    [x_r; x_e])."""
    st, rb = jst.schur, jst.rowblocks
    nr, ne = st.nr, st.ne
    vals, res, Jrows, rrows = [], [], [], []
    for k in range(len(rb["table"])):
        m = int(rb["nrows"][k])
        rows = np.zeros((m, nr + ne))
        for j in rb["blocks"][rb["ptr"][k] - 1: rb["ptr"][k + 1] - 1]:
            j = int(j)
            if j <= st.nrblocks:
                d, o = int(st.r_dims[j - 1]), int(st.boffsets_r[j - 1]) - 1
            else:
                e = j - st.nrblocks
                d, o = int(st.e_dims[e - 1]), nr + int(st.boffsets_e[e - 1]) - 1
            B = rng.standard_normal((m, d))
            vals.append(B.T.reshape(-1))
            rows[:, o:o + d] = B
        rr = rng.standard_normal(m)
        res.append(rr)
        Jrows.append(rows)
        rrows.append(rr)
    return (np.concatenate(vals), np.concatenate(res), np.vstack(Jrows), np.concatenate(rrows))


def test_fill_tables_reproduces_generator_tables(blsamd):
    """fill_tables (block / residual_block assignments, src/schur_jacobian.jl:264-387) from the generator's
    per-observation blocks writes exactly the tables the generator laid out by its own vectorised indexing."""
    import jacobian_ref as JR
    p = blsamd.generators.config4(scale=0.0002, ncam=30, kmax=12)
    jst, st = p.jst, p.structure
    vals, res = [], []
    for k in range(len(p.cam_of_obs)):
        d = int(st.r_dims[p.cam_of_obs[k] - 1])
        vals += [p.Jc[k, :, :d].T.reshape(-1), p.Jl[k].T.reshape(-1)]
        res.append(p.rr[k])
    raw, r = np.zeros(jst.raw_length()), np.zeros(jst.residual_length())
    n, nres = JR.fill_tables(jst, np.concatenate(vals), np.concatenate(res), raw, r)
    assert n == sum(len(v) for v in vals) and nres == 2 * len(p.cam_of_obs)
    assert np.array_equal(raw, p.data) and np.array_equal(r, p.r)


def test_fill_tables_equals_dense_rowblock_jacobian(blsamd):
    """With free tables (camera priors, camera-pair rows) and mixed camera sizes: the filled tables give the J'J and
    J'r of the dense J stacked from the row blocks themselves."""
    import jacobian_ref as JR
    from test_gpu_qr import _jac_with_free_rows
    p = _jac_with_free_rows(blsamd, seed=3)
    jst = p.jst
    assert jst.ntables > p.structure.neblocks
    vals, res, J, rr = _rowblock_values(jst, np.random.default_rng(1))
    p.data, p.r = np.zeros(jst.raw_length()), np.zeros(jst.residual_length())
    JR.fill_tables(jst, vals, res, p.data, p.r)
    Jt, rt = _dense_J(p)
    assert Jt.shape == J.shape
    assert np.allclose(Jt.T @ Jt, J.T @ J, rtol=1e-13, atol=1e-12)
    assert np.allclose(Jt.T @ rt, J.T @ rr, rtol=1e-13, atol=1e-12)
