"""The restatement of addblock! (oracle/assembly_ref.py, src/schur_assembly.jl:140-246), pinned by the reference
docstring's example and by the dense J'J of the stacked row blocks read back through E_data, C_data and R."""
import numpy as np
import pytest

import assembly_ref as A
from helpers import assembly_problem


def test_docstring_example(blsamd):
    """src/schur_assembly.jl:126-137: a point (3) and a camera (6) sharing one row block, B = ones(9, 9), the three
    addblock! calls of the example: sum(E_data), sum(C_data), sum(R) == (9, 18, 36)."""
    st = blsamd.structure.SchurStructure.from_description([3, 6], [[1, 2]], [True, False])
    p = np.argsort(np.asarray(st.blockorder()) - 1) + 1
    H = A.HessianStorage(st)
    B = np.ones((9, 9))
    A.addblock(H, p[0], p[0], B, list(range(0, 3)), list(range(0, 3)))
    A.addblock(H, p[1], p[1], B, list(range(3, 9)), list(range(3, 9)))
    A.addblock(H, p[1], p[0], B, list(range(3, 9)), list(range(0, 3)))
    assert (H.E.sum(), H.C.sum(), H.R.sum()) == (9.0, 18.0, 36.0)


def _dense_J(st, rowblocks, rb_rows, J, r):
    """The stacked Jacobian over [x_r; x_e] (layout order) and residual."""
    nr, ne = int(np.sum(st.r_dims)), int(np.sum(st.e_dims))
    dims = list(st.r_dims) + list(st.e_dims)
    start = lambda b: st.boffsets_r[b - 1] - 1 if b <= st.nrblocks else nr + st.boffsets_e[b - st.nrblocks - 1] - 1
    Jd = np.zeros((int(np.sum(rb_rows)), nr + ne))
    joff = row = 0
    for blocks, m in zip(rowblocks, rb_rows):
        W = sum(int(dims[b - 1]) for b in blocks)
        Jk = J[joff:joff + m * W].reshape(W, m).T
        c = 0
        for b in blocks:
            d = int(dims[b - 1])
            Jd[row:row + m, start(b):start(b) + d] = Jk[:, c:c + d]
            c += d
        joff += m * W
        row += m
    return Jd, r


@pytest.mark.parametrize("sparse", [False, True])
def test_oracle_matches_dense_JtJ(blsamd, sparse):
    """H from the addblock! restatement equals J'J of the stacked Jacobian in every stored block (E full, C_er
    d_e x d_r, R upper with full diagonal blocks or stored lower blocks), and b = J'r."""
    st, rowblocks, rb_rows, J, r, R_layout = assembly_problem(blsamd, 3, sparse=sparse)
    H = A.assemble(A.HessianStorage(st, *(R_layout or ())), rowblocks, rb_rows, J, r)
    Jd, rd_ = _dense_J(st, rowblocks, rb_rows, J, r)
    M, g = Jd.T @ Jd, Jd.T @ rd_
    nr = int(np.sum(st.r_dims))
    tol = 1e-12 * np.abs(M).max()
    for e in range(st.neblocks):
        d, o = int(st.e_dims[e]), nr + st.boffsets_e[e] - 1
        Eb = H.E[st.eoffsets[e] - 1:st.eoffsets[e] - 1 + d * d].reshape(d, d).T
        assert np.abs(Eb - M[o:o + d, o:o + d]).max() <= tol
        for c in range(st.eliminated_ptr[e] - 1, st.eliminated_ptr[e + 1] - 1):
            rr = st.entry_retained[c]
            dr, ro = int(st.r_dims[rr - 1]), st.boffsets_r[rr - 1] - 1
            Cb = H.C[st.coffsets[c] - 1:st.coffsets[c] - 1 + d * dr].reshape(dr, d).T   # d_e x d_r column-major
            assert np.abs(Cb - M[o:o + d, ro:ro + dr]).max() <= tol
    rdim, bo = st.r_dims, st.boffsets_r - 1
    if R_layout is None:
        for q in range(st.nrblocks):
            for p in range(q + 1):
                blk = H.R[bo[p]:bo[p] + rdim[p], bo[q]:bo[q] + rdim[q]]
                assert np.abs(blk - M[bo[p]:bo[p] + rdim[p], bo[q]:bo[q] + rdim[q]]).max() <= tol
        blk_of = np.searchsorted(bo, np.arange(nr), side="right") - 1   # the block of each scalar index
        assert np.all(H.R[blk_of[:, None] > blk_of[None, :]] == 0)         # nothing below the block diagonal
    else:
        colptr, rowval, offs = R_layout
        for q in range(st.nrblocks):
            for k in range(colptr[q] - 1, colptr[q + 1] - 1):
                p = rowval[k] - 1
                blk = H.R[offs[k] - 1:offs[k] - 1 + rdim[p] * rdim[q]].reshape(rdim[q], rdim[p]).T
                assert np.abs(blk - M[bo[p]:bo[p] + rdim[p], bo[q]:bo[q] + rdim[q]]).max() <= tol
    assert np.abs(H.b_r - g[:nr]).max() <= 1e-12 * np.abs(g).max()
    assert np.abs(H.b_e - g[nr:]).max() <= 1e-12 * np.abs(g).max()


def test_oracle_rejects_pairs_without_storage(blsamd):
    """check_addblock (src/schur_assembly.jl:161-181): two eliminated blocks, or a retained and an eliminated block
    that share no row block, have no storage."""
    st = blsamd.structure.SchurStructure.from_description([3, 3, 6, 6], [[1, 3], [2, 4]], [True, True, False, False])
    p = np.argsort(np.asarray(st.blockorder()) - 1) + 1
    H = A.HessianStorage(st)
    B = np.ones((18, 18))
    with pytest.raises(ValueError, match="both eliminated"):
        A.addblock(H, p[0], p[1], B, list(range(3)), list(range(3, 6)))
    with pytest.raises(ValueError, match="share no row block"):
        A.addblock(H, p[0], p[3], B, list(range(3)), list(range(6, 12)))


@pytest.mark.parametrize("sparse", [False, True])
def test_bilinear_matches_dense(blsamd, sparse):
    """schur_bilinear's restatement (src/linalg.jl:145-260) on the assembled storage equals x' (J'J) y, and x' H x
    for x == y (schur_quadform's value)."""
    st, rowblocks, rb_rows, J, r, R_layout = assembly_problem(blsamd, 4, sparse=sparse)
    H = A.assemble(A.HessianStorage(st, *(R_layout or ())), rowblocks, rb_rows, J, r)
    Jd, _ = _dense_J(st, rowblocks, rb_rows, J, r)
    M = Jd.T @ Jd
    rng = np.random.default_rng(1)
    x, y = rng.standard_normal(M.shape[0]), rng.standard_normal(M.shape[0])
    for a, b in ((x, y), (x, x)):
        want = a @ M @ b
        assert abs(A.bilinear(H, a, b) - want) <= 1e-12 * np.abs(M).sum() * max(1.0, np.abs(a).max() * np.abs(b).max())
