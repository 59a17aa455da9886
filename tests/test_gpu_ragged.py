"""Ragged structures through the storage queries and the table quantities: an eliminated block with no coupling
(observed by its own prior only), a retained block with no coupling (a prior only), mixed block dimensions. The
assembly, dot(x, H, y), getdiag!, the conditioning scale, jtr!, jtj_diag!, jtj_quadform and reduced_jacobian! against
the restatements in oracle/ (tolerances as in their own tests: 1e-12, copies bitwise)."""
import numpy as np
import pytest

import assembly_ref as A
import conditioning_ref as C
import jacobian_ref as JR

pytestmark = pytest.mark.gpu

# caller blocks: 1 landmark (3) seen only by its prior, 2 landmark (3) seen by camera 4, 3 landmark (2) seen by
# cameras 4 and 5, 4 camera (6), 5 camera (9), 6 camera (6) with a prior only
BLOCKS = [3, 3, 2, 6, 9, 6]
ELIM = [True, True, True, False, False, False]
ROWS = [([1], 3), ([2, 4], 2), ([3, 4], 2), ([3, 5], 2), ([4], 6), ([5], 9), ([6], 6), ([4, 5], 3)]


def _T(a, dtype=None):
    import torch
    t = torch.as_tensor(np.ascontiguousarray(a)).cuda()
    return t if dtype is None else t.to(dtype)


@pytest.mark.parametrize("sparse", [False, True])
def test_ragged_hessian_queries(blsamd, sparse):
    import torch
    dev = blsamd.device
    st = blsamd.structure.SchurStructure.from_description(BLOCKS, [sorted(r) for r, _ in ROWS], ELIM)
    layout = np.argsort(np.asarray(st.blockorder()) - 1) + 1
    rowblocks = [[int(layout[b - 1]) for b in r] for r, _ in ROWS]
    rb_rows = np.array([m for _, m in ROWS])
    dims = list(st.r_dims) + list(st.e_dims)
    rng = np.random.default_rng(3)
    J = rng.standard_normal(int(sum(m * sum(int(dims[b - 1]) for b in rb) for rb, m in zip(rowblocks, rb_rows))))
    r = rng.standard_normal(int(rb_rows.sum()))
    R_layout = None
    if sparse:   # the diagonal blocks and the pair (4, 5) that shares a row block, lower blocks as a block CSC
        nrb = st.nrblocks
        pairs = {(q, q) for q in range(1, nrb + 1)}
        for rb in rowblocks:
            ret = [b for b in rb if b <= nrb]
            pairs |= {(a, b) for a in ret for b in ret if a >= b}
        colptr, rowval, offs, off = [1], [], [], 1
        for q in range(1, nrb + 1):
            for p_ in sorted(p for (p, qq) in pairs if qq == q):
                rowval.append(p_); offs.append(off); off += int(st.r_dims[p_ - 1] * st.r_dims[q - 1])
            colptr.append(len(rowval) + 1)
        R_layout = (np.array(colptr), np.array(rowval), np.array(offs))
    h = dev.AssemblyHandle(st, rowblocks, rb_rows, R_layout)
    nan = lambda n: torch.full((n,), float("nan"), dtype=torch.float64, device="cuda")
    ne = int(np.sum(st.e_dims))
    E, Cd, R, br, be = nan(h.E_len), nan(h.C_len), nan(h.R_len), nan(h.nr), nan(ne)
    dev.assemble_hessian_(E, Cd, R, br, be, h, _T(J), _T(r))
    H = A.assemble(A.HessianStorage(st, *(R_layout or ())), rowblocks, rb_rows, J, r)
    Ro = H.R.reshape(-1, order="F") if R_layout is None else H.R
    for name, a, b in (("E", E, H.E), ("C", Cd, H.C), ("R", R, Ro), ("b_r", br, H.b_r), ("b_e", be, H.b_e)):
        a = a.cpu().numpy()
        assert not np.isnan(a).any(), name
        assert np.abs(a - b).max() <= 1e-12 * max(1.0, np.abs(b).max()), name
    n = h.nr + ne
    x, y = rng.standard_normal(n), rng.standard_normal(n)
    Hd = A.HessianStorage(st, *(R_layout or ()))
    Hd.E, Hd.C = E.cpu().numpy(), Cd.cpu().numpy()
    Hd.R = R.cpu().numpy().reshape(h.nr, h.nr, order="F") if R_layout is None else R.cpu().numpy()
    ref = A.bilinear(Hd, x, y)
    assert abs(dev.hessian_dot(_T(x), h, _T(y), E, Cd, R) - ref) <= 1e-12 * max(1.0, abs(ref))
    d = nan(n)
    dev.hessian_diag_(d, h, E, R)
    dref = A.getdiag(Hd)
    assert np.array_equal(d.cpu().numpy(), dref)
    c = dev.EliminationConditioning(n)
    s = dev.prepare_scale_(c, lambda dd: dev.hessian_diag_(dd, h, E, R), 0.25, None)
    assert np.array_equal(s.cpu().numpy(), C.conditioning_scale(dref + 0.25))


def test_ragged_table_quantities(blsamd):
    import torch
    dev = blsamd.device
    jst = blsamd.structure.SchurJacobianStructure.from_description(BLOCKS, [r for r, _ in ROWS], [m for _, m in ROWS],
                                                                    ELIM)
    st = jst.schur
    rng = np.random.default_rng(4)
    data = rng.standard_normal(jst.raw_length())
    for t in range(st.neblocks):           # the damping rows of the raw eliminated tables are zero
        M, Cc, top = jst.table_nrows[t], jst.table_ncols[t] - 1, jst.table_top[t]
        data[jst.raw_offsets[t] - 1:jst.raw_offsets[t] - 1 + M * Cc].reshape(Cc, M)[:, M - top:] = 0.0
    res = rng.standard_normal(jst.residual_length())
    h = dev.QRHandle(jst)
    F, rr = _T(data), _T(res)
    n = st.nr + st.ne
    v = rng.standard_normal(n)
    g = torch.full((n,), float("nan"), dtype=torch.float64, device="cuda")
    dg = torch.full((n,), float("nan"), dtype=torch.float64, device="cuda")
    dev.jtr_(g, h, F, rr)
    dev.jtj_diag_(dg, h, F)
    q = dev.jtj_quadform(h, F, _T(v))
    go, do, qo = JR.jtr(jst, data, res), JR.jtj_diag(jst, data), JR.jtj_quadform(jst, data, v)
    assert np.abs(g.cpu().numpy() - go).max() <= 1e-12 * np.abs(go).max()
    assert np.abs(dg.cpu().numpy() - do).max() <= 1e-12 * np.abs(do).max()
    assert abs(q - qo) <= 1e-12 * qo
    ws = dev.QRReductionWorkspace(h)
    assert dev.reduce_qr_(ws, h, F, rr, _T(np.ones(st.ne)), 0.5)
    m = int(jst.reduced_nrows())
    colptr, rowval = dev.build_reduced_jacobian_pattern(h)
    co, ro = JR.build_reduced_jacobian_pattern(jst)
    assert np.array_equal(colptr, co) and np.array_equal(rowval, ro)
    nz = torch.full((len(rowval),), float("nan"), dtype=torch.float64, device="cuda")
    rt = torch.full((m,), float("nan"), dtype=torch.float64, device="cuda")
    dev.reduced_jacobian_(nz, rt, h, ws, sparse=True)
    assert np.array_equal(nz.cpu().numpy(), JR.reduced_jacobian_sparse(jst, ws.factored.cpu().numpy(), co))
