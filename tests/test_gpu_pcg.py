"""GPU block-Jacobi PCG on reduced systems produced by the GPU Schur path: converges to the requested true
residual (FP64), agrees with a direct solve, handles batches of independent problems, and is deterministic.
(No reference counterpart: pinned by the residual, the direct solve and the oracle PCG's iteration count.)"""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


def _reduce(blsamd, p):
    import torch
    dev = blsamd.device
    st = p.structure
    h = dev.SchurHandle(st, p.ws_index, pattern=p.pattern, r_to_s=p.r_to_s)
    ws = dev.SchurReductionWorkspace(h)
    T = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64)).cuda()
    S = torch.empty(p.pattern.nnz, dtype=torch.float64, device="cuda")
    g = torch.empty(st.nr, dtype=torch.float64, device="cuda")
    dev.reset_schur_sparse_(S, T(p.R_data), h, T(p.scaling_r), p.lam)
    assert dev.reduce_schur_sparse_(S, g, T(p.b_r), T(p.b_e), T(p.scaling_e), p.lam, h, T(p.E_data), T(p.C_data), ws)
    return S, g


def _dense(p, S):
    st, pat = p.structure, p.pattern
    M = np.zeros((st.nr, st.nr))
    for j in range(st.nr):
        for k in range(pat.S_colptr[j] - 1, pat.S_colptr[j + 1] - 1):
            M[pat.S_rowval[k] - 1, j] = S[k]
    return np.triu(M) + np.triu(M, 1).T


def test_pcg_single_problem(blsamd):
    import torch
    p = blsamd.generators.config2(scale=0.01)
    S, g = _reduce(blsamd, p)
    pcg = blsamd.device.PCGHandle(p.structure, p.pattern)
    x = torch.zeros_like(g)
    ok, iters, rel = pcg.solve(S, g, x, tol=1e-8)
    assert ok and rel[0] <= 1e-8 and iters[0] > 0
    Sd = _dense(p, S.cpu().numpy())
    xd = np.linalg.solve(Sd, g.cpu().numpy())
    xg = x.cpu().numpy()
    assert np.linalg.norm(Sd @ xg - g.cpu().numpy()) <= 1.0001e-8 * np.linalg.norm(g.cpu().numpy())
    assert np.linalg.norm(xg - xd) <= 1e-5 * np.linalg.norm(xd)
    # the oracle PCG (same algorithm, same stopping rule) needs about as many iterations
    args = oracle.SchurArgs(p.structure, p.ws_index, p.pattern)
    xo, ito, relo = oracle.pcg(p.structure, args, S.cpu().numpy(), g.cpu().numpy(), tol=1e-8)
    assert abs(ito - iters[0]) <= max(2, 0.1 * ito)
    # relres: the true residual of a given x, against the host's dense product
    xr = torch.as_tensor(np.random.default_rng(1).standard_normal(p.structure.nr)).cuda()
    gh = g.cpu().numpy()
    want = np.linalg.norm(Sd @ xr.cpu().numpy() - gh) / np.linalg.norm(gh)
    assert abs(pcg.relres(S, g, xr)[0] - want) <= 1e-12 * want
    assert pcg.relres(S, g, x)[0] <= 1.0001e-8
    # deterministic
    x2 = torch.zeros_like(g)
    ok2, iters2, rel2 = pcg.solve(S, g, x2, tol=1e-8)
    assert np.array_equal(x2.cpu().numpy(), xg) and iters2[0] == iters[0]


def test_pcg_batched_problems(blsamd):
    """Three independent problems stacked block-diagonally, solved in one batch."""
    import torch
    S_ = blsamd.structure
    probs = [blsamd.generators.config5_problem(i, scale=0.02, ncam=[30, 55, 41][i]) for i in range(3)]
    # union structure: blocks of problem i after those of problem i-1
    st_list = [q.structure for q in probs]
    r_dims = np.concatenate([s.r_dims for s in st_list])
    e_dims = np.concatenate([s.e_dims for s in st_list])
    cpl, roff = [], 0
    for s in st_list:
        for e in range(s.neblocks):
            cpl.append((s.entry_retained[s.eliminated_ptr[e] - 1: s.eliminated_ptr[e + 1] - 1] + roff).tolist())
        roff += s.nrblocks
    st = S_.SchurStructure(r_dims, e_dims, cpl)
    pat = S_.build_schur_pattern(st)
    ws = S_.SchurReductionWorkspaceIndex(st, 1)
    Rs = S_.hessian_retained_storage(st, pat)
    R = np.zeros(Rs.nnzvals)
    boff = 0
    for q in probs:
        for r in range(q.structure.nrblocks):
            s0 = Rs.block_start(boff + r + 1, boff + r + 1)
            d = q.structure.r_dims[r]
            R[s0 - 1:s0 - 1 + d * d] = q.Rblocks[r, :d, :d].T.reshape(-1)
        boff += q.structure.nrblocks

    class U:
        pass
    u = U()
    u.structure, u.pattern, u.ws_index, u.R_data = st, pat, ws, R
    u.r_to_s = S_.build_r_to_s_map(Rs, pat, st.boffsets_r)
    u.E_data = np.concatenate([q.E_data for q in probs])
    u.C_data = np.concatenate([q.C_data for q in probs])
    u.b_r = np.concatenate([q.b_r for q in probs])
    u.b_e = np.concatenate([q.b_e for q in probs])
    u.scaling_r, u.scaling_e, u.lam = np.ones(st.nr), np.ones(st.ne), 1e-4
    S, g = _reduce(blsamd, u)
    pbp = np.cumsum([1] + [q.structure.nrblocks for q in probs])
    pcg = blsamd.device.PCGHandle(st, pat, pbp)
    x = torch.zeros_like(g)
    ok, iters, rel = pcg.solve(S, g, x, tol=1e-8)
    assert ok and (rel <= 1e-8).all() and (iters > 0).all()
    # each problem alone gives the same iterates (independent problems; per-problem reductions)
    for i, q in enumerate(probs):
        Sq, gq = _reduce(blsamd, q)
        xq = torch.zeros_like(gq)
        okq, itq, relq = blsamd.device.PCGHandle(q.structure, q.pattern).solve(Sq, gq, xq, tol=1e-8)
        assert okq and itq[0] == iters[i]
        a, b = st.boffsets_r[pbp[i] - 1] - 1, st.boffsets_r[pbp[i + 1] - 2] - 1 + st.r_dims[pbp[i + 1] - 2]
        assert np.allclose(x.cpu().numpy()[a:b], xq.cpu().numpy(), rtol=1e-9, atol=1e-12 * np.abs(xq.cpu().numpy()).max())


def test_pcg_mixed_block_dims(blsamd):
    """PCG on a reduced system whose camera blocks are 6 and 9 wide (the product with block dims that differ between
    block rows): S = J~'J~ + lambda I from the QR path on the device; true residual, a dense solve, relres against
    the host's product, and determinism."""
    import torch
    from test_oracle_qr import _small_jacobian, _qr_pattern
    dev = blsamd.device
    p = _small_jacobian(blsamd, 21, nland=300, ncam=12, kmax=6)
    st = p.structure
    assert len(set(int(d) for d in st.r_dims)) == 2
    h = dev.QRHandle(p.jst)
    pat = _qr_pattern(blsamd, p)
    h.set_sparse(pat)
    ws = dev.QRReductionWorkspace(h)
    T = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64)).cuda()
    assert dev.reduce_qr_(ws, h, T(p.data), T(p.r), T(p.scaling_e), 0.1)
    S = torch.empty(pat.nnz, dtype=torch.float64, device="cuda")
    g = torch.empty(st.nr, dtype=torch.float64, device="cuda")
    dev.reduced_normal_sparse_(S, g, h, ws)
    S[torch.as_tensor(pat.diag_inds - 1).cuda()] += 0.1
    pcg = dev.PCGHandle(st, pat)
    x = torch.zeros_like(g)
    ok, iters, rel = pcg.solve(S, g, x, tol=1e-8, maxit=2000)
    assert ok and rel[0] <= 1e-8
    M = np.zeros((st.nr, st.nr))
    Sh = S.cpu().numpy()
    for j in range(st.nr):
        for k in range(pat.S_colptr[j] - 1, pat.S_colptr[j + 1] - 1):
            M[pat.S_rowval[k] - 1, j] = Sh[k]
    Sd = np.triu(M) + np.triu(M, 1).T
    gh, xg = g.cpu().numpy(), x.cpu().numpy()
    xd = np.linalg.solve(Sd, gh)
    assert np.linalg.norm(Sd @ xg - gh) <= 1.0001e-8 * np.linalg.norm(gh)
    assert np.linalg.norm(xg - xd) <= 1e-5 * np.linalg.norm(xd)
    xr = torch.as_tensor(np.random.default_rng(3).standard_normal(st.nr)).cuda()
    want = np.linalg.norm(Sd @ xr.cpu().numpy() - gh) / np.linalg.norm(gh)
    assert abs(pcg.relres(S, g, xr)[0] - want) <= 1e-12 * want
    x2 = torch.zeros_like(g)
    pcg.solve(S, g, x2, tol=1e-8, maxit=2000)
    assert np.array_equal(x2.cpu().numpy(), xg)
