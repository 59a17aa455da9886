"""The entry point that allocates nothing (bls_reduce_schur_sparse_f64_noalloc; no counterpart in the reference):
its results against the allocating entry point it varies (bls_reduce_schur_sparse_f64_async) on two calls in a
row, the second on a smaller problem, both with split items so that the static scratch's row sums are used; and
its clear error when a work list needs more scratch than the library holds. The two entry points run the same
kernels in the same order, so their results must agree bit for bit (and with the oracle within 1e-12 per
block)."""
import numpy as np
import pytest

import oracle
from helpers import max_block_rel, sparse_upper_blocks

pytestmark = pytest.mark.gpu


def _inputs(torch, p):
    T = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64)).cuda()
    return dict(E=T(p.E_data), C=T(p.C_data), b_r=T(p.b_r), b_e=T(p.b_e), sr=T(p.scaling_r), se=T(p.scaling_e),
                R=T(p.R_data))


def _run(torch, dev, p, h, x, noalloc):
    st = p.structure
    ws = dev.SchurReductionWorkspace(h)
    S = torch.full((p.pattern.nnz,), float("nan"), dtype=torch.float64, device="cuda")
    g = torch.full((st.nr,), float("nan"), dtype=torch.float64, device="cuda")
    dev.reset_schur_sparse_(S, x["R"], h, x["sr"], p.lam)
    if noalloc:
        dev.reduce_schur_sparse_noalloc_(S, g, x["b_r"], x["b_e"], x["se"], p.lam, h, x["E"], x["C"], ws)
    else:
        dev.reduce_schur_sparse_(S, g, x["b_r"], x["b_e"], x["se"], p.lam, h, x["E"], x["C"], ws, sync=False)
    torch.cuda.synchronize()
    return (S.cpu().numpy(), g.cpu().numpy(), ws.E_inv_data.cpu().numpy(), ws.z_data.cpu().numpy(),
            int(ws.d_failed.item()))


def test_noalloc_matches_allocating_entry_point(blsamd):
    import torch
    dev = blsamd.device
    for scale in (0.01, 0.004):   # the second call on a smaller problem
        p = blsamd.generators.config2(scale=scale, seed=21)
        st = p.structure
        h = dev.SchurHandle(st, p.ws_index, pattern=p.pattern, r_to_s=p.r_to_s)
        h.set_tuning(32, 96)   # 32 landmarks per part: most items are split and merged
        x = _inputs(torch, p)
        S1, g1, Ei1, z1, f1 = _run(torch, dev, p, h, x, noalloc=True)
        S0, g0, Ei0, z0, f0 = _run(torch, dev, p, h, x, noalloc=False)
        assert f1 == 0 and f0 == 0
        for a, b in ((S1, S0), (g1, g0), (Ei1, Ei0), (z1, z0)):
            assert np.array_equal(a, b)
        ref = oracle.schur_reduce_problem(p)
        assert max_block_rel(sparse_upper_blocks(st, p.pattern, S1), sparse_upper_blocks(st, p.pattern, ref["S"])) <= 1e-12
        assert np.abs(g1 - ref["g"]).max() <= 1e-12 * np.abs(ref["g"]).max()


def test_noalloc_reports_failure_and_rejects_oversized_work_list(blsamd):
    import torch
    dev = blsamd.device
    # a singular landmark block: the lowest failing block comes back through the static flag, as from the
    # allocating entry point
    p = blsamd.generators.config2(scale=0.003, seed=22)
    p.E_data[:9] = 0.0
    p.lam = 0.0
    h = dev.SchurHandle(p.structure, p.ws_index, pattern=p.pattern, r_to_s=p.r_to_s)
    x = _inputs(torch, p)
    assert _run(torch, dev, p, h, x, noalloc=True)[4] == 1
    assert _run(torch, dev, p, h, x, noalloc=False)[4] == 1
    # a work list of 16-landmark parts over 40 000 landmarks needs far more than the 16 MiB the library holds
    q = blsamd.generators.config2(scale=0.2, seed=23)
    hq = dev.SchurHandle(q.structure, q.ws_index, pattern=q.pattern, r_to_s=q.r_to_s)
    hq.set_tuning(16, 96)
    xq = _inputs(torch, q)
    with pytest.raises(dev.BLSError, match="scratch"):
        _run(torch, dev, q, hq, xq, noalloc=True)
