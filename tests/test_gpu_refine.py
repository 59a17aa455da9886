"""The Float64 iterative refinement of the mixed reduction (no reference counterpart): factor_all_eliminated! on
the device against the reduction's own phase 1, the retained residual with x_e = z against the reduction's g_S,
the matrix-free residual g - S x against the one from the formed S, and the refined solution of the Float32-stored
blocks' system against the Float64 blocks' reduced system: ||g - S x|| / ||g|| <= 1e-8 per problem (the stated
bound), where the unrefined mixed solution stays near the inputs' rounding (~1e-6)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _setup(blsamd, p):
    import torch
    dev = blsamd.device
    st = p.structure
    h = dev.SchurHandle(st, p.ws_index, pattern=p.pattern, r_to_s=p.r_to_s)
    T64 = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64)).cuda()
    T32 = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32)).cuda()
    d = dict(E=T64(p.E_data), C=T64(p.C_data), R=T64(p.R_data), b_r=T64(p.b_r), b_e=T64(p.b_e),
             sr=T64(p.scaling_r), se=T64(p.scaling_e), E32=T32(p.E_data), C32=T32(p.C_data), R32=T32(p.R_data))
    nnz = p.pattern.nnz
    S64 = torch.empty(nnz, dtype=torch.float64, device="cuda")
    g64 = torch.empty(st.nr, dtype=torch.float64, device="cuda")
    ws64 = dev.SchurReductionWorkspace(h)
    dev.reset_schur_sparse_(S64, d["R"], h, d["sr"], p.lam)
    assert dev.reduce_schur_sparse_(S64, g64, d["b_r"], d["b_e"], d["se"], p.lam, h, d["E"], d["C"], ws64)
    Sm = torch.empty(nnz, dtype=torch.float64, device="cuda")
    gm = torch.empty(st.nr, dtype=torch.float64, device="cuda")
    wsm = dev.SchurReductionWorkspace(h, "mixed")
    dev.reset_schur_sparse_(Sm, d["R32"], h, d["sr"], p.lam)
    assert dev.reduce_schur_sparse_(Sm, gm, d["b_r"], d["b_e"], d["se"], p.lam, h, d["E32"], d["C32"], wsm)
    SR = torch.empty(nnz, dtype=torch.float64, device="cuda")
    dev.reset_schur_sparse_(SR, d["R"], h, d["sr"], p.lam)
    return h, d, S64, g64, ws64, Sm, gm, SR


def test_factor_and_retained_residual(blsamd):
    import torch
    dev = blsamd.device
    p = blsamd.generators.config2(scale=0.01)
    h, d, S64, g64, ws64, Sm, gm, SR = _setup(blsamd, p)
    ws = dev.SchurReductionWorkspace(h)
    assert dev.factor_eliminated_(h, ws, d["E"], d["b_e"], d["se"], p.lam)
    assert torch.equal(ws.E_inv_data, ws64.E_inv_data) and torch.equal(ws.z_data, ws64.z_data)
    g = torch.empty_like(g64)
    dev.retained_residual_(g, h, d["b_r"], ws.z_data, d["C"])
    assert float((g - g64).abs().max()) <= 1e-12 * float(g64.abs().max())
    # the matrix-free residual of a random x equals g - S x of the formed S
    refine = blsamd.refine
    pcg = dev.PCGHandle(p.structure, p.pattern)
    rd = refine.RefinementData(h, d["E"], d["C"], d["b_r"], d["b_e"], d["se"], SR, p.lam)
    assert refine.factor(rd)
    x = torch.randn(p.structure.nr, dtype=torch.float64, device="cuda", generator=torch.Generator("cuda").manual_seed(3))
    rel = refine.residual(pcg, rd, x, g64)
    want = pcg.relres(S64, g64, x)
    assert abs(rel[0] - want[0]) <= 1e-10 * want[0]
    # a failing landmark block is reported as the reduction reports it
    E_bad = d["E"].clone()
    E_bad[9:18] = 0.0
    assert not dev.factor_eliminated_(h, ws, E_bad, d["b_e"], d["se"], 0.0)
    assert ws.failed_block == 2


@pytest.mark.parametrize("union", [False, True])
def test_refined_mixed_solve_meets_float64_bound(blsamd, union):
    import torch
    refine = blsamd.refine
    dev = blsamd.device
    if union:
        p = blsamd.generators.config5_union(ncams=[60, 140, 90, 200], nlands=[20000, 30000, 15000, 40000], scale=0.5)
        pcg = dev.PCGHandle(p.structure, p.pattern, p.prob_block_ptr)
    else:
        p = blsamd.generators.config2(scale=0.02)
        pcg = dev.PCGHandle(p.structure, p.pattern)
    h, d, S64, g64, ws64, Sm, gm, SR = _setup(blsamd, p)
    x = torch.zeros_like(g64)
    ok, _, _ = pcg.solve(Sm, gm, x, tol=1e-8)
    assert ok
    unrefined = pcg.relres(S64, g64, x)
    assert unrefined.max() > 1e-8   # the Float32 inputs' rounding: the mixed system alone misses the bound
    rd = refine.RefinementData(h, d["E"], d["C"], d["b_r"], d["b_e"], d["se"], SR, p.lam)
    ok, steps, iters, rel = refine.refined_solve(pcg, Sm, gm, x, rd, tol=1e-8)
    assert ok and steps <= 4
    true_rel = pcg.relres(S64, g64, x)   # against the formed Float64 reduced system
    print(f"unrefined {unrefined.max():.2e}; refined in {steps} steps ({iters.max()} PCG iterations): "
          f"{true_rel.max():.2e}")
    assert true_rel.max() <= 1.0001e-8
