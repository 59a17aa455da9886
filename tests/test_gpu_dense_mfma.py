"""reduce_schur_dense! with phase 2 as a SYRK on the matrix cores (bls_reduce_schur_dense_mfma_{f64,mixed}; the
dense configuration's FP32-MFMA SYRK with Float64 cross-slice accumulation), through the C ABI against the CPU
oracle on the same inputs. The landmark blocks E and their factorisation are Float64 in both variants; the mixed
one takes the couplings C in Float32.

Tolerances. FP64 MFMA: S within 1e-12 of the oracle per upper block, relative to the magnitude the block is summed
from (|R| + sum |C_1|' |E^-1| |C_2|: the dense configuration's 1e2 / 1e-2 camera scales cancel summands, and two
orders of exact-arithmetic-equivalent sums differ by up to ~3e-9 relative to the result there); g, E_inv, z within
1e-12. Mixed (Float32 couplings, FP32 MFMA): S against the oracle's Float64 reduction of the same couplings within
2e-5 of that magnitude per block -- the products and each slice's tile sums are Float32 (eps 6e-8) over up to a few
thousand rows -- and g, E_inv, z (Float64 from the Float32 blocks) within 1e-12; after one Float64 refinement step
the dense solve's residual against the Float64 blocks' system is within 1e-10 (unrefined: ~1e-5 and worse)."""
import numpy as np
import pytest

import oracle
from helpers import (dense_magnitude_blocks, dense_upper_blocks, max_block_rel_to, problem_from_dense,
                     reference_schur_problem, schur_test_problem)

pytestmark = pytest.mark.gpu


def _rounded(p):
    class Q:
        pass
    q = Q()
    q.__dict__.update(p.__dict__)
    q.C_data = p.C_data.astype(np.float32).astype(np.float64)
    return q


def _run(blsamd, p, variant, nslices=0, lam=None):
    import torch
    dev = blsamd.device
    st = p.structure
    lam = p.lam if lam is None else lam
    h = dev.SchurHandle(st, p.ws_index, dense=True)
    ws = dev.SchurReductionWorkspace(h)
    dt = torch.float64 if variant == "f64" else torch.float32
    T = lambda a, d=torch.float64: torch.as_tensor(np.ascontiguousarray(a)).to(d).cuda()
    S = torch.empty(st.nr * st.nr, dtype=torch.float64, device="cuda")
    g = torch.full((st.nr,), float("nan"), dtype=torch.float64, device="cuda")
    dev.reset_schur_dense_(S, T(p.R.T.reshape(-1)), h, T(p.scaling_r), lam)
    ok = dev.reduce_schur_dense_mfma_(S, g, T(p.b_r), T(p.b_e), T(p.scaling_e), lam, h, T(p.E_data),
                                      T(p.C_data, dt), ws, nslices=nslices)
    torch.cuda.synchronize()
    return dict(S=S.cpu().numpy(), g=g.cpu().numpy(), E_inv=ws.E_inv_data.cpu().numpy(), z=ws.z_data.cpu().numpy(),
                ok=ok, failed_block=ws.failed_block)


def _compare(p, gpu, orc, tol_S, tol=1e-12):
    st = p.structure
    assert gpu["ok"] and orc["ok"]
    mag = dense_magnitude_blocks(p)
    err = max_block_rel_to(dense_upper_blocks(st, gpu["S"]), dense_upper_blocks(st, orc["S"]), mag)
    assert err <= tol_S, err
    for k in ("g", "E_inv", "z"):
        assert np.abs(gpu[k] - orc[k]).max() <= tol * max(1.0, np.abs(orc[k]).max()), k
    return err


@pytest.mark.parametrize("nslices", [0, 1, 7])
def test_mfma_f64_config3_shape(blsamd, nslices):
    """256 cameras (S 1536 x 1536: 78 upper tiles), k ~ U{2..64}, per-camera scales; several slice counts."""
    p = blsamd.generators.config3(scale=0.003)
    _compare(p, _run(blsamd, p, "f64", nslices), oracle.schur_reduce_problem(p), 1e-12)


def test_mfma_f64_small_and_mixed_dims(blsamd):
    """S smaller than one tile (config 1: 60 x 60), and the reference's hand-built structure (retained dims
    6/3/6/9, eliminated 3 and 1: tiles cut through blocks)."""
    p = blsamd.generators.config1()
    _compare(p, _run(blsamd, p, "f64"), oracle.schur_reduce_problem(p), 1e-12)
    rng = np.random.default_rng(11)
    r_dims, e_dims, cpl = reference_schur_problem()
    H, b = schur_test_problem(rng, r_dims, e_dims, cpl, rr_pairs=[(1, 2), (2, 3)])
    q = problem_from_dense(blsamd, H, b, r_dims, e_dims, cpl, 0.3, True, rr_pairs=[(1, 2), (2, 3)])
    _compare(q, _run(blsamd, q, "f64"), oracle.schur_reduce_problem(q), 1e-12)


def test_mfma_mixed_config3_shape(blsamd):
    p = blsamd.generators.config3(scale=0.003)
    q = _rounded(p)
    err = _compare(q, _run(blsamd, p, "mixed"), oracle.schur_reduce_problem(q), 2e-5)
    print(f"mixed MFMA: S within {err:.2e} of the magnitude per block")


def test_mfma_failure_leaves_S(blsamd):
    """A singular landmark block: BLS_FAILED with the lowest failing block; S keeps the reset and g_S is not
    written (the reference returns false before g_S)."""
    p = blsamd.generators.config1()
    p.E_data[p.structure.eoffsets[4] - 1: p.structure.eoffsets[4] - 1 + 9] = 0.0
    out = _run(blsamd, p, "f64", lam=0.0)
    assert not out["ok"] and out["failed_block"] == 5
    assert np.array_equal(out["S"].reshape(p.structure.nr, p.structure.nr).T[np.triu_indices(p.structure.nr)],
                          p.R[np.triu_indices(p.structure.nr)])
    assert np.all(np.isnan(out["g"]))


def test_mfma_mixed_refined_dense_solve(blsamd):
    """The dense configuration's shape solved through the Cholesky factor of the mixed (FP32 MFMA) S, with one
    Float64 refinement step (the residual from the Float64 blocks): the result's residual against the Float64
    blocks' reduced system (formed by the oracle) is below 1e-10; without the step it is not."""
    import torch
    dev = blsamd.device
    refine = blsamd.refine
    p = blsamd.generators.config3(scale=0.003)
    st = p.structure
    orc = oracle.schur_reduce_problem(p)
    S64 = orc["S"].reshape(st.nr, st.nr).T
    S64 = np.triu(S64) + np.triu(S64, 1).T
    g64 = orc["g"]
    h = dev.SchurHandle(st, p.ws_index, dense=True)
    T = lambda a, d=torch.float64: torch.as_tensor(np.ascontiguousarray(a)).to(d).cuda()
    Sm = torch.empty(st.nr * st.nr, dtype=torch.float64, device="cuda")
    gm = torch.empty(st.nr, dtype=torch.float64, device="cuda")
    Rv, sr = T(p.R.T.reshape(-1)), T(p.scaling_r)
    dev.reset_schur_dense_(Sm, Rv, h, sr, p.lam)
    wsm = dev.SchurReductionWorkspace(h)
    assert dev.reduce_schur_dense_mfma_(Sm, gm, T(p.b_r), T(p.b_e), T(p.scaling_e), p.lam, h, T(p.E_data),
                                        T(p.C_data, torch.float32), wsm)
    SR = torch.empty_like(Sm)
    dev.reset_schur_dense_(SR, T(p.R.T.reshape(-1)), h, sr, p.lam)
    rd = refine.RefinementData(h, T(p.E_data), T(p.C_data), T(p.b_r), T(p.b_e), T(p.scaling_e), SR, p.lam)
    assert refine.factor(rd)
    chol = dev.DenseCholesky(st.nr)
    assert chol.factor(Sm)[0]
    x = torch.empty(st.nr, dtype=torch.float64, device="cuda")
    rel = lambda xv: np.linalg.norm(S64 @ xv - g64) / np.linalg.norm(g64)
    refine.refined_dense_solve(chol, gm, x, rd, steps=0)
    unrefined = rel(x.cpu().numpy())
    refine.refined_dense_solve(chol, gm, x, rd, steps=1)
    refined = rel(x.cpu().numpy())
    print(f"dense mixed MFMA solve: residual against the Float64 system {unrefined:.2e} unrefined, {refined:.2e} "
          f"after one step")
    assert unrefined > 1e-10 and refined <= 1e-10
