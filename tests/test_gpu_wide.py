"""GPU parity of the run-time-size route: eliminated blocks above 8 and retained blocks above 16 (the reference's
dynamic path, src/schur_reduction.jl:164-225 phase 1, :262-354 phase 2, :578-614 backsubstitute!). A handle that
holds such a block reduces its whole instance with the wide kernels.

Tolerances as in test_gpu_schur.py: 1e-12 relative Frobenius per block of S and relative on g_S / E_inv / z / x_e
against the oracle on the same inputs; the full solve within the reference test's 1e-8 of the dense solve
(test/schur_problems.jl:28-40); the failing block exact. The f32 / mixed variants to the bars of test_gpu_precision.py.
"""
import numpy as np
import pytest

import oracle
from helpers import max_block_rel, problem_from_dense, schur_test_problem
from test_gpu_precision import _blocks, _check_f32, _check_mixed, _oracle, run_variant
from test_gpu_schur import _torch, compare, run_gpu

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _backsub(blsamd, p, gpu, x_r, s_r=None, s_e=None):
    torch = _torch()
    T = lambda a: None if a is None else torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64)).cuda()
    x_e = torch.full((p.structure.ne,), float("nan"), dtype=torch.float64, device="cuda")
    blsamd.device.backsubstitute_(x_e, T(x_r), gpu["h"], gpu["C"], gpu["ws"], s_r=T(s_r), s_e=T(s_e))
    torch.cuda.synchronize()
    return x_e.cpu().numpy()


def _sym(p, S):
    st = p.structure
    if p.dense:
        M = S.reshape(st.nr, st.nr).T
    else:
        pat = p.pattern
        M = np.zeros((st.nr, st.nr))
        for j in range(st.nr):
            for k in range(pat.S_colptr[j] - 1, pat.S_colptr[j + 1] - 1):
                M[pat.S_rowval[k] - 1, j] = S[k]
    return np.triu(M) + np.triu(M, 1).T


def _full_solve_check(blsamd, p, lam):
    """reduce on the device, x_r by a dense solve of S, x_e by the device's backsubstitute!; against the oracle's
    reduction and backsubstitute!, and the whole x against the dense solve of (H + lam I) x = b."""
    gpu, orc = run_gpu(blsamd, p), oracle.schur_reduce_problem(p)
    compare(p, gpu, orc)
    assert gpu["h"] is not None
    x_r = np.linalg.solve(_sym(p, gpu["S"]), gpu["g"])
    x_e = _backsub(blsamd, p, gpu, x_r)
    x_e_o = np.zeros(p.structure.ne)
    oracle.backsubstitute_schur(orc["args"], x_e_o, x_r, p.C_data, orc["E_inv"], orc["z"])
    assert np.abs(x_e - x_e_o).max() <= TOL * max(1.0, np.abs(x_e_o).max())
    x_ref = np.linalg.solve(p.H + lam * np.eye(len(p.b)), p.b)
    assert np.linalg.norm(np.concatenate([x_r, x_e]) - x_ref) <= 1e-8 * np.linalg.norm(x_ref)


@pytest.mark.parametrize("dense", [False, True])
@pytest.mark.parametrize("de", [1, 3, 4, 5, 7, 8, 9, 300])
def test_reference_eliminated_size_sweep(blsamd, de, dense):
    """test/schur_problems.jl:28-35: each eliminated block size on its own (static sizes, the first dynamic size 9,
    and 300 above the old UInt8 limit), retained blocks [6, 6], couplings [[1, 2], [1]], lambda 0.2."""
    rng = np.random.default_rng(de)
    cpl = [[1, 2], [1]]
    H, b = schur_test_problem(rng, [6, 6], [de, de], cpl)
    p = problem_from_dense(blsamd, H, b, [6, 6], [de, de], cpl, 0.2, dense)
    _full_solve_check(blsamd, p, 0.2)


@pytest.mark.parametrize("dense", [False, True])
def test_reference_wide_retained_and_eliminated(blsamd, dense):
    """test/schur_problems.jl:36-41 (R9): a retained block of 260 and an eliminated block of 300."""
    rng = np.random.default_rng(256)
    cpl = [[1, 2], [2]]
    H, b = schur_test_problem(rng, [6, 260], [300, 4], cpl)
    p = problem_from_dense(blsamd, H, b, [6, 260], [300, 4], cpl, 0.2, dense)
    _full_solve_check(blsamd, p, 0.2)


def _mixed_sizes(blsamd, seed, dense, nland=60, rdims=(6, 9, 20), edims=(2, 3, 9, 12)):
    rng = np.random.default_rng(seed)
    r_dims = [int(d) for d in rng.choice(rdims, size=8)]
    e_dims = [int(d) for d in rng.choice(edims, size=nland)]
    cpl = [sorted(rng.choice(np.arange(1, 9), size=int(rng.integers(1, 5)), replace=False).tolist())
           for _ in range(nland)]
    H, b = schur_test_problem(rng, r_dims, e_dims, cpl, rr_pairs=[(1, 3), (2, 7)])
    scaling = 0.5 + rng.random(len(b))
    return problem_from_dense(blsamd, H, b, r_dims, e_dims, cpl, 0.05, dense, rr_pairs=[(1, 3), (2, 7)],
                              scaling=scaling)


@pytest.mark.parametrize("dense", [False, True])
def test_mixed_block_sizes_with_conditioning_scales(blsamd, dense):
    """Many landmarks of d_e 2, 3, 9, 12 over retained blocks of 6, 9 and 20, retained-retained blocks, per-variable
    damping scales and conditioning scales s_r, s_e: S, g_S, E_inv, z and backsubstitute! against the oracle."""
    torch = _torch()
    dev = blsamd.device
    p = _mixed_sizes(blsamd, 11, dense)
    st = p.structure
    rng = np.random.default_rng(5)
    s_r, s_e = 10.0 ** rng.uniform(-1, 1, st.nr), 10.0 ** rng.uniform(-1, 1, st.ne)
    orc = _oracle(p, np.float64, s_r=s_r, s_e=s_e)
    assert orc["ok"]
    h = dev.SchurHandle(st, p.ws_index, pattern=p.pattern, r_to_s=getattr(p, "r_to_s", None), dense=dense)
    ws = dev.SchurReductionWorkspace(h)
    T = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64)).cuda()
    E, C, dsr, dse = T(p.E_data), T(p.C_data), T(s_r), T(s_e)
    g = torch.full((st.nr,), float("nan"), dtype=torch.float64, device="cuda")
    if dense:
        S = torch.empty(st.nr * st.nr, dtype=torch.float64, device="cuda")
        dev.reset_schur_dense_(S, T(p.R.T.reshape(-1)), h, T(p.scaling_r), p.lam, s_r=dsr)
        ok = dev.reduce_schur_dense_(S, g, T(p.b_r), T(p.b_e), T(p.scaling_e), p.lam, h, E, C, ws, s_r=dsr, s_e=dse)
    else:
        S = torch.empty(p.pattern.nnz, dtype=torch.float64, device="cuda")
        dev.reset_schur_sparse_(S, T(p.R_data), h, T(p.scaling_r), p.lam, s_r=dsr)
        ok = dev.reduce_schur_sparse_(S, g, T(p.b_r), T(p.b_e), T(p.scaling_e), p.lam, h, E, C, ws, s_r=dsr, s_e=dse)
    torch.cuda.synchronize()
    gpu = dict(S=S.cpu().numpy(), g=g.cpu().numpy(), E_inv=ws.E_inv_data.cpu().numpy(), z=ws.z_data.cpu().numpy(),
               ok=ok)
    compare(p, gpu, orc)
    x_r = np.random.default_rng(6).standard_normal(st.nr)
    x_e = _backsub(blsamd, p, dict(h=h, C=C, ws=ws), x_r, s_r=s_r, s_e=s_e)
    x_e_o = np.zeros(st.ne)
    oracle.backsubstitute_schur(orc["args"], x_e_o, x_r, p.C_data, orc["E_inv"], orc["z"], s_r=s_r, s_e=s_e)
    assert np.abs(x_e - x_e_o).max() <= TOL * max(1.0, np.abs(x_e_o).max())


def test_mixed_block_sizes_repeatable(blsamd):
    """The wide kernels sum every entry in a fixed order: repeated reductions are bitwise identical."""
    p = _mixed_sizes(blsamd, 12, False)
    runs = run_gpu(blsamd, p, repeat=3)
    for r in runs[1:]:
        assert np.array_equal(r["S"], runs[0]["S"]) and np.array_equal(r["g"], runs[0]["g"])


@pytest.mark.parametrize("dense", [False, True])
def test_mixed_block_sizes_f32_and_mixed(blsamd, dense):
    """The Float32 variant (factorisation in Float32, as the reference's SchurBlockHessian{Float32}) and the mixed
    variant on blocks above the register path."""
    p = _mixed_sizes(blsamd, 13, dense, nland=30)
    _check_f32(p, run_variant(blsamd, p, "f32", lam=1e-2), lam=1e-2)
    _check_mixed(p, run_variant(blsamd, p, "mixed"))


@pytest.mark.parametrize("dense", [False, True])
def test_wide_block_failure_is_lowest_and_leaves_S(blsamd, dense):
    """A rank-deficient 9 x 9 and a 12 x 12 landmark block (lambda = 0): both fail, the lowest is reported as the
    oracle reports it (src/schur_reduction.jl:242-253), and S keeps what the reset wrote."""
    torch = _torch()
    p = _mixed_sizes(blsamd, 14, dense, nland=40, edims=(3, 9, 12))
    st = p.structure
    bad = [e for e in range(st.neblocks) if st.e_dims[e] > 8][1:3]
    for e in bad:
        de, eo = st.e_dims[e], st.eoffsets[e] - 1
        v = np.random.default_rng(e).standard_normal(de)
        p.E_data[eo:eo + de * de] = np.outer(v, v).reshape(-1)   # rank 1
    orc = oracle.schur_reduce_problem(p, lam=0.0)
    assert not orc["ok"] and orc["failed_block"] == bad[0] + 1
    gpu = run_gpu(blsamd, p, lam=0.0)
    assert not gpu["ok"] and gpu["failed_block"] == bad[0] + 1
    # S as the reset left it
    dev = blsamd.device
    T = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64)).cuda()
    h = gpu["h"]
    if dense:
        S0 = torch.empty(st.nr * st.nr, dtype=torch.float64, device="cuda")
        dev.reset_schur_dense_(S0, T(p.R.T.reshape(-1)), h, T(p.scaling_r), 0.0)
    else:
        S0 = torch.empty(p.pattern.nnz, dtype=torch.float64, device="cuda")
        dev.reset_schur_sparse_(S0, T(p.R_data), h, T(p.scaling_r), 0.0)
    torch.cuda.synchronize()
    assert np.array_equal(gpu["S"], S0.cpu().numpy())
    assert max_block_rel(_blocks(p, gpu["S"]), _blocks(p, orc["S"])) == 0.0
