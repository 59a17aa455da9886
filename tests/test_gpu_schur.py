"""GPU parity of the Schur path: the HIP kernels through the C ABI against the CPU oracle on the same inputs.

Tolerance (floating point, FP64): 1e-12 relative Frobenius per block of S (upper blocks), 1e-12 relative on
g_S / E_inv / z / x_e. Integer outputs (failed_block, status) exact; repeated runs bitwise identical.
"""
import numpy as np
import pytest

import oracle
from helpers import (dense_magnitude_blocks, dense_upper_blocks, max_block_rel, max_block_rel_to, problem_from_dense,
                     reference_schur_problem, schur_test_problem, sparse_upper_blocks, standard_problem_dims)

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _torch():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return torch


def run_gpu(blsamd, p, lam=None, repeat=1, mode=None):
    """reset + reduce on the device; returns host copies of S, g, E_inv, z, status, failed_block. `mode`: the
    phase-2 kernel (bls_schur_set_phase2_mode), default automatic."""
    torch = _torch()
    dev = blsamd.device
    st = p.structure
    lam = p.lam if lam is None else lam
    h = dev.SchurHandle(st, p.ws_index, pattern=p.pattern, r_to_s=getattr(p, "r_to_s", None), dense=p.dense)
    if mode is not None:
        h.set_phase2_mode(mode)
    ws = dev.SchurReductionWorkspace(h)
    T = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64)).cuda()
    E, C, b_r, b_e = T(p.E_data), T(p.C_data), T(p.b_r), T(p.b_e)
    sr, se = T(p.scaling_r), T(p.scaling_e)
    outs = []
    for _ in range(repeat):
        g = torch.full((st.nr,), float("nan"), dtype=torch.float64, device="cuda")
        if p.dense:
            S = torch.empty(st.nr * st.nr, dtype=torch.float64, device="cuda")
            dev.reset_schur_dense_(S, T(p.R.T.reshape(-1)), h, sr, lam)
            ok = dev.reduce_schur_dense_(S, g, b_r, b_e, se, lam, h, E, C, ws)
        else:
            S = torch.empty(p.pattern.nnz, dtype=torch.float64, device="cuda")
            dev.reset_schur_sparse_(S, T(p.R_data), h, sr, lam)
            ok = dev.reduce_schur_sparse_(S, g, b_r, b_e, se, lam, h, E, C, ws)
        torch.cuda.synchronize()
        outs.append(dict(S=S.cpu().numpy(), g=g.cpu().numpy(), E_inv=ws.E_inv_data.cpu().numpy(),
                         z=ws.z_data.cpu().numpy(), ok=ok, failed_block=ws.failed_block, h=h, ws=ws, C=C))
    return outs if repeat > 1 else outs[0]


def oracle_run(p, lam=None):
    return oracle.schur_reduce_problem(p, lam=lam)


def compare(p, gpu, orc):
    st = p.structure
    assert gpu["ok"] == orc["ok"]
    if p.dense:
        A, B = dense_upper_blocks(st, gpu["S"]), dense_upper_blocks(st, orc["S"])
    else:
        A, B = sparse_upper_blocks(st, p.pattern, gpu["S"]), sparse_upper_blocks(st, p.pattern, orc["S"])
    assert max_block_rel(A, B) <= TOL
    for k in ("g", "E_inv", "z"):
        ref = orc[k]
        assert np.abs(gpu[k] - ref).max() <= TOL * max(1.0, np.abs(ref).max()), k


def test_config1_dense(blsamd):
    p = blsamd.generators.config1()
    compare(p, run_gpu(blsamd, p), oracle_run(p))


def test_config2_small_sparse(blsamd):
    p = blsamd.generators.config2(scale=0.01)
    compare(p, run_gpu(blsamd, p), oracle_run(p))


def test_config2_small_dense(blsamd):
    p = blsamd.generators.config2(scale=0.002, dense=True)
    compare(p, run_gpu(blsamd, p), oracle_run(p))


def test_config3_shape_dense(blsamd):
    """The dense configuration's shape at a reduced landmark count: 256 cameras, k ~ U{2..64} random cameras per
    landmark, Jacobian blocks scaled per camera by 10^U(-2, 2). Blocks joining a 1e2-scaled camera with a
    1e-2-scaled one cancel their summands to ~1e-9 of their size: two exact-arithmetic-equivalent orders (the
    oracle and a numpy LU reduction) differ there by 3e-9 relative per block, so S is compared relative to the
    magnitude it is summed from, |R| + sum |C_1|'|E^-1||C_2| (1e-12); g, E_inv and z as everywhere."""
    p = blsamd.generators.config3(scale=0.003)
    assert p.dense and p.structure.nrblocks == 256
    gpu, orc = run_gpu(blsamd, p), oracle_run(p)
    st = p.structure
    assert gpu["ok"] and orc["ok"]
    mag = dense_magnitude_blocks(p)
    assert max_block_rel_to(dense_upper_blocks(st, gpu["S"]), dense_upper_blocks(st, orc["S"]), mag) <= TOL
    for k in ("g", "E_inv", "z"):
        assert np.abs(gpu[k] - orc[k]).max() <= TOL * max(1.0, np.abs(orc[k]).max()), k


@pytest.mark.parametrize("dense", [True, False])
def test_reference_structure_mixed_dims(blsamd, dense):
    """The hand-built structure of test/schur.jl:12-66 (retained dims 6/3/6/9, eliminated 3 and 1)."""
    rng = np.random.default_rng(11)
    r_dims, e_dims, cpl = reference_schur_problem()
    H, b = schur_test_problem(rng, r_dims, e_dims, cpl, rr_pairs=[(1, 2), (2, 3), (3, 4)])
    p = problem_from_dense(blsamd, H, b, r_dims, e_dims, cpl, 0.3, dense, rr_pairs=[(1, 2), (2, 3), (3, 4)])
    compare(p, run_gpu(blsamd, p), oracle_run(p))


@pytest.mark.parametrize("de", [1, 2, 3, 4, 5, 7, 8])
def test_eliminated_dims(blsamd, de):
    """test/schur_problems.jl:13-41's sweep over eliminated block sizes (device path: d_e <= 8)."""
    rng = np.random.default_rng(100 + de)
    nrb, neb = 5, 12
    cpl = [sorted(rng.choice(np.arange(1, nrb + 1), size=rng.integers(1, 4), replace=False).tolist()) for _ in range(neb)]
    H, b = schur_test_problem(rng, [6] * nrb, [de] * neb, cpl)
    for dense in (True, False):
        p = problem_from_dense(blsamd, H, b, [6] * nrb, [de] * neb, cpl, 0.0, dense)
        compare(p, run_gpu(blsamd, p), oracle_run(p))


def test_standard_problem_mixed_eliminated(blsamd):
    rng = np.random.default_rng(5)
    r_dims, e_dims = standard_problem_dims()
    cpl = [sorted(rng.choice(np.arange(1, 5), size=rng.integers(1, 5), replace=False).tolist()) for _ in e_dims]
    H, b = schur_test_problem(rng, r_dims, e_dims, cpl)
    for dense in (True, False):
        p = problem_from_dense(blsamd, H, b, r_dims, e_dims, cpl, 1e-3, dense)
        compare(p, run_gpu(blsamd, p), oracle_run(p))


def test_bitwise_reproducible(blsamd):
    p = blsamd.generators.config2(scale=0.02)
    a, b = run_gpu(blsamd, p, repeat=2)
    assert np.array_equal(a["S"], b["S"]) and np.array_equal(a["g"], b["g"])


# Columns with more landmarks than one workgroup takes (384) are split into parts summed in part order; these
# workloads put 500-1000 landmarks on every camera so that every diagonal item is split.
SPLIT = 384


def _max_column_landmarks(p):
    return int(np.diff(p.ws_index.retained_ptr).max())


@pytest.mark.parametrize("dense", [False, True])
def test_split_columns(blsamd, dense):
    p = blsamd.generators.config2(scale=0.01 if not dense else 0.006, ncam=10 if not dense else 6, dense=dense)
    assert _max_column_landmarks(p) > 2 * SPLIT
    compare(p, run_gpu(blsamd, p), oracle_run(p))
    a, b = run_gpu(blsamd, p, repeat=2)
    assert np.array_equal(a["S"], b["S"]) and np.array_equal(a["g"], b["g"])


def test_split_columns_mixed_eliminated_dims(blsamd):
    """Landmark dimensions 1-4 mixed (the packed-record path) with 500+ landmarks per camera."""
    rng = np.random.default_rng(17)
    r_dims = [6, 9, 6, 9]
    e_dims = rng.integers(1, 5, size=900).tolist()
    cpl = [sorted(rng.choice(np.arange(1, 5), size=rng.integers(2, 5), replace=False).tolist()) for _ in e_dims]
    H, b = schur_test_problem(rng, r_dims, e_dims, cpl)
    p = problem_from_dense(blsamd, H, b, r_dims, e_dims, cpl, 1e-3, False)
    assert _max_column_landmarks(p) > SPLIT
    compare(p, run_gpu(blsamd, p), oracle_run(p))


def test_singular_block_reports_lowest(blsamd):
    """Rank-1 E at lambda = 0 fails with the lowest failing block (test/schur.jl:405-425)."""
    p = blsamd.generators.config2(scale=0.002)
    st = p.structure
    E = p.E_data.copy()
    bad = [7, 3, 12]   # 1-based eliminated blocks made rank 1
    for e in bad:
        v = np.arange(1.0, 4.0)
        eo = st.eoffsets[e - 1] - 1
        E[eo:eo + 9] = np.outer(v, v).T.reshape(-1)
    p.E_data = E
    gpu = run_gpu(blsamd, p, lam=0.0)
    orc = oracle_run(p, lam=0.0)
    assert orc["ok"] is False and orc["failed_block"] == 3
    assert gpu["ok"] is False and gpu["failed_block"] == 3
    assert np.isnan(gpu["g"]).all()   # g_S is not written when phase 1 fails


def test_backsubstitute(blsamd):
    torch = _torch()
    p = blsamd.generators.config2(scale=0.005)
    st = p.structure
    gpu = run_gpu(blsamd, p)
    orc = oracle_run(p)
    x_r = np.random.default_rng(0).standard_normal(st.nr)
    xe_o = np.zeros(st.ne)
    oracle.backsubstitute_schur(orc["args"], xe_o, x_r, p.C_data, orc["E_inv"], orc["z"])
    xe_g = torch.empty(st.ne, dtype=torch.float64, device="cuda")
    blsamd.device.backsubstitute_(xe_g, torch.as_tensor(x_r).cuda(), gpu["h"], gpu["C"], gpu["ws"])
    xe_g = xe_g.cpu().numpy()
    assert np.abs(xe_g - xe_o).max() <= TOL * np.abs(xe_o).max()


def test_full_solve_matches_dense_cholesky(blsamd):
    """reset -> reduce -> dense reduced solve -> backsubstitute equals the dense solve of the damped H
    (reference_solve, test/fixtures/SchurFixtures.jl:114-117)."""
    torch = _torch()
    rng = np.random.default_rng(3)
    r_dims, e_dims, cpl = reference_schur_problem()
    H, b = schur_test_problem(rng, r_dims, e_dims, cpl)
    lam = 0.3
    p = problem_from_dense(blsamd, H, b, r_dims, e_dims, cpl, lam, True)
    gpu = run_gpu(blsamd, p)
    st = p.structure
    S = gpu["S"].reshape(st.nr, st.nr).T
    S = np.triu(S) + np.triu(S, 1).T
    x_r = np.linalg.solve(S, gpu["g"])
    xe = torch.empty(st.ne, dtype=torch.float64, device="cuda")
    blsamd.device.backsubstitute_(xe, torch.as_tensor(x_r).cuda(), gpu["h"], gpu["C"], gpu["ws"])
    x = np.concatenate([x_r, xe.cpu().numpy()])
    x_ref = np.linalg.solve(H + lam * np.eye(len(b)), b)
    assert np.linalg.norm(x - x_ref) <= 1e-9 * np.linalg.norm(x_ref)


def test_copyto_symmetric_csc(blsamd):
    """copyto!(csc, S, scatter_map) of the reduced system into its full symmetric CSC equals the oracle's S
    symmetrised, bit for bit (a gather)."""
    torch = _torch()
    import scipy.sparse as sps
    p = blsamd.generators.config2(scale=0.005)
    st = p.structure
    gpu = run_gpu(blsamd, p)
    colptr, rowval, smap = blsamd.structure.reduced_symmetric_scatter(st, p.pattern)
    dmap = blsamd.device.scatter_map_to_device(smap)
    dest = torch.empty(len(smap), dtype=torch.float64, device="cuda")
    blsamd.device.copyto_(dest, torch.as_tensor(gpu["S"]).cuda(), dmap)
    vals = dest.cpu().numpy()
    assert np.array_equal(vals, gpu["S"][smap - 1])
    A = sps.csc_matrix((vals, rowval - 1, colptr - 1), shape=(st.nr, st.nr)).toarray()
    assert np.array_equal(A, A.T)
    orc = oracle_run(p)
    B = sps.csc_matrix((orc["S"][smap - 1], rowval - 1, colptr - 1), shape=(st.nr, st.nr)).toarray()
    assert np.abs(A - B).max() <= TOL * np.abs(B).max()


def test_null_required_arrays_rejected(blsamd):
    """A NULL scaling vector (or any required array) is an argument error raised on the host, before any launch."""
    torch = _torch()
    dev = blsamd.device
    p = blsamd.generators.config2(scale=0.001)
    st = p.structure
    h = dev.SchurHandle(st, p.ws_index, pattern=p.pattern, r_to_s=p.r_to_s)
    ws = dev.SchurReductionWorkspace(h)
    T = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64)).cuda()
    S = torch.zeros(p.pattern.nnz, dtype=torch.float64, device="cuda")
    g = torch.zeros(st.nr, dtype=torch.float64, device="cuda")
    with pytest.raises(dev.BLSError):
        dev.reset_schur_sparse_(S, T(p.R_data), h, None, p.lam, check=False)
    with pytest.raises(dev.BLSError):
        dev.reduce_schur_sparse_(S, g, T(p.b_r), T(p.b_e), None, p.lam, h, T(p.E_data), T(p.C_data), ws, check=False)
    with pytest.raises(ValueError):
        dev.reduce_schur_sparse_(S, g, T(p.b_r), T(p.b_e), T(p.scaling_e), -1.0, h, T(p.E_data), T(p.C_data), ws)


def _scales(p, seed):
    rng = np.random.default_rng(seed)
    st = p.structure
    return 10.0 ** rng.uniform(-1, 1, st.nr), 10.0 ** rng.uniform(-1, 1, st.ne)


@pytest.mark.parametrize("dense", [False, True])
def test_conditioning_scales(blsamd, dense):
    """Conditioning scales s_r, s_e (src/conditioning.jl:99-141): every stored value is read as
    scaled(a, s_i s_j); reset scales R and the damping by s_r^2; phase 1 scales E and b_e; phase 2 reads
    scaled C; backsubstitute! unscales nothing (x is in the scaled variables)."""
    torch = _torch()
    dev = blsamd.device
    p = blsamd.generators.config2(scale=0.01 if not dense else 0.003, dense=dense)
    st = p.structure
    s_r, s_e = _scales(p, 3)
    # oracle
    args = oracle.SchurArgs(st, p.ws_index, p.pattern)
    E_inv_o = np.zeros(len(p.E_data)); z_o = np.zeros(st.ne); W = np.zeros(len(p.C_data)); g_o = np.zeros(st.nr)
    if dense:
        S_o = np.zeros(st.nr * st.nr)
        oracle.reset_schur_dense(S_o, p.R.T.reshape(-1).copy(), p.scaling_r, p.lam, s_r=s_r)
    else:
        S_o = np.zeros(p.pattern.nnz)
        oracle.reset_schur_sparse(S_o, p.R_data, p.r_to_s, p.pattern, p.scaling_r, p.lam, s_r=s_r)
    ok_o, fb_o = oracle.reduce_schur(not dense, args, S_o, g_o, p.b_r, p.b_e, p.scaling_e, p.lam, p.E_data,
                                     p.C_data, E_inv_o, z_o, W, s_r=s_r, s_e=s_e)
    assert ok_o
    # device
    h = dev.SchurHandle(st, p.ws_index, pattern=p.pattern, r_to_s=getattr(p, "r_to_s", None), dense=dense)
    ws = dev.SchurReductionWorkspace(h)
    T = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64)).cuda()
    E, C, b_r, b_e, sr, se = T(p.E_data), T(p.C_data), T(p.b_r), T(p.b_e), T(p.scaling_r), T(p.scaling_e)
    dsr, dse = T(s_r), T(s_e)
    g = torch.empty(st.nr, dtype=torch.float64, device="cuda")
    if dense:
        S = torch.empty(st.nr * st.nr, dtype=torch.float64, device="cuda")
        dev.reset_schur_dense_(S, T(p.R.T.reshape(-1)), h, sr, p.lam, s_r=dsr)
        ok = dev.reduce_schur_dense_(S, g, b_r, b_e, se, p.lam, h, E, C, ws, s_r=dsr, s_e=dse)
    else:
        S = torch.empty(p.pattern.nnz, dtype=torch.float64, device="cuda")
        dev.reset_schur_sparse_(S, T(p.R_data), h, sr, p.lam, s_r=dsr)
        ok = dev.reduce_schur_sparse_(S, g, b_r, b_e, se, p.lam, h, E, C, ws, s_r=dsr, s_e=dse)
    torch.cuda.synchronize()
    gpu = dict(S=S.cpu().numpy(), g=g.cpu().numpy(), E_inv=ws.E_inv_data.cpu().numpy(), z=ws.z_data.cpu().numpy(), ok=ok)
    orc = dict(S=S_o, g=g_o, E_inv=E_inv_o, z=z_o, ok=ok_o)
    compare(p, gpu, orc)
    # the unscaled reduction differs: the scales are really applied
    base = oracle_run(p)
    assert not np.allclose(base["g"], g_o)
    # backsubstitute with the same scales
    x_r = np.random.default_rng(1).standard_normal(st.nr)
    xe_o = np.zeros(st.ne)
    oracle.backsubstitute_schur(args, xe_o, x_r, p.C_data, E_inv_o, z_o, s_r=s_r, s_e=s_e)
    xe = torch.empty(st.ne, dtype=torch.float64, device="cuda")
    dev.backsubstitute_(xe, T(x_r), h, C, ws, s_r=dsr, s_e=dse)
    torch.cuda.synchronize()
    assert np.abs(xe.cpu().numpy() - xe_o).max() <= TOL * max(1.0, np.abs(xe_o).max())


@pytest.mark.parametrize("variant", ["f64", "mixed"])
def test_graph_captured_step_replays(blsamd, variant):
    """reset_schur_sparse! + reduce_schur_sparse! (asynchronous status) + copyto! captured once in a HIP graph and
    replayed three times, with eager work on the same buffers between the replays: every replay equals the eager
    step bit for bit (the path records no memset node: flags and S are cleared by kernels)."""
    torch = _torch()
    dev = blsamd.device
    p = blsamd.generators.config2(scale=0.02)
    st = p.structure
    h = dev.SchurHandle(st, p.ws_index, pattern=p.pattern, r_to_s=p.r_to_s)
    ws = dev.SchurReductionWorkspace(h, variant)
    bt = torch.float32 if variant == "mixed" else torch.float64
    T = lambda a, d=torch.float64: torch.as_tensor(np.ascontiguousarray(a)).to(d).cuda()
    E, C, Rd = T(p.E_data, bt), T(p.C_data, bt), T(p.R_data, bt)
    b_r, b_e, sr, se = T(p.b_r), T(p.b_e), T(p.scaling_r), T(p.scaling_e)
    colptr, rowval, smap = blsamd.structure.reduced_symmetric_scatter(st, p.pattern)
    dmap = dev.scatter_map_to_device(smap)
    S = torch.empty(p.pattern.nnz, dtype=torch.float64, device="cuda")
    g = torch.empty(st.nr, dtype=torch.float64, device="cuda")
    csc = torch.empty(len(smap), dtype=torch.float64, device="cuda")

    def step():
        dev.reset_schur_sparse_(S, Rd, h, sr, p.lam, check=False)
        dev.reduce_schur_sparse_(S, g, b_r, b_e, se, p.lam, h, E, C, ws, sync=False, check=False)
        dev.copyto_(csc, S, dmap)

    step()
    torch.cuda.synchronize()
    ref = [x.clone() for x in (S, g, csc)]
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        step()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        step()
    for rep in range(3):
        S.fill_(float("nan"))
        g.fill_(float("nan"))
        csc.fill_(0.0)
        ws.d_failed.fill_(-1)
        graph.replay()
        torch.cuda.synchronize()
        assert int(ws.d_failed.item()) == 0, rep
        for a, b in zip((S, g, csc), ref):
            assert torch.equal(a, b), rep
        step()   # eager work between replays
        torch.cuda.synchronize()


def test_autotune_shapes_agree(blsamd):
    """The phase-2 work-list shape changes the grouping of the sums, not their value: every tuned shape reproduces the
    oracle within the Float64 bound, each shape is deterministic, and the autotuner keeps one of its candidates."""
    torch = _torch()
    dev = blsamd.device
    p = blsamd.generators.config2(scale=0.01, ncam=10)   # > 384 landmarks per camera: split items
    st = p.structure
    orc = oracle_run(p)
    h = dev.SchurHandle(st, p.ws_index, pattern=p.pattern, r_to_s=p.r_to_s)
    ws = dev.SchurReductionWorkspace(h)
    T = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64)).cuda()
    E, C, b_r, b_e, sr, se, Rd = (T(x) for x in (p.E_data, p.C_data, p.b_r, p.b_e, p.scaling_r, p.scaling_e, p.R_data))
    S = torch.empty(p.pattern.nnz, dtype=torch.float64, device="cuda")
    g = torch.empty(st.nr, dtype=torch.float64, device="cuda")
    cands = [(384, 96), (128, 48), (1024, 256), (64, 16)]
    h.set_phase2_mode(1)   # the work list (the shapes are its shapes)
    for part, lanes in cands:
        h.set_tuning(part, lanes)
        outs = []
        for _ in range(2):
            dev.reset_schur_sparse_(S, Rd, h, sr, p.lam)
            assert dev.reduce_schur_sparse_(S, g, b_r, b_e, se, p.lam, h, E, C, ws)
            torch.cuda.synchronize()
            outs.append((S.cpu().numpy(), g.cpu().numpy()))
        assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[1][1])
        compare(p, dict(S=outs[0][0], g=outs[0][1], E_inv=ws.E_inv_data.cpu().numpy(), z=ws.z_data.cpu().numpy(),
                        ok=True), orc)
    dev.reset_schur_sparse_(S, Rd, h, sr, p.lam)
    part, lanes, ms = h.autotune(cands, S, g, b_r, b_e, se, p.lam, E, C, ws)
    assert (part, lanes) in cands and ms > 0
    assert h.phase2_info()[0] == 1   # mode 1 set by the caller: the tiles are not timed
    # in automatic mode the tiles are timed as well; the kernel kept reproduces the oracle
    h.set_phase2_mode(0)
    dev.reset_schur_sparse_(S, Rd, h, sr, p.lam)
    part, lanes, ms = h.autotune(cands, S, g, b_r, b_e, se, p.lam, E, C, ws)
    assert (part, lanes) in cands and ms > 0 and h.phase2_info()[0] in (1, 2)
    dev.reset_schur_sparse_(S, Rd, h, sr, p.lam)
    assert dev.reduce_schur_sparse_(S, g, b_r, b_e, se, p.lam, h, E, C, ws)
    torch.cuda.synchronize()
    compare(p, dict(S=S.cpu().numpy(), g=g.cpu().numpy(), E_inv=ws.E_inv_data.cpu().numpy(),
                    z=ws.z_data.cpu().numpy(), ok=True), orc)


@pytest.mark.parametrize("shape", [None, (1536, 96), (2304, 64)])
def test_config5_union_matches_oracle(blsamd, shape):
    """The batched configuration's union structure (problems side by side, one reduction for all of them) on 8
    problems of uneven size, at the default work-list shape and at the shapes the bench's tuner picks for the
    union and for the sequential-SLAM workload. The dense columns of the small problems (about 500-900 landmarks
    per camera) are split into parts and merged, so the split path is covered too. S within 1e-12 relative
    Frobenius per block of the oracle's; g, E_inv and z as everywhere."""
    p = blsamd.generators.config5_union(ncams=[50, 80, 120, 60, 200, 90, 150, 70],
                                        nlands=[24000, 8000, 20000, 6000, 15000, 9000, 11000, 7000])
    st = p.structure
    per_col = np.bincount(np.asarray(st.entry_retained) - 1, minlength=st.nrblocks)
    assert per_col.max() > 2304 and p.nprob == 8   # columns longer than any part: every shape splits some
    torch = _torch()
    dev = blsamd.device
    h = dev.SchurHandle(st, p.ws_index, pattern=p.pattern, r_to_s=p.r_to_s)
    if shape is not None:
        h.set_tuning(*shape)
    ws = dev.SchurReductionWorkspace(h)
    T = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64)).cuda()
    S = torch.empty(p.pattern.nnz, dtype=torch.float64, device="cuda")
    g = torch.empty(st.nr, dtype=torch.float64, device="cuda")
    dev.reset_schur_sparse_(S, T(p.R_data), h, T(p.scaling_r), p.lam)
    ok = dev.reduce_schur_sparse_(S, g, T(p.b_r), T(p.b_e), T(p.scaling_e), p.lam, h, T(p.E_data), T(p.C_data), ws)
    gpu = dict(S=S.cpu().numpy(), g=g.cpu().numpy(), E_inv=ws.E_inv_data.cpu().numpy(), z=ws.z_data.cpu().numpy(),
               ok=ok)
    compare(p, gpu, oracle.schur_reduce_problem(p, nthreads=4))
