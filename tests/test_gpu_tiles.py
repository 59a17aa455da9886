"""GPU parity of phase 2 on 16 x 16 tiles of S (the matrix-core kernel, bls_schur_set_phase2_mode 2) against the
oracle and against the work-list kernel (mode 1), through the C ABI.

The reference's phase 2 (src/schur_reduction.jl:315-380) is one computation, S[r1, r2] -= W_{c1}' C_{c2} and
g_S = b_r - sum C' z; the two device kernels sum it in different fixed orders. Tolerance (FP64): 1e-12 relative
Frobenius per block of S against the oracle and between the kernels, 1e-12 relative on g_S; each kernel bitwise
identical from run to run. The Float32 and mixed variants to the bounds of the precision tests.
"""
import numpy as np
import pytest

from helpers import max_block_rel, sparse_upper_blocks
from test_gpu_precision import _check_f32, _check_mixed, LAM32, run_variant
from test_gpu_schur import _scales, _torch, compare, oracle_run, run_gpu

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _blocks(p, S):
    return sparse_upper_blocks(p.structure, p.pattern, np.asarray(S, dtype=np.float64))


def _agree(p, a, b):
    """the two kernels' results: S per block and g within the Float64 bound"""
    assert max_block_rel(_blocks(p, a["S"]), _blocks(p, b["S"])) <= TOL
    assert np.abs(a["g"] - b["g"]).max() <= TOL * max(1.0, np.abs(b["g"]).max())


def _info(blsamd, p):
    _torch()   # the runtime initialised by torch before the library's first allocation
    h = blsamd.device.SchurHandle(p.structure, p.ws_index, pattern=p.pattern, r_to_s=p.r_to_s)
    return h, h.phase2_info()


def test_plan_selection(blsamd):
    """The tiles run by default where the layout fits them: the sequential-SLAM structure (contiguous windows,
    d_e 3, d_r 6) does; uniformly random camera sets and mixed landmark sizes do not, and forcing mode 2 there is
    an error."""
    gen = blsamd.generators
    p = gen.config2(scale=0.01)
    h, info = _info(blsamd, p)
    assert info[0] == 2 and info[1] > 0 and info[2] >= p.structure.neblocks and info[3] >= info[2]
    h.set_phase2_mode(1)
    assert h.phase2_info()[0] == 1
    h.set_phase2_mode(2)
    assert h.phase2_info()[0] == 2
    q = gen.config5_union(ncams=[40, 60], nlands=[3000, 4000])
    assert _info(blsamd, q)[1][0] == 2
    # random camera sets: windows are not contiguous
    rng = np.random.default_rng(3)
    ncam, nland = 20, 300
    k, ptr, cams = gen._tracks_uniform_random(rng, ncam, nland, 3, 6)
    r = gen._hessian_from_tracks(rng, ncam, np.full(ncam, 6, dtype=np.int64), nland, k, ptr, cams, 1e-4, False)
    h, info = _info(blsamd, r)
    assert info[0] == 1 and info[1] == 0
    with pytest.raises(blsamd.device.BLSError, match="contiguous"):
        h.set_phase2_mode(2)
    h.set_phase2_mode(1)


@pytest.mark.parametrize("de,dr", [(3, 6), (1, 6), (2, 9), (4, 12), (3, 16), (2, 7), (4, 8)])
def test_block_sizes_match_oracle_and_work_list(blsamd, de, dr):
    """Every landmark size the tiles take (1-4) on camera sizes that place block boundaries at every offset inside
    a tile (6, 7, 8, 9, 12, 16); windows up to 40 cameras reach many tile columns from one tile row."""
    p = blsamd.generators.window_problem(seed=10 + de * dr, ncam=48, nland=1200, de=de, dr=dr,
                                         kmin=3 if de == 4 else 2)
    a, a2 = run_gpu(blsamd, p, mode=2, repeat=2)
    assert a["h"].phase2_info()[0] == 2
    assert np.array_equal(a["S"], a2["S"]) and np.array_equal(a["g"], a2["g"])
    compare(p, a, oracle_run(p))
    _agree(p, a, run_gpu(blsamd, p, mode=1))


@pytest.mark.parametrize("ncam,de,dr,nland", [(2, 3, 6, 40), (3, 2, 16, 60), (1, 2, 6, 30)])
def test_tiny_layouts(blsamd, ncam, de, dr, nland):
    """Layouts of one or two tile rows (fewer scalar rows than a tile, a single camera): most XCDs get only empty
    units, every tile is clipped at nr."""
    p = blsamd.generators.window_problem(seed=ncam + de + dr, ncam=ncam, nland=nland, de=de, dr=dr,
                                         kmin=1 if ncam == 1 else 2, kmax=ncam)
    a = run_gpu(blsamd, p, mode=2)
    assert a["h"].phase2_info()[0] == 2
    compare(p, a, oracle_run(p))
    _agree(p, a, run_gpu(blsamd, p, mode=1))


def test_unobserved_cameras(blsamd):
    """Cameras seen by no landmark (a run in the middle, the last one): their tile rows hold no landmark, and g_S
    there is b_r alone; S blocks no landmark reaches are left as reset wrote them."""
    p = blsamd.generators.window_problem(seed=5, ncam=60, nland=1500, unobserved=list(range(30, 40)) + [59])
    a = run_gpu(blsamd, p, mode=2)
    assert a["h"].phase2_info()[0] == 2
    compare(p, a, oracle_run(p))
    rows = np.arange(30 * 6, 40 * 6)
    assert np.array_equal(a["g"][rows], p.b_r[rows])
    _agree(p, a, run_gpu(blsamd, p, mode=1))


def test_sequential_slam_shape(blsamd):
    """The sequential-SLAM configuration's generator (1000-camera shape cut to 100 cameras, 4000 landmarks) with the
    tiles, against the oracle; the tiles and the work list agree; each is bitwise reproducible."""
    p = blsamd.generators.config2(scale=0.02)
    a, a2 = run_gpu(blsamd, p, mode=2, repeat=2)
    assert np.array_equal(a["S"], a2["S"]) and np.array_equal(a["g"], a2["g"])
    compare(p, a, oracle_run(p))
    _agree(p, a, run_gpu(blsamd, p, mode=1))


def test_union_structure(blsamd):
    """The batched configuration's union structure (problems side by side): tiles never join two problems' blocks
    with a landmark, and S stays block diagonal over problems."""
    p = blsamd.generators.config5_union(ncams=[50, 80, 30, 64], nlands=[6000, 9000, 2500, 7000])
    a = run_gpu(blsamd, p, mode=2)
    compare(p, a, oracle_run(p))
    _agree(p, a, run_gpu(blsamd, p, mode=1))


def test_conditioning_scales(blsamd):
    """Conditioning scales s_r, s_e (src/conditioning.jl:99-141) through the tiles: C is read scaled."""
    import torch
    import oracle
    dev = blsamd.device
    p = blsamd.generators.config2(scale=0.01)
    st = p.structure
    s_r, s_e = _scales(p, 3)
    args = oracle.SchurArgs(st, p.ws_index, p.pattern)
    E_inv_o, z_o, W, g_o = np.zeros(len(p.E_data)), np.zeros(st.ne), np.zeros(len(p.C_data)), np.zeros(st.nr)
    S_o = np.zeros(p.pattern.nnz)
    oracle.reset_schur_sparse(S_o, p.R_data, p.r_to_s, p.pattern, p.scaling_r, p.lam, s_r=s_r)
    ok_o, _ = oracle.reduce_schur(True, args, S_o, g_o, p.b_r, p.b_e, p.scaling_e, p.lam, p.E_data, p.C_data,
                                  E_inv_o, z_o, W, s_r=s_r, s_e=s_e)
    T = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64)).cuda()
    outs = {}
    for mode in (2, 1):
        h = dev.SchurHandle(st, p.ws_index, pattern=p.pattern, r_to_s=p.r_to_s)
        h.set_phase2_mode(mode)
        ws = dev.SchurReductionWorkspace(h)
        S = torch.empty(p.pattern.nnz, dtype=torch.float64, device="cuda")
        g = torch.empty(st.nr, dtype=torch.float64, device="cuda")
        dev.reset_schur_sparse_(S, T(p.R_data), h, T(p.scaling_r), p.lam, s_r=T(s_r))
        ok = dev.reduce_schur_sparse_(S, g, T(p.b_r), T(p.b_e), T(p.scaling_e), p.lam, h, T(p.E_data), T(p.C_data),
                                      ws, s_r=T(s_r), s_e=T(s_e))
        torch.cuda.synchronize()
        outs[mode] = dict(S=S.cpu().numpy(), g=g.cpu().numpy(), E_inv=ws.E_inv_data.cpu().numpy(),
                          z=ws.z_data.cpu().numpy(), ok=ok)
    compare(p, outs[2], dict(S=S_o, g=g_o, E_inv=E_inv_o, z=z_o, ok=ok_o))
    _agree(p, outs[2], outs[1])


def test_singular_block_leaves_outputs(blsamd):
    """A failing landmark block (rank-1 E at lambda = 0, test/schur.jl:405-425): the lowest failing block is
    reported and the tiles write neither S nor g, as the work list does."""
    p = blsamd.generators.config2(scale=0.002)
    st = p.structure
    E = p.E_data.copy()
    for e in (9, 4, 15):
        v = np.arange(1.0, 4.0)
        eo = st.eoffsets[e - 1] - 1
        E[eo:eo + 9] = np.outer(v, v).T.reshape(-1)
    p.E_data = E
    a, b = run_gpu(blsamd, p, lam=0.0, mode=2), run_gpu(blsamd, p, lam=0.0, mode=1)
    assert a["ok"] is False and a["failed_block"] == 4 and b["failed_block"] == 4
    assert np.isnan(a["g"]).all()
    assert np.array_equal(a["S"], b["S"])


@pytest.mark.parametrize("de,dr", [(3, 6), (2, 9)])
def test_f32_and_mixed(blsamd, de, dr):
    """The Float32 variant (Float32 E_inv, C and S; Float64 sums) and the mixed variant (Float32 C, Float64 E_inv and
    S) through the tiles, to the precision tests' bounds."""
    p = blsamd.generators.window_problem(seed=3 * de + dr, ncam=40, nland=1000, de=de, dr=dr)
    d32 = run_variant(blsamd, p, "f32", lam=LAM32, mode=2)
    assert d32["h"].phase2_info()[0] == 2
    _check_f32(p, d32)
    _check_mixed(p, run_variant(blsamd, p, "mixed", mode=2))
