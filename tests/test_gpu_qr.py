"""GPU parity of reduce_qr!: the HIP table kernel through the C ABI against the CPU oracle on the same inputs.

Tolerance (FP64): per table, ||F_gpu - F_oracle||_F <= 1e-12 ||F_oracle||_F and the same for the factored
residual; failing block exact; raw tables untouched; repeated runs bitwise identical."""
import numpy as np
import pytest

import oracle
from test_oracle_qr import _qr, _small_jacobian

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _jac_with_free_rows(blsamd, seed=5, nland=60, ncam=8, tall_pairs=(), wide=0):
    """A description with observations (camera, landmark), camera-camera rows and camera priors (free groups).
    tall_pairs: (a, b, n) adds n relative-pose row blocks (3 rows each) between cameras a and b, one free table of
    3 n rows (the cameras then have 9 parameters each). wide: two row blocks of 3 rows over cameras 0 .. wide - 1,
    one free table with that many retained blocks."""
    rng = np.random.default_rng(seed)
    S = blsamd.structure
    cam_dims = np.where(rng.random(ncam) < 0.5, 6, 9)
    if tall_pairs:
        cam_dims[:] = 9
    blocksizes = np.concatenate([cam_dims, np.full(nland, 3)])
    eliminated = np.concatenate([np.zeros(ncam, bool), np.ones(nland, bool)])
    rowblocks, rowsz = [], []
    for l in range(nland):
        for c in rng.choice(ncam, size=int(rng.integers(2, 5)), replace=False):
            rowblocks.append([int(c) + 1, ncam + l + 1]); rowsz.append(2)
    for c in range(ncam):
        rowblocks.append([c + 1]); rowsz.append(int(cam_dims[c]))          # prior on camera c
    for _ in range(5):
        a, b = rng.choice(ncam, size=2, replace=False)
        rowblocks.append([int(a) + 1, int(b) + 1]); rowsz.append(3)       # relative-pose rows
    for a, b, n in tall_pairs:
        for _ in range(n):
            rowblocks.append([a + 1, b + 1]); rowsz.append(3)
    for _ in range(2 if wide else 0):
        rowblocks.append(list(range(1, wide + 1))); rowsz.append(3)
    jst = S.SchurJacobianStructure.from_description(blocksizes, rowblocks, rowsz, eliminated)
    data = rng.standard_normal(jst.raw_length())
    # damping rows of eliminated tables are zero in the raw storage
    for t in range(jst.schur.neblocks):
        M, C, top = jst.table_nrows[t], jst.table_ncols[t] - 1, jst.table_top[t]
        off = jst.raw_offsets[t] - 1
        T = data[off:off + M * C].reshape(C, M)
        T[:, M - top:] = 0.0
    r = rng.standard_normal(jst.residual_length())
    for t in range(jst.schur.neblocks):
        M, top = jst.table_nrows[t], jst.table_top[t]
        r[jst.residual_offsets[t] - 1 + M - top: jst.residual_offsets[t] - 1 + M] = 0.0

    class P:
        pass
    p = P()
    p.jst, p.structure, p.data, p.r, p.scaling_e = jst, jst.schur, data, r, np.ones(jst.schur.ne)
    return p


def run_gpu(blsamd, p, lam, repeat=1):
    import torch
    dev = blsamd.device
    h = dev.QRHandle(p.jst)
    ws = dev.QRReductionWorkspace(h)
    T = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64)).cuda()
    raw, r, sc = T(p.data), T(p.r), T(p.scaling_e)
    outs = []
    for _ in range(repeat):
        ok = dev.reduce_qr_(ws, h, raw, r, sc, lam)
        torch.cuda.synchronize()
        outs.append((ok, ws.failed_block, ws.factored.cpu().numpy(), ws.fres.cpu().numpy()))
    assert np.array_equal(raw.cpu().numpy(), p.data, equal_nan=True)   # raw tables are read only
    return outs if repeat > 1 else outs[0]


def compare_tables(p, Fg, Frg, Fo, Fro):
    jst = p.jst
    for t in range(jst.ntables):
        a, b = jst.raw_offsets[t] - 1, jst.raw_offsets[t + 1] - 1
        nrm = np.linalg.norm(Fo[a:b])
        assert np.linalg.norm(Fg[a:b] - Fo[a:b]) <= TOL * max(nrm, 1e-300), t
        a, b = jst.residual_offsets[t] - 1, jst.residual_offsets[t + 1] - 1
        nrm = np.linalg.norm(Fro[a:b])
        assert np.linalg.norm(Frg[a:b] - Fro[a:b]) <= TOL * max(nrm, 1e-300), t


@pytest.mark.parametrize("seed", [0, 1])
def test_qr_matches_oracle(blsamd, seed):
    p = _small_jacobian(blsamd, seed, nland=80, ncam=10, kmax=6)
    ok, fb, Fg, Frg = run_gpu(blsamd, p, 0.1)
    oko, fbo, Fo, Fro, _, _ = _qr(p, 0.1)
    assert ok == oko and fb == fbo == 0
    compare_tables(p, Fg, Frg, Fo, Fro)


def test_qr_free_tables_and_mixed_dims(blsamd):
    p = _jac_with_free_rows(blsamd)
    assert p.jst.ntables > p.structure.neblocks
    ok, fb, Fg, Frg = run_gpu(blsamd, p, 0.3)
    oko, fbo, Fo, Fro, _, _ = _qr(p, 0.3)
    assert ok and oko
    compare_tables(p, Fg, Frg, Fo, Fro)


def test_qr_config4_tall_tables(blsamd):
    """Config-4 shaped tables (k up to 30, rows up to 63, cameras of dim 6 and 9)."""
    p = blsamd.generators.config4(scale=0.0005, ncam=40, kmax=30)
    assert p.jst.table_nrows[:p.structure.neblocks].max() > 32
    ok, fb, Fg, Frg = run_gpu(blsamd, p, 1e-4)
    oko, fbo, Fo, Fro, _, _ = _qr(p, 1e-4)
    assert ok and oko
    compare_tables(p, Fg, Frg, Fo, Fro)


def test_qr_failures_lowest_block(blsamd):
    p = _small_jacobian(blsamd, 3)
    jst, st = p.jst, p.structure
    for e in (7, 11):
        M, C, de = jst.table_nrows[e - 1], jst.table_ncols[e - 1] - 1, st.e_dims[e - 1]
        off = jst.raw_offsets[e - 1] - 1
        p.data[off + (C - de) * M: off + C * M] = 0.0
    p.data[jst.raw_offsets[20] - 1] = np.nan
    ok, fb, _, _ = run_gpu(blsamd, p, 0.0)
    assert not ok and fb == 7
    ok, fb, _, _ = run_gpu(blsamd, p, 0.5)
    assert not ok and fb == 21


def test_qr_bitwise_reproducible(blsamd):
    p = _small_jacobian(blsamd, 6, nland=200, ncam=12, kmax=8)
    (a, _, Fa, Fra), (b, _, Fb, Frb) = run_gpu(blsamd, p, 0.1, repeat=2)
    assert np.array_equal(Fa, Fb) and np.array_equal(Fra, Frb)


@pytest.mark.parametrize("free", ["none", "free", "tall"])
def test_reduced_products_match_oracle(blsamd, free):
    """reduced_mul!, reduced_tmul!, reduced_residual!, reduced_diagonal_blocks! (src/qr_reduced.jl:161-243,
    :484-512) on the device, from the device-factored tables, against the oracle on the same factored tables.
    "tall" adds free tables of 210, 300 and 1200 rows (more than one 64-row chunk of the diagonal-block pass)."""
    import torch
    dev = blsamd.device
    if free == "tall":
        p = _jac_with_free_rows(blsamd, seed=9, tall_pairs=((0, 1, 70), (2, 3, 100), (4, 5, 400)))
    elif free == "free":
        p = _jac_with_free_rows(blsamd)
    else:
        p = _small_jacobian(blsamd, 7, nland=90, ncam=9, kmax=6)
    ok, fb, Fg, Frg = run_gpu(blsamd, p, 0.2)
    assert ok
    jst, st = p.jst, p.structure
    sargs = oracle.SchurArgs(st, blsamd.structure.SchurReductionWorkspaceIndex(st, 1))
    jargs = oracle.JacobianArgs(jst)
    h = dev.QRHandle(jst)
    T = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64)).cuda()
    F, Fr = T(Fg), T(Frg)
    nred = jst.reduced_nrows()
    rng = np.random.default_rng(2)
    v = rng.standard_normal(st.nr)
    w = rng.standard_normal(nred)
    y = torch.empty(nred, dtype=torch.float64, device="cuda")
    dev.reduced_mul_(y, h, F, T(v))
    yo = np.zeros(nred)
    oracle.reduced_mul(sargs, jargs, yo, Fg, v)
    g = torch.empty(st.nr, dtype=torch.float64, device="cuda")
    dev.reduced_tmul_(g, h, F, T(w))
    go = np.zeros(st.nr)
    oracle.reduced_tmul(sargs, jargs, go, Fg, w)
    rt = torch.empty(nred, dtype=torch.float64, device="cuda")
    dev.reduced_residual_(rt, h, Fr)
    rto = np.zeros(nred)
    oracle.reduced_residual(jargs, rto, Frg)
    D = torch.empty(int((st.r_dims ** 2).sum()), dtype=torch.float64, device="cuda")
    dev.reduced_diagonal_blocks_(D, h, F)
    Do = np.zeros(int((st.r_dims ** 2).sum()))
    oracle.reduced_diagonal_blocks(sargs, jargs, Do, Fg)
    torch.cuda.synchronize()
    for name, a, b in (("mul", y, yo), ("tmul", g, go), ("residual", rt, rto), ("diag", D, Do)):
        a = a.cpu().numpy()
        assert np.abs(a - b).max() <= TOL * max(1.0, np.abs(b).max()), name
    assert np.array_equal(rt.cpu().numpy(), rto)


def test_qr_conditioning_scale(blsamd):
    """reduce_qr! with a conditioning scale s (src/qr_reduction.jl:162-255) on the device against the oracle, on
    eliminated and free tables."""
    import torch
    from test_oracle_qr import _pow2_scale
    dev = blsamd.device
    rng = np.random.default_rng(11)
    for p in (_small_jacobian(blsamd, 8, nland=70, ncam=9, kmax=6), _jac_with_free_rows(blsamd)):
        st, jst = p.structure, p.jst
        s = _pow2_scale(rng, st.nr + st.ne) * rng.uniform(0.5, 2.0, st.nr + st.ne)
        h = dev.QRHandle(jst)
        ws = dev.QRReductionWorkspace(h)
        T = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64)).cuda()
        ok = dev.reduce_qr_(ws, h, T(p.data), T(p.r), T(p.scaling_e), 0.25, s=T(s))
        torch.cuda.synchronize()
        sargs = oracle.SchurArgs(st, blsamd.structure.SchurReductionWorkspaceIndex(st, 1))
        Fo = np.zeros(jst.raw_length()); Fro = np.zeros(jst.residual_length())
        oko, fbo = oracle.reduce_qr(sargs, oracle.JacobianArgs(jst), Fo, Fro, p.data, p.r, p.scaling_e, 0.25, cs=s)
        assert ok and oko
        compare_tables(p, ws.factored.cpu().numpy(), ws.fres.cpu().numpy(), Fo, Fro)


@pytest.mark.parametrize("table_type", ["f64", "f32"])
@pytest.mark.parametrize("ncam,kmax", [(9, 6), (40, 30)])
def test_qr_backsubstitute_matches_oracle(blsamd, table_type, ncam, kmax):
    """backsubstitute!(x_e, x_r, ws::QRReductionWorkspace) (src/qr_reduced.jl:24-110) on the device from the
    device-factored tables, against the oracle on the same tables (1e-12 relative; Float64 accumulation for both
    table types). kmax 30: tables of up to 30 retained entries (63 rows), so the sixteen lanes of a table take
    several entries each."""
    import torch
    dev = blsamd.device
    p = _small_jacobian(blsamd, 12, nland=90, ncam=ncam, kmax=kmax)
    if kmax > 16:
        assert np.diff(p.jst.table_ptr).max() > 16
    jst, st = p.jst, p.structure
    dt = torch.float64 if table_type == "f64" else torch.float32
    h = dev.QRHandle(jst)
    ws = dev.QRReductionWorkspace(h, dtype=dt)
    T = lambda a, d=torch.float64: torch.as_tensor(np.ascontiguousarray(a)).to(d).cuda()
    assert dev.reduce_qr_(ws, h, T(p.data, dt), T(p.r), T(p.scaling_e), 0.2)
    x_r = np.random.default_rng(3).standard_normal(st.nr)
    x_e = torch.empty(st.ne, dtype=torch.float64, device="cuda")
    dev.qr_backsubstitute_(x_e, T(x_r), h, ws)
    torch.cuda.synchronize()
    F, Fr = ws.factored.cpu().numpy(), ws.fres.cpu().numpy()
    sargs = oracle.SchurArgs(st, blsamd.structure.SchurReductionWorkspaceIndex(st, 1))
    xo = np.zeros(st.ne)
    oracle.qr_backsubstitute(sargs, oracle.JacobianArgs(jst), xo, x_r, F, Fr)
    assert np.abs(x_e.cpu().numpy() - xo).max() <= TOL * np.abs(xo).max()


def test_reduced_normal_rejects_too_wide_tables(blsamd):
    """A free table over 70 cameras is past reduced_normal's 64 retained blocks per table: the call raises a clear
    error from the host-side plan, before any launch."""
    import torch
    from test_oracle_qr import _qr_pattern
    dev = blsamd.device
    p = _jac_with_free_rows(blsamd, seed=17, nland=40, ncam=72, wide=70)
    h = dev.QRHandle(p.jst)
    h.set_sparse(_qr_pattern(blsamd, p))
    ws = dev.QRReductionWorkspace(h)
    T = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64)).cuda()
    assert dev.reduce_qr_(ws, h, T(p.data), T(p.r), T(p.scaling_e), 0.2)
    S = torch.empty(_qr_pattern(blsamd, p).nnz, dtype=torch.float64, device="cuda")
    g = torch.empty(p.structure.nr, dtype=torch.float64, device="cuda")
    with pytest.raises(blsamd.device.BLSError, match="64"):
        dev.reduced_normal_sparse_(S, g, h, ws)


def _gram_classes(jst):
    """Launch classes of reduced_normal's Gram pass that the tables of jst fall in: staged size (reduced rows rounded
    up to 4) x (entry columns rounded up to 16, plus one) against 4096, 9472 and 19456 doubles."""
    tp = np.asarray(jst.table_ptr) - 1
    ec = np.asarray(jst.entry_col) - 1
    rd = np.asarray(jst.schur.r_dims)[np.asarray(jst.entry_retained) - 1]
    out = set()
    for t in range(jst.ntables):
        u0, u1 = tp[t], tp[t + 1]
        if u1 == u0:
            continue
        W = ec[u1 - 1] + rd[u1 - 1] - ec[u0]
        n = jst.table_nrows[t] - jst.table_top[t]
        st = ((n + 3) // 4 * 4) * ((W + 15) // 16 * 16 + 1)
        out.add(next((k for k, cap in enumerate((4096, 9472, 19456)) if st <= cap), 3))
    return out


@pytest.mark.parametrize("table_type", ["f64", "f32"])
def test_reduced_normal_matches_oracle(blsamd, table_type):
    """reduced_normal_dense! / reduced_normal_sparse! (src/qr_reduced.jl:373-480) on the device from the
    device-factored tables (eliminated and free ones), against the oracle on the same tables. Float64: 1e-12
    relative Frobenius per block. Float32 (S in Float32, sums in Float64 before the per-table rounding): within
    64 eps(Float32) of the magnitude J~'J~ of |F| per block; g_S (Float64 accumulation) within 1e-12."""
    import torch
    from helpers import dense_upper_blocks, sparse_upper_blocks, max_block_rel, max_block_rel_to
    from test_oracle_qr import _qr_pattern
    dev = blsamd.device
    # the third problem puts about 250 tables on every camera: its columns are split into parts and merged
    # the fourth problem adds camera-pair tables of 210, 300 and 1200 rows: with the landmark tables they fill all
    # four launch classes of the Gram pass (staged block up to 4096, 9472, 19456 doubles, and global memory)
    tall = _jac_with_free_rows(blsamd, seed=9, tall_pairs=((0, 1, 70), (2, 3, 100), (4, 5, 400)))
    assert _gram_classes(tall.jst) == {0, 1, 2, 3}
    # the fifth has a free table over 40 cameras (40 retained blocks, about 300 columns)
    wide = _jac_with_free_rows(blsamd, seed=13, nland=60, ncam=45, wide=40)
    for p in (_small_jacobian(blsamd, 14, nland=80, ncam=9, kmax=6), _jac_with_free_rows(blsamd),
              _small_jacobian(blsamd, 15, nland=500, ncam=6, kmax=4), tall, wide):
        jst, st = p.jst, p.structure
        dt = torch.float64 if table_type == "f64" else torch.float32
        npdt = np.float64 if table_type == "f64" else np.float32
        h = dev.QRHandle(jst)
        pat = _qr_pattern(blsamd, p)
        h.set_sparse(pat)
        ws = dev.QRReductionWorkspace(h, dtype=dt)
        T = lambda a, d=torch.float64: torch.as_tensor(np.ascontiguousarray(a)).to(d).cuda()
        assert dev.reduce_qr_(ws, h, T(p.data, dt), T(p.r), T(p.scaling_e), 0.2)
        Sd = torch.full((st.nr * st.nr,), float("nan"), dtype=dt, device="cuda")
        Ss = torch.full((pat.nnz,), float("nan"), dtype=dt, device="cuda")
        gd = torch.empty(st.nr, dtype=torch.float64, device="cuda")
        gs = torch.empty(st.nr, dtype=torch.float64, device="cuda")
        dev.reduced_normal_dense_(Sd, gd, h, ws)
        dev.reduced_normal_sparse_(Ss, gs, h, ws)
        Ss2 = torch.full_like(Ss, float("nan"))
        gs2 = torch.empty_like(gs)
        dev.reduced_normal_sparse_(Ss2, gs2, h, ws)   # a second call: bitwise the same (fixed-order sums)
        torch.cuda.synchronize()
        assert torch.equal(Ss, Ss2) and torch.equal(gs, gs2)
        F, Fr = ws.factored.cpu().numpy(), ws.fres.cpu().numpy()
        sargs = oracle.SchurArgs(st, blsamd.structure.SchurReductionWorkspaceIndex(st, 1), pat)
        jargs = oracle.JacobianArgs(jst)
        So, go = np.zeros(st.nr * st.nr, npdt), np.zeros(st.nr)
        oracle.reduced_normal(False, sargs, jargs, So, go, F, Fr)
        Sdh = Sd.cpu().numpy()
        # the dense variant zeroes whole scalar columns, lower triangle included
        L = np.tril(Sdh.reshape(st.nr, st.nr).T, -1)
        assert not np.isnan(Sdh).any() and np.all(L == 0)
        A, B = dense_upper_blocks(st, Sdh.astype(np.float64)), dense_upper_blocks(st, So.astype(np.float64))
        As = sparse_upper_blocks(st, pat, Ss.cpu().numpy().astype(np.float64))
        if table_type == "f64":
            assert max_block_rel(A, B) <= TOL
            assert max_block_rel({k: As[k] for k in As}, {k: B[k] for k in As}) <= TOL
        else:
            Mo = np.zeros(st.nr * st.nr)
            oracle.reduced_normal(False, sargs, jargs, Mo, np.zeros(st.nr), np.abs(F).astype(np.float64), np.abs(Fr))
            Mag = dense_upper_blocks(st, Mo)
            eps = 64 * float(np.finfo(np.float32).eps)
            assert max_block_rel_to(A, B, Mag) <= eps
            assert max_block_rel_to({k: As[k] for k in As}, {k: B[k] for k in As}, Mag) <= eps
        for gg in (gd, gs):
            assert np.abs(gg.cpu().numpy() - go).max() <= TOL * np.abs(go).max()


@pytest.mark.parametrize("case", ["f64", "f32", "free"])
def test_reduced_jacobian_matches_oracle(blsamd, case):
    """build_reduced_jacobian_pattern, check_reduced_jacobian_pattern and reduced_jacobian! (sparse and dense,
    src/qr_reduced.jl:247-370) through the C ABI: the pattern equal to the oracle's, every stored entry written (NaN
    before), values bitwise equal to the oracle's copies of the device-factored tables, rt == reduced_residual!, and
    the reference's rejections (test/schur_qr.jl:257-269)."""
    import torch
    import jacobian_ref as JR
    dev = blsamd.device
    p = _jac_with_free_rows(blsamd, seed=12) if case == "free" else _small_jacobian(blsamd, 6, nland=70, ncam=8)
    jst, st = p.jst, p.structure
    dt = torch.float32 if case == "f32" else torch.float64
    h = dev.QRHandle(jst)
    ws = dev.QRReductionWorkspace(h, dtype=dt)
    T = lambda a, d=torch.float64: torch.as_tensor(np.ascontiguousarray(a)).to(d).cuda()
    assert dev.reduce_qr_(ws, h, T(p.data, dt), T(p.r), T(p.scaling_e), 0.3)
    m = int(jst.reduced_nrows())
    colptr, rowval = dev.build_reduced_jacobian_pattern(h)
    co, ro = JR.build_reduced_jacobian_pattern(jst)
    assert np.array_equal(colptr, co) and np.array_equal(rowval, ro)
    dev.check_reduced_jacobian_pattern(h, m, st.nr, colptr, rowval)
    nz = torch.full((len(rowval),), float("nan"), dtype=dt, device="cuda")
    rt = torch.full((m,), float("nan"), dtype=torch.float64, device="cuda")
    dev.reduced_jacobian_(nz, rt, h, ws, sparse=True)
    Jd = torch.full((m * st.nr,), float("nan"), dtype=dt, device="cuda")
    rt2 = torch.full((m,), float("nan"), dtype=torch.float64, device="cuda")
    dev.reduced_jacobian_(Jd, rt2, h, ws, sparse=False)
    F = ws.factored.cpu().numpy()
    rto = torch.empty(m, dtype=torch.float64, device="cuda")
    dev.reduced_residual_(rto, h, ws.fres)
    torch.cuda.synchronize()
    assert np.array_equal(nz.cpu().numpy(), JR.reduced_jacobian_sparse(jst, F, co))
    assert np.array_equal(Jd.cpu().numpy().reshape(st.nr, m).T, JR.reduced_jacobian_dense(jst, F))
    assert np.array_equal(rt.cpu().numpy(), rto.cpu().numpy()) and np.array_equal(rt2.cpu().numpy(), rto.cpu().numpy())
    # a matrix without the pattern is rejected
    with pytest.raises(dev.BLSError, match="DimensionMismatch"):
        dev.check_reduced_jacobian_pattern(h, m, st.nr + 1, np.concatenate([colptr, colptr[-1:]]), rowval)
    full = np.arange(st.nr + 1, dtype=np.int64) * m + 1                         # every column full: wrong heights
    with pytest.raises(dev.BLSError, match="column 1 stores"):
        dev.check_reduced_jacobian_pattern(h, m, st.nr, full, np.tile(np.arange(1, m + 1), st.nr))
    bad = rowval.copy()
    c = int(np.nonzero(np.diff(colptr))[0][0])
    bad[colptr[c] - 1] = 1 if bad[colptr[c] - 1] == m else m                    # a run that starts at the wrong row
    with pytest.raises(dev.BLSError, match="stores row"):
        dev.check_reduced_jacobian_pattern(h, m, st.nr, colptr, bad)
