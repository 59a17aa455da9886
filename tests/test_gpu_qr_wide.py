"""GPU parity of the QR path on tables the register kernel does not take: eliminated tables of more than 64 rows
(tracks of more than 30 observations) and eliminated blocks of d_e > 8. qr_table! (src/qr_reduction.jl:50-131) and
backsubstitute! (src/qr_reduced.jl:24-110) have no size limit; the device works such tables by one workgroup each
(k_qr_table_wide, k_qr_backsub_wide).

Tolerances as in test_gpu_qr.py: per table ||F_gpu - F_oracle||_F <= 1e-12 ||F_oracle||_F (and the residual), x_e
within 1e-12, the failing block exact; Float32 tables within twice the oracle's own Float32 error plus 4 eps.
"""
import numpy as np
import pytest

import oracle
from test_gpu_precision import EPS32, _qr_oracle, _qr_tables, _table_err
from test_gpu_qr import TOL, compare_tables, run_gpu
from test_oracle_qr import _pow2_scale, _qr, _qr_pattern

pytestmark = pytest.mark.gpu


def _tracks(blsamd, seed, nland, ncam, kmin, kmax, de):
    rng = np.random.default_rng(seed)
    cam_dims = np.where(rng.random(ncam) < 0.5, 6, 9).astype(np.int64)
    k = rng.integers(kmin, kmax + 1, size=nland)
    ptr = np.concatenate([[0], np.cumsum(k)])
    cams = np.concatenate([rng.choice(ncam, size=int(kk), replace=False) for kk in k])
    return blsamd.generators.jacobian_from_tracks(rng, ncam, cam_dims, nland, k, ptr, cams, 0.1, de=de)


# seed, landmarks, cameras, k range, d_e: tracks of up to 45 observations (93 rows) beside short ones; d_e 9 on
# short tables; d_e 12 on long ones
CASES = {"long": (21, 60, 50, 2, 45, 3), "de9": (22, 40, 20, 2, 12, 9), "de12_long": (23, 30, 45, 20, 40, 12)}


def _check_wide(p):
    st, jst = p.structure, p.jst
    M = np.asarray(jst.table_nrows[:st.neblocks])
    assert (M > 64).any() or (np.asarray(st.e_dims) > 8).any()


@pytest.mark.parametrize("case", list(CASES))
def test_wide_tables_match_oracle(blsamd, case):
    p = _tracks(blsamd, *CASES[case])
    _check_wide(p)
    ok, fb, Fg, Frg = run_gpu(blsamd, p, 0.1)
    oko, fbo, Fo, Fro, _, _ = _qr(p, 0.1)
    assert ok and oko and fb == fbo == 0
    compare_tables(p, Fg, Frg, Fo, Fro)


def test_wide_tables_bitwise_reproducible(blsamd):
    p = _tracks(blsamd, *CASES["long"])
    (a, _, Fa, Fra), (b, _, Fb, Frb) = run_gpu(blsamd, p, 0.1, repeat=2)
    assert a and b and np.array_equal(Fa, Fb) and np.array_equal(Fra, Frb)


def test_wide_tables_failure_lowest_block(blsamd):
    """A singular wide table and a singular register-kernel table (eliminated columns zero, lambda = 0): the lower
    one is reported, whichever kernel found it; a NaN in a wide table is a failure of that table."""
    p = _tracks(blsamd, *CASES["long"])
    jst, st = p.jst, p.structure
    M = np.asarray(jst.table_nrows[:st.neblocks])
    wide = [int(e) for e in np.nonzero(M > 64)[0]]
    narrow = [int(e) for e in np.nonzero(M <= 64)[0]]
    assert wide and narrow
    for e in (wide[1], narrow[-1]):
        Me, C, de = jst.table_nrows[e], jst.table_ncols[e] - 1, st.e_dims[e]
        off = jst.raw_offsets[e] - 1
        p.data[off + (C - de) * Me: off + C * Me] = 0.0
    ok, fb, _, _ = run_gpu(blsamd, p, 0.0)
    assert not ok and fb == min(wide[1], narrow[-1]) + 1
    oko, fbo, _, _, _, _ = _qr(p, 0.0)
    assert fbo == fb
    q = _tracks(blsamd, *CASES["long"])
    q.data[jst.raw_offsets[wide[0]] - 1 + 5] = np.nan   # row 5 of column 0: a data row of the first wide table
    ok, fb, _, _ = run_gpu(blsamd, q, 0.5)
    assert not ok and fb == wide[0] + 1


def test_wide_tables_conditioning_scale(blsamd):
    """reduce_qr! with a conditioning scale s (src/qr_reduction.jl:236-255) on wide tables."""
    import torch
    dev = blsamd.device
    rng = np.random.default_rng(31)
    for case in ("long", "de12_long"):
        p = _tracks(blsamd, *CASES[case])
        st, jst = p.structure, p.jst
        s = _pow2_scale(rng, st.nr + st.ne) * rng.uniform(0.5, 2.0, st.nr + st.ne)
        h = dev.QRHandle(jst)
        ws = dev.QRReductionWorkspace(h)
        T = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64)).cuda()
        ok = dev.reduce_qr_(ws, h, T(p.data), T(p.r), T(p.scaling_e), 0.25, s=T(s))
        torch.cuda.synchronize()
        sargs = oracle.SchurArgs(st, blsamd.structure.SchurReductionWorkspaceIndex(st, 1))
        Fo, Fro = np.zeros(jst.raw_length()), np.zeros(jst.residual_length())
        oko, _ = oracle.reduce_qr(sargs, oracle.JacobianArgs(jst), Fo, Fro, p.data, p.r, p.scaling_e, 0.25, cs=s)
        assert ok and oko
        compare_tables(p, ws.factored.cpu().numpy(), ws.fres.cpu().numpy(), Fo, Fro)


def test_wide_tables_f32_within_reference_float32_error(blsamd):
    import torch
    for case in ("long", "de12_long"):
        p = _tracks(blsamd, *CASES[case])
        jst = p.jst
        ok, fb, F, Fr, h = _qr_tables(blsamd, p, torch.float32, 0.1)
        ok64, _, F64, Fr64, _, _ = _qr_oracle(p, np.float64, 0.1, data=p.data.astype(np.float32).astype(np.float64))
        ok32, _, F32, Fr32, _, _ = _qr_oracle(p, np.float32, 0.1)
        assert ok and ok64 and ok32
        e_dev, e_orc = _table_err(jst, F.cpu().numpy(), F64), _table_err(jst, F32, F64)
        assert e_dev <= 2 * e_orc + 4 * EPS32, (case, e_dev, e_orc)


@pytest.mark.parametrize("table_type", ["f64", "f32"])
@pytest.mark.parametrize("case", list(CASES))
def test_wide_backsubstitute_matches_oracle(blsamd, case, table_type):
    import torch
    dev = blsamd.device
    p = _tracks(blsamd, *CASES[case])
    jst, st = p.jst, p.structure
    dt = torch.float64 if table_type == "f64" else torch.float32
    h = dev.QRHandle(jst)
    ws = dev.QRReductionWorkspace(h, dtype=dt)
    T = lambda a, d=torch.float64: torch.as_tensor(np.ascontiguousarray(a)).to(d).cuda()
    assert dev.reduce_qr_(ws, h, T(p.data, dt), T(p.r), T(p.scaling_e), 0.2)
    x_r = np.random.default_rng(3).standard_normal(st.nr)
    x_e = torch.full((st.ne,), float("nan"), dtype=torch.float64, device="cuda")
    dev.qr_backsubstitute_(x_e, T(x_r), h, ws)
    torch.cuda.synchronize()
    F, Fr = ws.factored.cpu().numpy(), ws.fres.cpu().numpy()
    sargs = oracle.SchurArgs(st, blsamd.structure.SchurReductionWorkspaceIndex(st, 1))
    xo = np.zeros(st.ne)
    oracle.qr_backsubstitute(sargs, oracle.JacobianArgs(jst), xo, x_r, F, Fr)
    assert np.abs(x_e.cpu().numpy() - xo).max() <= TOL * np.abs(xo).max()


@pytest.mark.parametrize("case", ["long", "de12_long"])
def test_wide_tables_reduced_products_and_normal(blsamd, case):
    """The reduced products and reduced_normal_sparse! (src/qr_reduced.jl:161-243, :409-480) on device-factored
    wide tables, against the oracle on the same tables."""
    import torch
    from helpers import max_block_rel, sparse_upper_blocks, dense_upper_blocks
    dev = blsamd.device
    p = _tracks(blsamd, *CASES[case])
    jst, st = p.jst, p.structure
    h = dev.QRHandle(jst)
    pat = _qr_pattern(blsamd, p)
    h.set_sparse(pat)
    ws = dev.QRReductionWorkspace(h)
    T = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64)).cuda()
    assert dev.reduce_qr_(ws, h, T(p.data), T(p.r), T(p.scaling_e), 0.2)
    F, Fr = ws.factored, ws.fres
    nred = jst.reduced_nrows()
    rng = np.random.default_rng(4)
    v, w = rng.standard_normal(st.nr), rng.standard_normal(nred)
    y = torch.empty(nred, dtype=torch.float64, device="cuda")
    g = torch.empty(st.nr, dtype=torch.float64, device="cuda")
    dev.reduced_mul_(y, h, F, T(v))
    dev.reduced_tmul_(g, h, F, T(w))
    S = torch.full((pat.nnz,), float("nan"), dtype=torch.float64, device="cuda")
    gS = torch.empty(st.nr, dtype=torch.float64, device="cuda")
    dev.reduced_normal_sparse_(S, gS, h, ws)
    torch.cuda.synchronize()
    Fh, Frh = F.cpu().numpy(), Fr.cpu().numpy()
    sargs = oracle.SchurArgs(st, blsamd.structure.SchurReductionWorkspaceIndex(st, 1), pat)
    jargs = oracle.JacobianArgs(jst)
    yo, go = np.zeros(nred), np.zeros(st.nr)
    oracle.reduced_mul(sargs, jargs, yo, Fh, v)
    oracle.reduced_tmul(sargs, jargs, go, Fh, w)
    assert np.abs(y.cpu().numpy() - yo).max() <= TOL * np.abs(yo).max()
    assert np.abs(g.cpu().numpy() - go).max() <= TOL * np.abs(go).max()
    So, gSo = np.zeros(st.nr * st.nr), np.zeros(st.nr)
    oracle.reduced_normal(False, sargs, jargs, So, gSo, Fh, Frh)
    A = sparse_upper_blocks(st, pat, S.cpu().numpy())
    B = dense_upper_blocks(st, So)
    assert max_block_rel(A, {k: B[k] for k in A}) <= TOL
    assert np.abs(gS.cpu().numpy() - gSo).max() <= TOL * np.abs(gSo).max()
