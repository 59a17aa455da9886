"""GPU parity of the conditioning scale (bls_prepare_scale_f64, src/conditioning.jl:24-87), getdiag! over a Schur
Hessian's storage (bls_asm_diag_f64) and the table quantities jtr!, jtj_diag!, jtj_quadform (bls_jtr_*,
bls_jtj_diag_*, bls_jtj_quadform_*, src/qr_reduced.jl:574-645) against the restatements in oracle/.

Tolerances: the scale bitwise equal to the oracle's on the same diagonal (every operation is IEEE-rounded); getdiag!
bitwise (copies); the table quantities within 1e-12 of the largest value (FP64 sums in another order)."""
import numpy as np
import pytest

import assembly_ref as A
import conditioning_ref as C
import jacobian_ref as JR
from helpers import assembly_problem
from test_conditioning import _drift_sequence, _scale_cases
from test_oracle_qr import _small_jacobian

pytestmark = pytest.mark.gpu


def _T(a, dtype=np.float64):
    import torch
    return torch.as_tensor(np.ascontiguousarray(a, dtype=dtype)).cuda()


@pytest.mark.parametrize("elem", ["f64", "f32"])
def test_scale_matches_oracle_bitwise(blsamd, elem):
    """A fresh fill: s_j = conditioning_scale(d_j + lambda sc_j) for d over 1e-40 .. 1e40, both signs, zeros, powers
    of two and their neighbours, inf and NaN; with sc and without (ones); Float32's floor for elem f32."""
    import torch
    dev = blsamd.device
    rng = np.random.default_rng(4)
    d = np.concatenate([_scale_cases(rng), [np.inf, -np.inf, np.nan]])
    n = len(d)
    sc = rng.uniform(0.25, 4.0, n)
    T, tt = (np.float64, None) if elem == "f64" else (np.float32, torch.float32)
    for lam, scv in ((0.0, None), (0.37, sc), (-1.0, sc), (3e-7, None)):
        c = dev.EliminationConditioning(n)
        s = dev.prepare_scale_(c, lambda dd: dd.copy_(_T(d)), lam, None if scv is None else _T(scv), tt)
        co = C.EliminationConditioning(n)
        so = C.prepare_scale(co, lambda dd: dd.__setitem__(slice(None), d), lam, scv, T)
        assert np.array_equal(s.cpu().numpy(), so), lam
        assert not c.stale


def test_prepare_scale_band_and_recomputation(blsamd):
    """test/conditioning.jl:72-80 on the device: the same scale within the band, one recomputation far outside it,
    the first scale again on returning (nrescale 0, 0, 1, 2); diag! only while stale; markstale! recomputes."""
    dev = blsamd.device
    rng = np.random.default_rng(1)
    n = 70000    # more than one workgroup of the drift test
    d = rng.uniform(0.1, 10.0, n) * 10.0 ** rng.integers(-2, 3, n)
    sc = rng.uniform(0.5, 2.0, n)
    c = dev.EliminationConditioning(n)
    calls = []

    def diag_(dd):
        calls.append(1)
        dd.copy_(_T(d))
    scd = _T(sc)
    prep = lambda lam: dev.prepare_scale_(c, diag_, lam, scd).cpu().numpy()
    s0 = _drift_sequence(prep, d, c)
    assert len(calls) == 1
    assert np.array_equal(s0, C.conditioning_scale(d))
    c.markstale()
    assert c.stale
    s = prep(0.5)
    assert len(calls) == 2 and np.array_equal(s, C.conditioning_scale(d + 0.5 * sc))


@pytest.mark.parametrize("sparse", [False, True])
def test_hessian_diag_matches_oracle(blsamd, sparse):
    """getdiag! over the device-assembled storage against the oracle's getdiag of the oracle's assembly."""
    import torch
    dev = blsamd.device
    st, rowblocks, rb_rows, J, r, R_layout = assembly_problem(blsamd, 8, nland=60, sparse=sparse)
    h = dev.AssemblyHandle(st, rowblocks, rb_rows, R_layout)
    z = lambda n: torch.zeros(n, dtype=torch.float64, device="cuda")
    E, Cd, R, br, be = z(h.E_len), z(h.C_len), z(h.R_len), z(h.nr), z(int(np.sum(st.e_dims)))
    dev.assemble_hessian_(E, Cd, R, br, be, h, _T(J), _T(r))
    d = torch.full((h.nr + int(np.sum(st.e_dims)),), float("nan"), dtype=torch.float64, device="cuda")
    dev.hessian_diag_(d, h, E, R)
    # the oracle's getdiag on the device's own storage: a copy, so bitwise
    H = A.HessianStorage(st, *(R_layout or ()))
    H.E, H.C = E.cpu().numpy(), Cd.cpu().numpy()
    H.R = R.cpu().numpy().reshape(h.nr, h.nr, order="F") if R_layout is None else R.cpu().numpy()
    assert np.array_equal(d.cpu().numpy(), A.getdiag(H))
    # and within 1e-12 of the oracle's assembly
    Ho = A.assemble(A.HessianStorage(st, *(R_layout or ())), rowblocks, rb_rows, J, r)
    ref = A.getdiag(Ho)
    assert np.abs(d.cpu().numpy() - ref).max() <= 1e-12 * np.abs(ref).max()


def test_hessian_diag_missing_block_reads_zero(blsamd):
    """A sparse R without the diagonal block of retained block 1 (a layout-only handle): zeros there."""
    dev = blsamd.device
    st, _, _, _, _, (cp, rv, ro) = assembly_problem(blsamd, 4, nland=10, sparse=True)
    keep = [k for k in range(len(rv)) if not (rv[k] == 1 and k < cp[1] - 1)]
    cp2 = np.array([1] + [1 + sum(1 for k in keep if k < cp[q + 1] - 1) for q in range(len(cp) - 1)])
    layout = (cp2, np.asarray(rv)[keep], np.asarray(ro)[keep])
    h = dev.AssemblyHandle(st, [], [], layout)
    H = A.HessianStorage(st, *layout)
    rng = np.random.default_rng(6)
    H.R = rng.standard_normal(h.R_len)
    H.E = rng.standard_normal(h.E_len)
    d = _T(np.full(h.nr + int(np.sum(st.e_dims)), np.nan))
    dev.hessian_diag_(d, h, _T(H.E), _T(H.R))
    out = d.cpu().numpy()
    assert np.array_equal(out, A.getdiag(H)) and np.all(out[:int(st.r_dims[0])] == 0.0)


@pytest.mark.parametrize("tables", ["f64", "f32", "free"])
def test_table_quantities_match_oracle(blsamd, tables):
    """jtr!, jtj_diag!, jtj_quadform on the raw tables (Float64 and Float32 tables; eliminated and free tables)."""
    import torch
    from test_gpu_qr import _jac_with_free_rows
    dev = blsamd.device
    p = _jac_with_free_rows(blsamd) if tables == "free" else _small_jacobian(blsamd, 5, nland=120, ncam=10, kmax=7)
    jst, st = p.jst, p.structure
    dt = np.float32 if tables == "f32" else np.float64
    data = p.data.astype(dt)
    h = dev.QRHandle(jst)
    F, r = _T(data, dt), _T(p.r)
    n = st.nr + st.ne
    v = np.random.default_rng(9).standard_normal(n)
    g = torch.full((n,), float("nan"), dtype=torch.float64, device="cuda")
    d = torch.full((n,), float("nan"), dtype=torch.float64, device="cuda")
    dev.jtr_(g, h, F, r)
    dev.jtj_diag_(d, h, F)
    q = dev.jtj_quadform(h, F, _T(v))
    go, do, qo = JR.jtr(jst, data, p.r), JR.jtj_diag(jst, data), JR.jtj_quadform(jst, data, v)
    for name, a, b in (("jtr", g.cpu().numpy(), go), ("jtj_diag", d.cpu().numpy(), do)):
        assert not np.isnan(a).any(), name
        assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max(), name
    assert abs(q - qo) <= 1e-12 * qo
    assert dev.jtj_quadform(h, F, _T(v)) == q          # fixed order: bitwise reproducible


@pytest.mark.parametrize("sparse", [False, True])
def test_hessian_dot_and_diag_float32(blsamd, sparse):
    """dot(x, H, y) and getdiag! over a SchurBlockHessian{Float32}'s storage: the oracle on the same Float32 values
    (widened): getdiag! bitwise, dot within 1e-12 of the sum of the terms' magnitudes (both sum in Float64)."""
    import torch
    dev = blsamd.device
    st, rowblocks, rb_rows, J, r, R_layout = assembly_problem(blsamd, 9, nland=50, sparse=sparse)
    h = dev.AssemblyHandle(st, rowblocks, rb_rows, R_layout)
    Ho = A.assemble(A.HessianStorage(st, *(R_layout or ())), rowblocks, rb_rows, J, r)
    E32 = Ho.E.astype(np.float32)
    C32 = Ho.C.astype(np.float32)
    R32 = (Ho.R.reshape(-1, order="F") if R_layout is None else Ho.R).astype(np.float32)
    H = A.HessianStorage(st, *(R_layout or ()))
    H.E, H.C = E32.astype(np.float64), C32.astype(np.float64)
    H.R = R32.astype(np.float64).reshape(h.nr, h.nr, order="F") if R_layout is None else R32.astype(np.float64)
    n = h.nr + int(np.sum(st.e_dims))
    rng = np.random.default_rng(10)
    x, y = rng.standard_normal(n), rng.standard_normal(n)
    Ed, Cd, Rd = _T(E32, np.float32), _T(C32, np.float32), _T(R32, np.float32)
    ref = A.bilinear(H, x, y)
    Habs = A.HessianStorage(st, *(R_layout or ()))
    Habs.E, Habs.C, Habs.R = np.abs(H.E), np.abs(H.C), np.abs(H.R)
    scale = A.bilinear(Habs, np.abs(x), np.abs(y))          # the sum of the terms' magnitudes
    assert abs(dev.hessian_dot(_T(x), h, _T(y), Ed, Cd, Rd) - ref) <= 1e-12 * scale
    d = torch.full((n,), float("nan"), dtype=torch.float64, device="cuda")
    dev.hessian_diag_(d, h, Ed, Rd)
    assert np.array_equal(d.cpu().numpy(), A.getdiag(H))
