"""GPU parity of the table fill: block(J, k, blockentry(jst, k, j)) .= J_kj and residual_block(J, k) .= r_k for every
row block (src/schur_jacobian.jl:264-387) by bls_jfill_*, against the oracle's fill_tables (itself pinned against
the generator's tables and the dense row-block J in test_oracle_qr.py). A fill is a copy: bit for bit."""
import numpy as np
import pytest

import jacobian_ref as JR
from test_gpu_qr import _jac_with_free_rows
from test_oracle_qr import _rowblock_values

pytestmark = pytest.mark.gpu


def _device_fill(blsamd, jst, vals, res, raw0, r0, dtype):
    import torch
    dev = blsamd.device
    fill = dev.JacobianFill(jst)
    assert fill.nvals == len(vals) and fill.nres == len(res)
    raw = torch.as_tensor(raw0).to(dtype).cuda()
    r = torch.as_tensor(r0).cuda()
    dev.fill_jacobian_(raw, r, fill, torch.as_tensor(vals).to(dtype).cuda(), torch.as_tensor(res).cuda())
    torch.cuda.synchronize()
    return raw.cpu().numpy(), r.cpu().numpy()


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_fill_reproduces_generator_tables(blsamd, dtype):
    """The config-4 shaped generator's per-observation blocks (cameras of 6 and 9, landmarks of 3): the device fill
    writes the tables the generator laid out itself."""
    import torch
    dt = torch.float64 if dtype == "f64" else torch.float32
    p = blsamd.generators.config4(scale=0.001, ncam=60, kmax=30)
    jst, st = p.jst, p.structure
    vals, res = [], []
    for k in range(len(p.cam_of_obs)):
        d = int(st.r_dims[p.cam_of_obs[k] - 1])
        vals += [p.Jc[k, :, :d].T.reshape(-1), p.Jl[k].T.reshape(-1)]
        res.append(p.rr[k])
    vals, res = np.concatenate(vals), np.concatenate(res)
    raw, r = _device_fill(blsamd, jst, vals, res, np.zeros(jst.raw_length()), np.zeros(jst.residual_length()), dt)
    want = p.data if dtype == "f64" else p.data.astype(np.float32)
    assert np.array_equal(raw, want) and np.array_equal(r, p.r)


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_fill_free_tables_matches_oracle(blsamd, dtype):
    """Camera priors and camera-pair rows (free tables), mixed camera sizes, a 70-block free row block: the device
    fill equals the oracle's bit for bit, and leaves every other row (damping rows) as it was."""
    import torch
    dt = torch.float64 if dtype == "f64" else torch.float32
    npdt = np.float64 if dtype == "f64" else np.float32
    for p in (_jac_with_free_rows(blsamd, seed=4), _jac_with_free_rows(blsamd, seed=5, nland=30, ncam=72, wide=70),
              _jac_with_free_rows(blsamd, seed=6, tall_pairs=((0, 1, 30),))):
        jst = p.jst
        vals, res, _, _ = _rowblock_values(jst, np.random.default_rng(2))
        vals = vals.astype(npdt).astype(np.float64)
        raw0 = np.full(jst.raw_length(), 7.0)
        r0 = np.full(jst.residual_length(), -3.0)
        rawo, ro = raw0.copy(), r0.copy()
        JR.fill_tables(jst, vals, res, rawo, ro)
        raw, r = _device_fill(blsamd, jst, vals, res, raw0, r0, dt)
        assert np.array_equal(raw.astype(np.float64), rawo) and np.array_equal(r, ro)


def test_fill_rejects_wrong_lengths(blsamd):
    import torch
    dev = blsamd.device
    p = _jac_with_free_rows(blsamd, seed=4)
    jst = p.jst
    fill = dev.JacobianFill(jst)
    f64 = lambda n: torch.zeros(n, dtype=torch.float64, device="cuda")
    raw, r = f64(jst.raw_length()), f64(jst.residual_length())
    for args in ((raw, r, f64(fill.nvals - 1), f64(fill.nres)), (raw, r, f64(fill.nvals), f64(fill.nres + 1)),
                 (f64(jst.raw_length() - 1), r, f64(fill.nvals), f64(fill.nres)),
                 (raw.float(), r, f64(fill.nvals), f64(fill.nres))):
        with pytest.raises(dev.BLSError):
            dev.fill_jacobian_(args[0], args[1], fill, args[2], args[3])
