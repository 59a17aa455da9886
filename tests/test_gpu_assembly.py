"""GPU assembly of a Schur Hessian from row blocks (bls_asm_hessian_f64, the batched addblock! of
src/schur_assembly.jl:140-246) against the restatement in oracle/assembly_ref.py: every stored value within 1e-12 of
the largest, b within 1e-12, bitwise reproducible, and the reference's no-storage errors at registration."""
import numpy as np
import pytest

import assembly_ref as A
from helpers import assembly_problem

pytestmark = pytest.mark.gpu


def _run(blsamd, st, rowblocks, rb_rows, J, r, R_layout):
    import torch
    dev = blsamd.device
    h = dev.AssemblyHandle(st, rowblocks, rb_rows, R_layout)
    nan = lambda n: torch.full((n,), float("nan"), dtype=torch.float64, device="cuda")
    E, C, R, br, be = nan(h.E_len), nan(h.C_len), nan(h.R_len), nan(h.nr), nan(int(np.sum(st.e_dims)))
    T = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64)).cuda()
    Jd, rd = T(J), T(r)
    dev.assemble_hessian_(E, C, R, br, be, h, Jd, rd)
    out = [x.cpu().numpy() for x in (E, C, R, br, be)]
    E2, C2, R2, br2, be2 = nan(h.E_len), nan(h.C_len), nan(h.R_len), nan(h.nr), nan(int(np.sum(st.e_dims)))
    dev.assemble_hessian_(E2, C2, R2, br2, be2, h, Jd, rd)
    torch.cuda.synchronize()
    again = [x.cpu().numpy() for x in (E2, C2, R2, br2, be2)]
    for a, b in zip(out, again):
        assert np.array_equal(a, b)
    return out


@pytest.mark.parametrize("sparse", [False, True])
@pytest.mark.parametrize("case", ["light", "heavy", "tall", "shared"])
def test_assembly_matches_oracle(blsamd, sparse, case):
    """Dense and sparse R. Observation row blocks (2 rows, <= 16 columns) take the 16-lane path, 9-row priors and
    wide pair rows the Gram pass. light: 20 landmarks, every destination block has <= 32 contributions (one
    wavefront per block); heavy: 150 landmarks, the camera blocks of R and b_r have more (runs over a workgroup);
    tall: camera-pair row blocks of 250, 400 and 1500 rows put row blocks in every launch class of the Gram pass.
    A coupling block with one contribution from an observation is written in place by the 16-lane path; shared:
    couplings with two contributions (an observation made twice) and couplings from row blocks over a landmark and
    two cameras, which take the ordered sums."""
    nland, tall, dup, triple = {"light": (20, (), 0, 0), "heavy": (150, (), 0, 0), "tall": (150, (250, 400, 1500), 0, 0),
                                "shared": (20, (), 6, 5)}[case]
    st, rowblocks, rb_rows, J, r, R_layout = assembly_problem(blsamd, 5 + nland, nland=nland, sparse=sparse, tall=tall,
                                                              dup=dup, triple=triple)
    E, C, R, br, be = _run(blsamd, st, rowblocks, rb_rows, J, r, R_layout)
    H = A.assemble(A.HessianStorage(st, *(R_layout or ())), rowblocks, rb_rows, J, r)
    Ro = H.R.reshape(-1, order="F") if R_layout is None else H.R
    for name, a, b in (("E", E, H.E), ("C", C, H.C), ("R", R, Ro), ("b_r", br, H.b_r), ("b_e", be, H.b_e)):
        assert not np.isnan(a).any(), name
        assert np.abs(a - b).max() <= 1e-12 * max(1.0, np.abs(b).max()), name


def test_assembly_rejects_pairs_without_storage(blsamd):
    """check_addblock (src/schur_assembly.jl:161-181): a row block over two eliminated blocks, or over a retained
    and an eliminated block that share no row block of the structure, fails at registration."""
    dev = blsamd.device
    st = blsamd.structure.SchurStructure.from_description([3, 3, 6, 6], [[1, 3], [2, 4]], [True, True, False, False])
    p = np.argsort(np.asarray(st.blockorder()) - 1) + 1
    with pytest.raises(dev.BLSError, match="both eliminated"):
        dev.AssemblyHandle(st, [[int(p[0]), int(p[1])]], [2])
    with pytest.raises(dev.BLSError, match="share no row block"):
        dev.AssemblyHandle(st, [[int(p[0]), int(p[3])]], [2])


@pytest.mark.parametrize("sparse", [False, True])
def test_hessian_dot_matches_oracle(blsamd, sparse):
    """dot(x, H, y) (src/linalg.jl:133-260) on the device, from a layout-only handle (no row blocks), over the
    storage the oracle assembled: against the restatement of schur_bilinear within 1e-12 of the sum of the terms'
    magnitudes, for x != y and x == y, and bitwise reproducible."""
    import torch
    dev = blsamd.device
    st, rowblocks, rb_rows, J, r, R_layout = assembly_problem(blsamd, 9, nland=80, sparse=sparse)
    H = A.assemble(A.HessianStorage(st, *(R_layout or ())), rowblocks, rb_rows, J, r)
    h = dev.AssemblyHandle(st, [], [], R_layout)
    T = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64)).cuda()
    Ro = H.R.reshape(-1, order="F") if R_layout is None else H.R
    E, C, R = T(H.E), T(H.C), T(Ro)
    n = H.nr + int(np.sum(st.e_dims))
    rng = np.random.default_rng(2)
    x, y = rng.standard_normal(n), rng.standard_normal(n)
    Habs = A.HessianStorage(st, *(R_layout or ()))
    Habs.E, Habs.C, Habs.R = np.abs(H.E), np.abs(H.C), np.abs(H.R)
    for a, b in ((x, y), (x, x)):
        got = dev.hessian_dot(T(a), h, T(b), E, C, R)
        want = A.bilinear(H, a, b)
        scale = A.bilinear(Habs, np.abs(a), np.abs(b))
        assert abs(got - want) <= 1e-12 * scale
        assert dev.hessian_dot(T(a), h, T(b), E, C, R) == got


@pytest.mark.parametrize("de,dr", [(3, 6), (2, 9), (4, 7), (1, 6)])
def test_hessian_dot_contiguous_windows(blsamd, de, dr):
    """dot(x, H, y) on a layout of uniform blocks whose landmarks couple contiguous camera windows (the
    coalesced per-row path of the C ABI's dot), against x' H y with H formed densely from the same storage
    (R's diagonal blocks, E, and each coupling block in both triangles): within 1e-12 of |x|' |H| |y|, bitwise
    reproducible."""
    import torch
    dev = blsamd.device
    p = blsamd.generators.window_problem(seed=de * 10 + dr, ncam=24, nland=300, de=de, dr=dr,
                                         kmin=3 if de == 4 else 2, dense=True)
    st = p.structure
    nr, ne = st.nr, st.ne
    n = nr + ne
    H = np.zeros((n, n))
    H[:nr, :nr] = p.R
    for e in range(st.neblocks):
        d = st.e_dims[e]
        eo, bo = st.eoffsets[e] - 1, nr + st.boffsets_e[e] - 1
        H[bo:bo + d, bo:bo + d] = p.E_data[eo:eo + d * d].reshape(d, d, order="F")
        for c in range(st.eliminated_ptr[e] - 1, st.eliminated_ptr[e + 1] - 1):
            r = st.entry_retained[c] - 1
            ro, dr_ = st.boffsets_r[r] - 1, st.r_dims[r]
            co = st.coffsets[c] - 1
            blk = p.C_data[co:co + d * dr_].reshape(d, dr_, order="F")   # d_e x d_r column-major
            H[bo:bo + d, ro:ro + dr_] = blk
            H[ro:ro + dr_, bo:bo + d] = blk.T
    h = dev.AssemblyHandle(st, [], [], None)
    T = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64)).cuda()
    E, C, R = T(p.E_data), T(p.C_data), T(p.R.reshape(-1, order="F"))
    rng = np.random.default_rng(de + dr)
    x, y = rng.standard_normal(n), rng.standard_normal(n)
    for a, b in ((x, y), (x, x)):
        got = dev.hessian_dot(T(a), h, T(b), E, C, R)
        want = a @ H @ b
        scale = np.abs(a) @ np.abs(H) @ np.abs(b)
        assert abs(got - want) <= 1e-12 * scale
        assert dev.hessian_dot(T(a), h, T(b), E, C, R) == got
