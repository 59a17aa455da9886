"""CPU tests of the integer layout: the host mirror (vectorised) against the reference's known-answer literals and
against the loop restatement under oracle/ (bit-exact maps)."""
import json
import os

import numpy as np
import pytest

import oracle
import structure_ref as ref

HERE = os.path.dirname(os.path.abspath(__file__))
LIT = json.load(open(os.path.join(HERE, "golden", "reference_literals.json")))


def test_sort_schur_blocks_literal(blsamd):
    L = LIT["sort_schur_blocks"]
    got = blsamd.structure.sort_schur_blocks(L["is_eliminated"], L["block_sizes"], L["couplings"])
    assert got.tolist() == L["expected"]
    assert ref.sort_schur_blocks(L["is_eliminated"], L["block_sizes"], L["couplings"]) == L["expected"]


def test_expand_block_perm_literal(blsamd):
    L = LIT["expand_block_perm"]
    assert blsamd.structure.expand_block_perm(L["block_perm"], L["block_sizes"]).tolist() == L["expected"]


def test_schur_structure_doctest(blsamd):
    L = LIT["schur_structure_doctest"]
    st = blsamd.structure.SchurStructure.from_description(L["blocksizes"], L["rowblocks"], L["eliminated"])
    assert st.blockorder().tolist() == L["blockorder"]
    assert st.nretained() == L["nretained"]
    assert st.blocksizes().tolist() == L["blocksizes_layout"]


def test_jacobian_structure_doctest(blsamd):
    L = LIT["jacobian_structure_doctest"]
    jst = blsamd.structure.SchurJacobianStructure.from_description(L["blocksizes"], L["rowblocks"], L["rowblocksizes"],
                                                                   L["eliminated"])
    assert jst.schur.blockorder().tolist() == L["blockorder"]
    assert jst.schur.nretained() == L["nretained"] and jst.schur.nblocks() == L["nblocks"]


def test_jacobian_layout_literals(blsamd):
    L = LIT["jacobian_layout"]
    S = blsamd.structure
    st = S.SchurStructure(L["r_dims"], L["e_dims"], L["couplings"])
    jst = S.SchurJacobianStructure(st, L["e_rows"], free_rows=L["free_rows"], free_retained=L["free_retained"])
    for k in ("table_nrows", "table_ncols", "table_top", "entry_retained", "entry_col", "entry_table", "table_ptr",
              "reduced_offsets", "doffsets"):
        assert getattr(jst, k).tolist() == L[k], k
    assert jst.table_offsets.tolist() == np.cumsum([1] + list(jst.table_nrows * jst.table_ncols)).tolist()
    assert jst.raw_offsets.tolist() == np.cumsum([1] + list(jst.table_nrows * (jst.table_ncols - 1))).tolist()
    assert jst.residual_offsets.tolist() == np.cumsum([1] + list(jst.table_nrows)).tolist()
    assert [r for (a, b) in jst.col_chunks for r in range(a, b + 1)] == L["col_chunks_flat"]
    for t, r, u in L["coupling_entry"]:
        assert jst.coupling_entry(t, r) == u
    with pytest.raises(ValueError):
        jst.coupling_entry(2, 1)
    with pytest.raises(IndexError):
        jst.coupling_entry(6, 1)
    assert len(S.SchurJacobianStructure(st, L["e_rows"], nchunks=3).col_chunks) == 3
    r = ref.schur_jacobian_structure(ref.schur_structure(L["r_dims"], L["e_dims"], L["couplings"]), L["e_rows"],
                                     L["free_rows"], L["free_retained"])
    for k in ("table_nrows", "table_ncols", "entry_retained", "entry_col", "table_ptr", "reduced_offsets", "doffsets",
              "retained_ptr", "retained_entry", "raw_offsets", "residual_offsets"):
        assert getattr(jst, k).tolist() == r[k], k


def test_jacobian_structure_rejections(blsamd):
    S = blsamd.structure
    st = S.SchurStructure([6, 3, 6], [3, 3, 2], [[3, 1], [2], []])
    with pytest.raises(S.DimensionMismatch):
        S.SchurJacobianStructure(st, [5, 4])
    with pytest.raises(ValueError):
        S.SchurJacobianStructure(st, [5, 0, 2])
    with pytest.raises(ValueError):
        S.SchurJacobianStructure(st, [5, 4, 2], free_rows=[3], free_retained=[[1, 4]])
    with pytest.raises(ValueError):
        S.SchurJacobianStructure(st, [5, 4, 2], free_rows=[3], free_retained=[[2, 1, 2]])
    with pytest.raises(ValueError):
        S.SchurJacobianStructure(st, [5, 4, 2], nchunks=0)


def test_chunks_literal(blsamd):
    L = LIT["chunks_K4"]
    S = blsamd.structure
    assert [(a, b) for (a, b, _d) in S.runs(L["e_dims"])] == [tuple(x) for x in L["runs"]]
    assert [d for (_a, _b, d) in S.runs(L["e_dims"])] == L["run_sizes"]
    cpl = [[1, 2]] * len(L["e_dims"])
    # the literal is hand-built for the internal constructor (test/schur.jl:85-89); its empty ranges (11:10,
    # 11:10) differ in start from the positional constructor's cld rule (11:10, 12:11): compare non-empty ranges
    # exactly and require the others to be empty
    norm = lambda grp: [list(c) if c[0] <= c[1] else "empty" for c in grp]
    for K, key in ((1, "chunks_K1"), (4, "chunks_K4")):
        st = S.SchurStructure([6, 3, 6, 9], L["e_dims"], cpl, nchunks=K)
        assert [norm(g) for g in st.chunks] == [norm(g) for g in L[key]]
        assert st.chunks == ref.schur_structure([6, 3, 6, 9], L["e_dims"], cpl, nchunks=K)["chunks"]


def test_description_rejections(blsamd):
    S = blsamd.structure
    with pytest.raises(ValueError):      # two eliminated blocks in one row block
        S.SchurStructure.from_description([3, 6, 3], [[1, 3]], [True, False, True])
    with pytest.raises(ValueError):      # repeated column block
        S.SchurStructure.from_description([3, 6], [[1, 1, 2]], [True, False])
    with pytest.raises(ValueError):      # out of range
        S.SchurStructure.from_description([3, 6], [[1, 3]], [True, False])
    with pytest.raises(ValueError):      # no retained block
        S.SchurStructure.from_description([3, 6], [[1, 2]], [True, True])
    with pytest.raises(ValueError):      # block size outside 1..65535
        S.SchurStructure.from_description([3, 65536], [[1, 2]], [True, False])
    with pytest.raises(S.DimensionMismatch):
        S.SchurStructure.from_description([3, 6], [[1, 2]], [True])
    with pytest.raises(ValueError):
        S.SchurStructure([6], [3], [[1]], nchunks=0)


def _random_description(rng, ncam, nland, kmax, rr=0, mixed=False):
    sizes_c = rng.choice([6, 9, 3], size=ncam) if mixed else np.full(ncam, 6)
    sizes_l = rng.choice([3, 2, 1], size=nland) if mixed else np.full(nland, 3)
    # interleave caller numbering: cameras and landmarks shuffled
    perm = rng.permutation(ncam + nland)
    blocksizes = np.empty(ncam + nland, dtype=np.int64)
    eliminated = np.zeros(ncam + nland, bool)
    cam_id = perm[:ncam] + 1
    land_id = perm[ncam:] + 1
    blocksizes[cam_id - 1] = sizes_c
    blocksizes[land_id - 1] = sizes_l
    eliminated[land_id - 1] = True
    rowblocks = []
    for l in range(nland):
        k = rng.integers(1, kmax + 1)
        for c in rng.choice(ncam, size=min(k, ncam), replace=False):
            rowblocks.append([int(cam_id[c]), int(land_id[l])] if rng.random() < 0.5 else [int(land_id[l]), int(cam_id[c])])
    for _ in range(rr):
        a, b = rng.choice(ncam, size=2, replace=False)
        rowblocks.append([int(cam_id[a]), int(cam_id[b])])
    return blocksizes, rowblocks, eliminated


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_mirror_matches_loop_restatement(blsamd, seed):
    """Vectorised host mirror == loop restatement, bit for bit: order, couplings, offsets, pattern, r_to_s,
    workspace index, column chunks."""
    rng = np.random.default_rng(seed)
    S = blsamd.structure
    bs, rows, el = _random_description(rng, ncam=7, nland=30, kmax=4, rr=3, mixed=(seed % 2 == 1))
    nchunks = 1 + seed
    st = S.SchurStructure.from_description(bs, rows, el, nchunks=nchunks)
    d = ref.describe_elimination(bs.tolist(), rows, el.tolist())
    assert st.order.tolist() == d["order"]
    assert st.retained_pairs == d["retained_pairs"]
    r_dims = [int(bs[b - 1]) for b in d["order"][:d["nretained"]]]
    e_dims = [int(bs[b - 1]) for b in d["order"][d["nretained"]:]]
    rs = ref.schur_structure(r_dims, e_dims, d["couplings"], nchunks=nchunks)
    for k in ("boffsets_r", "boffsets_e", "eoffsets", "eliminated_ptr", "entry_retained", "coffsets"):
        assert getattr(st, k).tolist() == rs[k], k
    assert [[tuple(c) for c in g] for g in st.chunks] == rs["chunks"]
    pat = S.build_schur_pattern(st)
    rp = ref.build_schur_pattern(rs, d["retained_pairs"])
    for k in ("block_colptr", "block_rowval", "block_row_offsets", "col_heights", "S_colptr", "S_rowval", "diag_inds"):
        assert pat[k].tolist() == rp[k], k
    ws = S.SchurReductionWorkspaceIndex(st, nchunks)
    rw = ref.workspace_index(rs, nchunks)
    assert ws.retained_ptr.tolist() == rw["retained_ptr"]
    assert ws.retained_entry.tolist() == rw["retained_entry"]
    assert ws.entry_eliminated.tolist() == rw["entry_eliminated"]
    assert ws.col_chunks == rw["col_chunks"]
    Rs = S.hessian_retained_storage(st, pat)
    blocks, nnzv = ref.retained_storage(rp, r_dims)
    assert Rs.nnzvals == nnzv
    assert list(zip(Rs.brow.tolist(), Rs.bcol.tolist(), Rs.start.tolist())) == blocks
    assert S.build_r_to_s_map(Rs, pat, st.boffsets_r).tolist() == ref.build_r_to_s_map(blocks, nnzv, rp, r_dims, rs["boffsets_r"])
    cp, rv, nz = S.makesparseindices(Rs, True)
    cp2, rv2, nz2 = ref.makesparseindices(blocks, r_dims, True)
    assert cp.tolist() == cp2 and rv.tolist() == rv2 and nz.tolist() == nz2


def test_partition_columns_matches_loop(blsamd):
    rng = np.random.default_rng(7)
    for _ in range(50):
        n = int(rng.integers(1, 30))
        counts = rng.integers(0, 9, size=n)
        rp = np.concatenate([[1], 1 + np.cumsum(counts)])
        K = int(rng.integers(1, 12))
        assert blsamd.structure.partition_columns(rp, K) == ref.partition_columns(rp.tolist(), K)


def test_r_to_s_count_rule(blsamd):
    """count(>(0), r_to_s) == 4*21 + 36*#retained pairs (test/schur_problems.jl:284-285), with extra pairs."""
    S = blsamd.structure
    st = S.SchurStructure([6, 6, 6, 6], [3] * 5, [[1], [2], [3], [4], [1, 2]], retained_pairs=[(1, 3), (2, 4)])
    pat = S.build_schur_pattern(st)
    Rs = S.hessian_retained_storage(st, pat)
    rts = S.build_r_to_s_map(Rs, pat, st.boffsets_r)
    # R's storage has a block for every pattern pair: (1,2) via the shared landmark, (1,3), (2,4)
    assert int((rts > 0).sum()) == 4 * 21 + 36 * 3


def _bsm_from_literal(blsamd, L, rbs, cbs):
    S = blsamd.structure
    rows = np.array(L["pattern_rows"])
    coords = [(i + 1, j + 1) for i in range(rows.shape[0]) for j in range(rows.shape[1]) if rows[i, j]]
    bsm = S.SparseBlockLinearSystem.from_blocks(coords, rbs, cbs)
    data = np.zeros(bsm.nnzvals)
    for key, blk in L["blocks"].items():
        i, j = map(int, key.split(","))
        s0 = bsm.block_start(i, j)
        blk = np.array(blk, dtype=float)
        data[s0 - 1:s0 - 1 + blk.size] = blk.T.reshape(-1)     # column-major
    return bsm, data


def test_copyto_literal_out(blsamd):
    """sparse(b) through makesparseindices + copyto! (the oracle's gather) equals the literal `out`."""
    import scipy.sparse as sps
    L = LIT["sparse_block_out"]
    bsm, data = _bsm_from_literal(blsamd, L, L["rowblocksizes"], L["colblocksizes"])
    cp, rv, smap = blsamd.structure.sparse_with_map(bsm, symmetric=False)
    dest = np.zeros(len(smap))
    oracle.copyto(dest, data, smap)
    A = sps.csc_matrix((dest, rv - 1, cp - 1), shape=(bsm.m, bsm.n)).toarray()
    assert np.array_equal(A, np.array(L["expected"], dtype=float))


def test_copyto_literal_out2_symmetric(blsamd):
    import scipy.sparse as sps
    L = LIT["sparse_block_out2"]
    bsm, data = _bsm_from_literal(blsamd, L, L["blocksizes"], L["blocksizes"])
    assert bsm.nnzvals == L["nnz_bsm"]
    cp, rv, smap = blsamd.structure.sparse_with_map(bsm, symmetric=True)
    assert len(smap) == L["nnz_symmetrifysparse"]
    dest = np.zeros(len(smap))
    oracle.copyto(dest, data, smap)
    A = sps.csc_matrix((dest, rv - 1, cp - 1), shape=(bsm.m, bsm.n)).toarray()
    out2 = np.array(L["out2"], dtype=float)
    assert np.array_equal(A, np.maximum(out2, out2.T))
    # in-place update after modifying a block (test/SparseBlockLinearSystem.jl:226-229)
    s0 = bsm.block_start(2, 1)
    data[s0 - 1:s0 + 5] += 5
    oracle.copyto(dest, data, smap)
    B = sps.csc_matrix((dest, rv - 1, cp - 1), shape=(bsm.m, bsm.n)).toarray()
    assert np.array_equal(B, B.T) and B[2, 0] == 6 and B[0, 2] == 6


def test_reduced_symmetric_scatter_is_symmetric_full_pattern(blsamd):
    """The reduced system's symmetric scatter map reproduces S (upper blocks, diagonal blocks mirrored)."""
    import scipy.sparse as sps
    p = blsamd.generators.config2(scale=0.002)
    st = p.structure
    res = oracle.schur_reduce_problem(p)
    cp, rv, smap = blsamd.structure.reduced_symmetric_scatter(st, p.pattern)
    dest = np.zeros(len(smap))
    oracle.copyto(dest, res["S"], smap)
    A = sps.csc_matrix((dest, rv - 1, cp - 1), shape=(st.nr, st.nr)).toarray()
    assert np.array_equal(A, A.T)
    # the upper triangle read straight from the pattern
    U = np.zeros((st.nr, st.nr))
    for j in range(st.nr):
        for k in range(p.pattern.S_colptr[j] - 1, p.pattern.S_colptr[j + 1] - 1):
            i = p.pattern.S_rowval[k] - 1
            if i <= j:
                U[i, j] = res["S"][k]
    assert np.array_equal(np.triu(A), U)
