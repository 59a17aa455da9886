"""Shared test helpers: block-wise relative errors and problem builders (the reference's fixtures restated)."""
import numpy as np


def dense_upper_blocks(st, S_flat):
    """Upper blocks (r1 <= r2) of a dense column-major flat S as a dict (r1, r2) -> block (rows r1, cols r2)."""
    nr = st.nr
    S = S_flat.reshape(nr, nr).T
    out = {}
    for r2 in range(st.nrblocks):
        c0, d2 = st.boffsets_r[r2] - 1, st.r_dims[r2]
        for r1 in range(r2 + 1):
            o1, d1 = st.boffsets_r[r1] - 1, st.r_dims[r1]
            blk = S[o1:o1 + d1, c0:c0 + d2].copy()
            if r1 == r2:
                blk = np.triu(blk)
            out[(r1, r2)] = blk
    return out


def sparse_upper_blocks(st, pattern, nzval):
    out = {}
    for r2 in range(st.nrblocks):
        d2 = st.r_dims[r2]
        base = pattern.S_colptr[st.boffsets_r[r2] - 1] - 1
        h = pattern.col_heights[r2]
        for k in range(pattern.block_colptr[r2] - 1, pattern.block_colptr[r2 + 1] - 1):
            r1 = pattern.block_rowval[k] - 1
            d1 = st.r_dims[r1]
            ro = pattern.block_row_offsets[k]
            blk = np.empty((d1, d2))
            for c in range(d2):
                blk[:, c] = nzval[base + c * h + ro: base + c * h + ro + d1]
            if r1 == r2:
                blk = np.triu(blk)
            out[(r1, r2)] = blk
    return out


def max_block_rel(A, B):
    """max over blocks of ||A_b - B_b||_F / ||B_b||_F (absolute where B_b = 0)."""
    worst = 0.0
    for k, b in B.items():
        nb = np.linalg.norm(b)
        d = np.linalg.norm(A[k] - b)
        worst = max(worst, d / nb if nb > 0 else d)
    return worst


def schur_test_problem(rng, r_dims, e_dims, couplings, rr_pairs=()):
    """SchurTestProblem (test/fixtures/SchurFixtures.jl:26-54): a random strictly diagonally dominant bipartite H,
    retained first, and b. Returns (H, b)."""
    r_dims, e_dims = list(r_dims), list(e_dims)
    nr, ne = sum(r_dims), sum(e_dims)
    n = nr + ne
    roff = np.cumsum([0] + r_dims[:-1])
    eoff = nr + np.cumsum([0] + e_dims[:-1])
    H = np.zeros((n, n))

    def fill(r0, nrow, c0, ncol):
        H[r0:r0 + nrow, c0:c0 + ncol] = 2 * rng.random((nrow, ncol)) - 1

    for r, d in enumerate(r_dims):
        fill(roff[r], d, roff[r], d)
    for e, d in enumerate(e_dims):
        fill(eoff[e], d, eoff[e], d)
    for e, cpl in enumerate(couplings):
        for r in cpl:
            fill(roff[r - 1], r_dims[r - 1], eoff[e], e_dims[e])
    for (r1, r2) in rr_pairs:
        fill(roff[r1 - 1], r_dims[r1 - 1], roff[r2 - 1], r_dims[r2 - 1])
    H = np.triu(H) + np.triu(H, 1).T
    for i in range(n):
        H[i, i] = np.abs(H[i]).sum() - abs(H[i, i]) + 0.5 + rng.random()
    return H, rng.standard_normal(n)


class Problem:
    pass


def problem_from_dense(blsamd, H, b, r_dims, e_dims, couplings, lam, dense, nchunks=1, rr_pairs=(), scaling=None):
    """Lay a dense H out as a SchurBlockHessian (fill_ce!, test/fixtures/SchurFixtures.jl:59-79) and build the
    pieces the kernels take: structure, workspace index, pattern, R storage, r_to_s, E_data, C_data, b_r, b_e."""
    S_ = blsamd.structure
    st = S_.SchurStructure(r_dims, e_dims, couplings, nchunks=nchunks, retained_pairs=[tuple(p) for p in rr_pairs])
    nr = st.nr
    E_data = np.zeros(st.eliminated_storage_length())
    C_data = np.zeros(st.coupling_storage_length())
    for e in range(st.neblocks):
        de = st.e_dims[e]
        eo = st.eoffsets[e] - 1
        eo_h = nr + st.boffsets_e[e] - 1
        E_data[eo:eo + de * de] = H[eo_h:eo_h + de, eo_h:eo_h + de].T.reshape(-1)
        for c in range(st.eliminated_ptr[e] - 1, st.eliminated_ptr[e + 1] - 1):
            r = st.entry_retained[c] - 1
            dr = st.r_dims[r]
            ro = st.boffsets_r[r] - 1
            co = st.coffsets[c] - 1
            C_data[co:co + dr * de] = H[ro:ro + dr, eo_h:eo_h + de].reshape(-1)   # (d_e x d_r) column-major
    p = Problem()
    p.structure, p.lam, p.dense = st, lam, dense
    p.E_data, p.C_data = E_data, C_data
    p.b_r, p.b_e = b[:nr].copy(), b[nr:].copy()
    p.scaling_r = np.ones(nr) if scaling is None else scaling[:nr].copy()
    p.scaling_e = np.ones(st.ne) if scaling is None else scaling[nr:].copy()
    p.ws_index = S_.SchurReductionWorkspaceIndex(st, nchunks)
    p.H, p.b = H, b
    Rfull = H[:nr, :nr]
    if dense:
        p.R = np.triu(Rfull) + np.tril(np.zeros_like(Rfull))
        # the dense retained part stores the upper triangle of off-diagonal blocks and full diagonal blocks
        for r in range(st.nrblocks):
            o, d = st.boffsets_r[r] - 1, st.r_dims[r]
            p.R[o:o + d, o:o + d] = Rfull[o:o + d, o:o + d]
        p.pattern = None
    else:
        pat = S_.build_schur_pattern(st)
        Rs = S_.hessian_retained_storage(st, pat)
        Rdata = np.zeros(Rs.nnzvals)
        for brow, bcol, s0 in zip(Rs.brow, Rs.bcol, Rs.start):
            ro, rd_ = st.boffsets_r[brow - 1] - 1, st.r_dims[brow - 1]
            co, cd = st.boffsets_r[bcol - 1] - 1, st.r_dims[bcol - 1]
            Rdata[s0 - 1:s0 - 1 + rd_ * cd] = Rfull[ro:ro + rd_, co:co + cd].T.reshape(-1)
        p.pattern, p.Rsys, p.R_data = pat, Rs, Rdata
        p.r_to_s = S_.build_r_to_s_map(Rs, pat, st.boffsets_r)
    return p


def standard_problem_dims():
    """standard_problem (test/fixtures/SchurFixtures.jl:119-130)-like: retained 4 x 6, 20 eliminated of dims 3/2/1."""
    return [6, 6, 6, 6], [3] * 10 + [2] * 6 + [1] * 4


def reference_schur_problem():
    """The hand-built structure of test/schur.jl:12-66: r_dims [6,3,6,9], e_dims 8 x 3 + 2 x 1, couplings."""
    couplings = [[1, 2], [1, 3], [2, 3, 4], [1, 4], [1, 2, 3], [2, 4], [3, 4], [1, 2, 3, 4], [1, 3], [2, 4]]
    return [6, 3, 6, 9], [3] * 8 + [1, 1], couplings


def dense_magnitude_blocks(p):
    """Per upper block of a dense problem, the magnitude its entries are summed from: |R + damping| plus, over the
    landmarks, |C_1|' |(E + damping)^-1| |C_2| (entrywise). Rounding in any summation order is bounded by a small
    multiple of eps times this (the standard bound |fl(sum) - sum| <= gamma_n sum |terms|), which is what a
    relative error per block cannot promise where the terms cancel."""
    st = p.structure
    nr = st.nr
    M = np.abs(p.R) + p.lam * np.diag(np.asarray(p.scaling_r))
    for e in range(st.neblocks):
        de = st.e_dims[e]
        eo = st.eoffsets[e] - 1
        bo = st.boffsets_e[e] - 1
        E = p.E_data[eo:eo + de * de].reshape(de, de).T + p.lam * np.diag(p.scaling_e[bo:bo + de])
        cols, idx = [], []
        for c in range(st.eliminated_ptr[e] - 1, st.eliminated_ptr[e + 1] - 1):
            r = st.entry_retained[c] - 1
            dr = st.r_dims[r]
            co = st.coffsets[c] - 1
            cols.append(p.C_data[co:co + de * dr].reshape(dr, de).T)
            idx.extend(range(st.boffsets_r[r] - 1, st.boffsets_r[r] - 1 + dr))
        Ca = np.abs(np.hstack(cols))
        M[np.ix_(idx, idx)] += Ca.T @ np.abs(np.linalg.inv(E)) @ Ca
    return dense_upper_blocks(st, M.T.reshape(-1))


def max_block_rel_to(A, B, Mag):
    """max over blocks of ||A_b - B_b||_F / ||Mag_b||_F."""
    worst = 0.0
    for k, b in B.items():
        nb = np.linalg.norm(Mag[k])
        d = np.linalg.norm(A[k] - b)
        worst = max(worst, d / nb if nb > 0 else d)
    return worst


def assembly_problem(blsamd, seed, ncam=6, nland=40, kmax=4, npairs=4, sparse=False, tall=(), dup=0, triple=0):
    """A description with observations (2 rows over [landmark; camera], the landmark listed first or second at
    random), camera priors and camera-pair rows; its SchurStructure, the row blocks as layout block numbers in each
    Jacobian's column order, random J and r, and (sparse) the stored lower blocks of R: every retained pair that
    shares a row block plus the diagonal, as a block CSC with 1-based data offsets. tall: row counts of extra
    row blocks over cameras 1 and 2 (which then have 9 parameters each). dup: the first `dup` observations are
    made a second time (their coupling blocks then have two contributions); triple: landmarks 1..triple each get a
    2-row block over themselves and two cameras."""
    rng = np.random.default_rng(seed)
    cam = np.where(rng.random(ncam) < 0.5, 6, 9)
    if tall:
        cam[:2] = 9
    blocksizes = np.concatenate([cam, np.full(nland, 3)])
    eliminated = np.concatenate([np.zeros(ncam, bool), np.ones(nland, bool)])
    rows, sizes = [], []
    for l in range(nland):
        for c in rng.choice(ncam, size=int(rng.integers(2, kmax + 1)), replace=False):
            rb = [ncam + l + 1, int(c) + 1]
            rows.append(rb if rng.random() < 0.5 else rb[::-1]); sizes.append(2)
    for i in range(dup):
        rows.append(list(rows[i])); sizes.append(2)
    for l in range(triple):
        a, b = rng.choice(ncam, size=2, replace=False)
        rows.append([ncam + l + 1, int(a) + 1, int(b) + 1]); sizes.append(2)
    for c in range(ncam):
        rows.append([c + 1]); sizes.append(int(cam[c]))
    for _ in range(npairs):
        a, b = rng.choice(ncam, size=2, replace=False)
        rows.append([int(a) + 1, int(b) + 1]); sizes.append(3)
    for m in tall:
        rows.append([2, 1]); sizes.append(int(m))
    st = blsamd.structure.SchurStructure.from_description(blocksizes, [sorted(r) for r in rows], eliminated)
    layout = np.argsort(np.asarray(st.blockorder()) - 1) + 1     # invperm: layout number of each caller block
    rowblocks = [[int(layout[b - 1]) for b in r] for r in rows]
    dims = list(st.r_dims) + list(st.e_dims)
    W = [sum(int(dims[b - 1]) for b in r) for r in rowblocks]
    J = rng.standard_normal(int(sum(m * w for m, w in zip(sizes, W))))
    r = rng.standard_normal(int(sum(sizes)))
    R_layout = None
    if sparse:
        nrb = st.nrblocks
        pairs = {(q, q) for q in range(1, nrb + 1)}
        for rb in rowblocks:
            ret = [b for b in rb if b <= nrb]
            for a in ret:
                for b in ret:
                    if a >= b:
                        pairs.add((a, b))
        colptr, rowval, offs, off = [1], [], [], 1
        rd = st.r_dims
        for q in range(1, nrb + 1):
            for p_ in sorted(p for (p, qq) in pairs if qq == q):
                rowval.append(p_); offs.append(off); off += int(rd[p_ - 1] * rd[q - 1])
            colptr.append(len(rowval) + 1)
        R_layout = (np.array(colptr), np.array(rowval), np.array(offs))
    return st, rowblocks, np.array(sizes), J, r, R_layout
