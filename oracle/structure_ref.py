"""Loop-by-loop restatement of the reference's integer layout builders (TEST INFRASTRUCTURE ONLY).

Plain Python loops, written line by line from the Julia source, for small inputs: the checker of the vectorised
host mirror (blocklinearsystems.jl_amd/structure.py), which must produce the same maps bit for bit. Every value is
the reference's 1-based Int64. Each function cites the reference lines it follows.

Parity pinning: the known-answer literals of the reference's tests (tests/golden/reference_literals.json) and the
two restatements agreeing with each other on random descriptions (tests/test_structure.py).
"""


def sort_schur_blocks(is_eliminated, block_sizes, couplings):
    """src/schur_structure.jl:204-228 (Julia's sort! is stable; Python's sorted is stable)."""
    n = len(is_eliminated)
    r_indices = [i for i in range(1, n + 1) if not is_eliminated[i - 1]]
    r_indices = sorted(r_indices, key=lambda i: -int(block_sizes[i - 1]))
    e_indices = [i for i in range(1, n + 1) if is_eliminated[i - 1]]

    def key(i):
        coupled = couplings[i - 1]
        r_min = min(coupled) if len(coupled) else 2 ** 63 - 1
        return (-int(block_sizes[i - 1]), r_min)

    e_indices = sorted(e_indices, key=key)
    return r_indices + e_indices


def describe_elimination(blocksizes, rowblocks, eliminated):
    """src/schur_description.jl:16-89 (without the row-size checks of the Jacobian path)."""
    n = len(blocksizes)
    K = len(rowblocks)
    seen = [0] * (n + 1)
    rowblock_eliminated = [0] * K
    coupled = [[] for _ in range(n + 1)]
    for k in range(1, K + 1):
        e = 0
        for b in rowblocks[k - 1]:
            if not (1 <= b <= n):
                raise ValueError("column block out of range")
            if seen[b] == k:
                raise ValueError("repeated column block")
            seen[b] = k
            if eliminated[b - 1]:
                if e != 0:
                    raise ValueError("two eliminated blocks")
                e = b
        rowblock_eliminated[k - 1] = e
        if e != 0:
            for b in rowblocks[k - 1]:
                if not eliminated[b - 1]:
                    coupled[e].append(b)
    for b in range(1, n + 1):
        coupled[b] = sorted(set(coupled[b]))
    nret = sum(1 for x in eliminated if not x)
    order = sort_schur_blocks(eliminated, blocksizes, [coupled[b] for b in range(1, n + 1)])
    position = [0] * (n + 1)
    for p, b in enumerate(order, start=1):
        position[b] = p
    couplings = [sorted(position[r] for r in coupled[order[p - 1]]) for p in range(nret + 1, n + 1)]
    pairs = []
    for k in range(1, K + 1):
        rs = [position[b] for b in rowblocks[k - 1] if not eliminated[b - 1]]
        for a in range(2, len(rs) + 1):
            for c in range(1, a):
                pairs.append((min(rs[a - 1], rs[c - 1]), max(rs[a - 1], rs[c - 1])))
    pairs = sorted(set(pairs))
    return dict(nretained=nret, order=order, position=position, rowblock_eliminated=rowblock_eliminated,
                couplings=couplings, retained_pairs=pairs)


def schur_structure(r_dims, e_dims, couplings, nchunks=1):
    """Positional SchurStructure (src/schur_structure.jl:70-166)."""
    nrblocks, neblocks = len(r_dims), len(e_dims)
    boffsets_r, acc = [], 1
    for d in r_dims:
        boffsets_r.append(acc)
        acc += d
    nr = acc - 1
    boffsets_e, acc = [], 1
    for d in e_dims:
        boffsets_e.append(acc)
        acc += d
    eoffsets = []
    if neblocks > 0:
        eoffsets.append(1)
        for i in range(2, neblocks + 1):
            eoffsets.append(eoffsets[-1] + e_dims[i - 2] ** 2)
    eliminated_ptr = [1]
    for e in range(1, neblocks + 1):
        eliminated_ptr.append(eliminated_ptr[e - 1] + len(couplings[e - 1]))
    entry_retained, coffsets = [], []
    curr = 1
    for e in range(1, neblocks + 1):
        de = e_dims[e - 1]
        for r in sorted(couplings[e - 1]):
            entry_retained.append(r)
            coffsets.append(curr)
            curr += r_dims[r - 1] * de
    # runs of equal e_dims, then chunks (cld)
    runs, e = [], 1
    while e <= neblocks:
        f = e
        while e <= neblocks and e_dims[e - 1] == e_dims[f - 1]:
            e += 1
        runs.append((f, e - 1))
    chunks = []
    for (start, stop) in runs:
        ln = stop - start + 1
        csz = -(-ln // nchunks)
        grp = []
        for k in range(1, nchunks + 1):
            first = start + (k - 1) * csz
            last = min(stop, first + csz - 1)
            grp.append((first, last) if first <= last else (first, first - 1))
        chunks.append(grp)
    return dict(nr=nr, nrblocks=nrblocks, neblocks=neblocks, boffsets_r=boffsets_r, boffsets_e=boffsets_e,
                eoffsets=eoffsets, eliminated_ptr=eliminated_ptr, entry_retained=entry_retained, coffsets=coffsets,
                chunks=chunks, r_dims=list(r_dims), e_dims=list(e_dims))


def build_schur_pattern(st, retained_pairs=(), extra_blocks=()):
    """src/schur_structure.jl:241-323 (the `sparse(I, J, true, n, n, |)` is a sorted set of (row, col))."""
    nrb = st["nrblocks"]
    pairs = set()
    ep, er = st["eliminated_ptr"], st["entry_retained"]
    for e in range(1, st["neblocks"] + 1):
        c_start, c_end = ep[e - 1], ep[e] - 1
        for idx2 in range(c_start, c_end + 1):
            r2 = er[idx2 - 1]
            for idx1 in range(c_start, idx2 + 1):
                pairs.add((er[idx1 - 1], r2))
    for (i, j) in retained_pairs:
        pairs.add((i, j))
    for (i, j) in extra_blocks:
        pairs.add((min(i, j), max(i, j)))
    for r in range(1, nrb + 1):
        pairs.add((r, r))
    bycol = sorted(pairs, key=lambda p: (p[1], p[0]))
    block_colptr = [1]
    block_rowval = []
    for j in range(1, nrb + 1):
        rows = [i for (i, jj) in bycol if jj == j]
        block_rowval.extend(rows)
        block_colptr.append(block_colptr[-1] + len(rows))
    col_heights = [0] * nrb
    block_row_offsets = [0] * len(block_rowval)
    for r2 in range(1, nrb + 1):
        h = 0
        for idx in range(block_colptr[r2 - 1], block_colptr[r2]):
            r1 = block_rowval[idx - 1]
            block_row_offsets[idx - 1] = h
            h += st["r_dims"][r1 - 1]
        col_heights[r2 - 1] = h
    # scalar pattern
    entries = []
    bo = st["boffsets_r"]
    for r2 in range(1, nrb + 1):
        for idx in range(block_colptr[r2 - 1], block_colptr[r2]):
            r1 = block_rowval[idx - 1]
            for c in range(1, st["r_dims"][r2 - 1] + 1):
                for rr in range(1, st["r_dims"][r1 - 1] + 1):
                    entries.append((bo[r1 - 1] + rr - 1, bo[r2 - 1] + c - 1))
    entries = sorted(set(entries), key=lambda p: (p[1], p[0]))
    nr = st["nr"]
    S_colptr, S_rowval = [1], []
    k = 0
    for j in range(1, nr + 1):
        while k < len(entries) and entries[k][1] == j:
            S_rowval.append(entries[k][0])
            k += 1
        S_colptr.append(len(S_rowval) + 1)
    diag_inds = []
    for col in range(1, nr + 1):
        for idx in range(S_colptr[col - 1], S_colptr[col]):
            if S_rowval[idx - 1] == col:
                diag_inds.append(idx)
                break
    return dict(block_colptr=block_colptr, block_rowval=block_rowval, block_row_offsets=block_row_offsets,
                col_heights=col_heights, S_colptr=S_colptr, S_rowval=S_rowval, diag_inds=diag_inds)


def workspace_index(st, nchunks):
    """SchurReductionWorkspace's index (src/schur_reduction.jl:71-89) and partition_columns (:30-49)."""
    ncoup = len(st["entry_retained"])
    entry_eliminated = [0] * ncoup
    counts = [0] * (st["nrblocks"] + 1)
    ep = st["eliminated_ptr"]
    for e in range(1, st["neblocks"] + 1):
        for c in range(ep[e - 1], ep[e]):
            entry_eliminated[c - 1] = e
            counts[st["entry_retained"][c - 1]] += 1
    retained_ptr = [1]
    for r in range(1, st["nrblocks"] + 1):
        retained_ptr.append(retained_ptr[-1] + counts[r])
    retained_entry = [0] * ncoup
    fill_pos = list(retained_ptr)
    for c in range(1, ncoup + 1):
        r = st["entry_retained"][c - 1]
        retained_entry[fill_pos[r - 1] - 1] = c
        fill_pos[r - 1] += 1
    return dict(entry_eliminated=entry_eliminated, retained_ptr=retained_ptr, retained_entry=retained_entry,
                col_chunks=partition_columns(retained_ptr, max(nchunks, 1)))


def partition_columns(retained_ptr, K):
    """src/schur_reduction.jl:30-49."""
    n = len(retained_ptr) - 1
    total = retained_ptr[-1] - retained_ptr[0] + n
    ranges, start = [], 1
    for k in range(1, K + 1):
        stop = start - 1
        if k == K:
            stop = n
        else:
            target = (total * k) / K
            while stop < n and (retained_ptr[stop + 1] - retained_ptr[0] + (stop + 1)) <= target:
                stop += 1
        ranges.append((start, stop))
        start = stop + 1
    return ranges


def retained_storage(pattern, r_dims):
    """SparseBlockLinearSystem{T}(TransposedPattern(block_pattern), r_dims) (src/SparseBlockLinearSystem.jl:43-68):
    stored blocks (brow, bcol) for pattern entries (bcol, brow), data starts in (brow, bcol ascending) order."""
    blocks, start = [], 1
    nrb = len(r_dims)
    for row in range(1, nrb + 1):
        for idx in range(pattern["block_colptr"][row - 1], pattern["block_colptr"][row]):
            col = pattern["block_rowval"][idx - 1]
            blocks.append((row, col, start))
            start += r_dims[row - 1] * r_dims[col - 1]
    return blocks, start - 1


def build_r_to_s_map(blocks, nnzvals, pattern, r_dims, boffsets_r):
    """src/schur_structure.jl:406-431."""
    r_to_s = [0] * nnzvals
    S_colptr = pattern["S_colptr"]
    for (brow, bcol, data_start) in blocks:
        if brow < bcol:
            continue
        drow, dcol = r_dims[brow - 1], r_dims[bcol - 1]
        lo, hi = pattern["block_colptr"][brow - 1], pattern["block_colptr"][brow] - 1
        k = None
        for idx in range(lo, hi + 1):
            if pattern["block_rowval"][idx - 1] == bcol:
                k = idx
        if k is None:
            raise ValueError("block missing")
        row_offset = pattern["block_row_offsets"][k - 1]
        base = S_colptr[boffsets_r[brow - 1] - 1] + row_offset
        stride = pattern["col_heights"][brow - 1]
        if brow == bcol:
            for col in range(1, dcol + 1):
                for row in range(1, col + 1):
                    r_to_s[data_start + (col - 1) * drow + row - 2] = base + (col - 1) * stride + row - 1
        else:
            for col in range(1, dcol + 1):
                for row in range(1, drow + 1):
                    r_to_s[data_start + (col - 1) * drow + row - 2] = base + (row - 1) * stride + col - 1
    return r_to_s


def makesparseindices(lower_blocks, rbs, symmetrify):
    """src/SparseBlockLinearSystem.jl:233-282 for a block-square storage given as (brow, bcol, start) blocks."""
    nb = len(rbs)
    startrow = [1]
    for d in rbs:
        startrow.append(startrow[-1] + d)
    by_col = {c: sorted((r, s) for (r, cc, s) in lower_blocks if cc == c) for c in range(1, nb + 1)}      # indices
    by_row = {r: sorted((c, s) for (rr, c, s) in lower_blocks if rr == r) for r in range(1, nb + 1)}      # transposed
    rows, idx, colptr = [], [], [1]
    for c in range(1, nb + 1):
        cbs = rbs[c - 1]
        lower = by_col[c]
        upper = by_row[c] if symmetrify else []
        if symmetrify and lower and lower[0][0] == c:
            upper = upper[:-1]
        for inner in range(cbs):
            for (row, s0) in upper:
                v = s0 + inner
                for i in range(rbs[row - 1]):
                    rows.append(startrow[row - 1] + i)
                    idx.append(v)
                    v += cbs
            for (row, s0) in lower:
                cc = rbs[row - 1]
                v = s0 + inner * cc
                for i in range(cc):
                    rows.append(startrow[row - 1] + i)
                    idx.append(v + i)
            colptr.append(len(rows) + 1)
    return colptr, rows, idx


def schur_jacobian_structure(st, e_rows, free_rows=(), free_retained=(), nchunks=None):
    """Positional SchurJacobianStructure (src/schur_jacobian.jl:59-160)."""
    ne, nrb = st["neblocks"], st["nrblocks"]
    ntables = ne + len(free_rows)
    sorted_free = [sorted(rs) for rs in free_retained]
    table_nrows, table_ncols, table_top, table_ptr = [], [], [], [1]
    entry_retained, entry_col, entry_table = [], [], []
    ep = st["eliminated_ptr"]
    for t in range(1, ntables + 1):
        rs = st["entry_retained"][ep[t - 1] - 1: ep[t] - 1] if t <= ne else sorted_free[t - ne - 1]
        col = 1
        for r in rs:
            entry_retained.append(r)
            entry_col.append(col)
            entry_table.append(t)
            col += st["r_dims"][r - 1]
        table_ptr.append(len(entry_retained) + 1)
        de = st["e_dims"][t - 1] if t <= ne else 0
        table_top.append(de)
        table_nrows.append(e_rows[t - 1] + de if t <= ne else free_rows[t - ne - 1])
        table_ncols.append(col - 1 + de + 1)
    raw_offsets, residual_offsets, reduced_offsets = [1], [1], [1]
    for t in range(ntables):
        raw_offsets.append(raw_offsets[-1] + table_nrows[t] * (table_ncols[t] - 1))
        residual_offsets.append(residual_offsets[-1] + table_nrows[t])
        reduced_offsets.append(reduced_offsets[-1] + table_nrows[t] - table_top[t])
    counts = [0] * (nrb + 1)
    for r in entry_retained:
        counts[r] += 1
    retained_ptr = [1]
    for r in range(1, nrb + 1):
        retained_ptr.append(retained_ptr[-1] + counts[r])
    retained_entry = [0] * len(entry_retained)
    fill = list(retained_ptr)
    for u, r in enumerate(entry_retained, start=1):
        retained_entry[fill[r - 1] - 1] = u
        fill[r - 1] += 1
    doffsets = [1]
    for d in st["r_dims"]:
        doffsets.append(doffsets[-1] + d * d)
    if nchunks is None:
        nchunks = max([len(c) for c in st["chunks"]] + [1])
    return dict(ntables=ntables, table_nrows=table_nrows, table_ncols=table_ncols, table_top=table_top,
                table_ptr=table_ptr, entry_retained=entry_retained, entry_col=entry_col, entry_table=entry_table,
                raw_offsets=raw_offsets, residual_offsets=residual_offsets, reduced_offsets=reduced_offsets,
                retained_ptr=retained_ptr, retained_entry=retained_entry, doffsets=doffsets,
                col_chunks=partition_columns(retained_ptr, nchunks))
