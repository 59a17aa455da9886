"""Layout of an elimination: the integer maps the device kernels read.

Host-side mirror of the reference's layout types, with the reference's field names and its conventions kept
verbatim: every index array holds **1-based** Int64 values, exactly what the Julia structures store, so the
same arrays cross the C ABI whether Julia or Python built them. Built with vectorised numpy (setup, not timed).

Reference: src/schur_description.jl (describe_elimination, public constructors), src/schur_structure.jl
(positional SchurStructure, sort_schur_blocks, build_schur_pattern, build_r_to_s_map), src/schur_reduction.jl
(SchurReductionWorkspace index, partition_columns), src/SparseBlockLinearSystem.jl (block storage,
makesparseindices, sparse_with_map), src/schur_jacobian.jl (SchurJacobianStructure).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

MAX_BLOCK_DIM = 65535          # BlockSize = UInt16 (src/BlockRLE.jl:33-35)
DEFAULT_MIN_WORK = 4096        # src/threading.jl:24
I64 = np.int64


class DimensionMismatch(ValueError):
    """Julia's DimensionMismatch."""


def _csr(lists):
    """(ptr, flat) of a list of integer lists; ptr 0-based with len(lists)+1 entries."""
    lens = np.fromiter((len(x) for x in lists), dtype=I64, count=len(lists))
    ptr = np.zeros(len(lists) + 1, dtype=I64)
    np.cumsum(lens, out=ptr[1:])
    flat = np.fromiter((int(v) for x in lists for v in x), dtype=I64, count=int(ptr[-1]))
    return ptr, flat


def _as_csr(rows):
    """Accept a list of lists or a (ptr0, flat) pair with 0-based ptr; returns (ptr0, flat) int64."""
    if isinstance(rows, tuple) and len(rows) == 2 and isinstance(rows[0], np.ndarray):
        return np.asarray(rows[0], dtype=I64), np.asarray(rows[1], dtype=I64)
    return _csr(rows)


def _check_dims(d, what="Block"):
    d = np.asarray(d, dtype=I64)
    bad = np.nonzero((d < 1) | (d > MAX_BLOCK_DIM))[0]
    if bad.size:
        i = int(bad[0])
        raise ValueError(f"{what} {i + 1} has dimension {int(d[i])}; dimensions must be between 1 and {MAX_BLOCK_DIM}")
    return d


def startindices(sizes):
    """computestartindices (src/utils.jl:297-305): 1-based start of each block."""
    sizes = np.asarray(sizes, dtype=I64)
    out = np.ones(len(sizes), dtype=I64)
    if len(sizes) > 1:
        out[1:] = 1 + np.cumsum(sizes[:-1])
    return out


def runs(dims):
    """Runs of equal block size (BlockRLE, src/BlockRLE.jl:49-67): list of (first, last, size), 1-based."""
    dims = np.asarray(dims, dtype=I64)
    if len(dims) == 0:
        return []
    cut = np.nonzero(np.diff(dims))[0] + 1
    starts = np.concatenate([[0], cut])
    stops = np.concatenate([cut, [len(dims)]])
    return [(int(a) + 1, int(b), int(dims[a])) for a, b in zip(starts, stops)]


def partition_columns(retained_ptr, K):
    """partition_columns (src/schur_reduction.jl:30-49): K contiguous ranges of retained blocks of roughly equal
    work (couplings plus the diagonal); Float64 target as in the reference. Returns a list of (lo, hi), 1-based,
    hi = lo - 1 for an empty range."""
    rp = np.asarray(retained_ptr, dtype=I64)
    n = len(rp) - 1
    total = int(rp[-1] - rp[0]) + n
    # work after block `stop` (1-based) = rp[stop+1] - rp[1] + stop  (the reference's retained_ptr[stop+2]
    # - retained_ptr[1] + stop + 1 for the next candidate)
    cum = (rp[1:] - rp[0]) + np.arange(1, n + 1, dtype=I64)   # cum[s-1]: work of blocks 1..s
    ranges = []
    start = 1
    for k in range(1, K + 1):
        stop = start - 1
        if k == K:
            stop = n
        else:
            target = (total * k) / K
            # advance while stop < n and cum[stop] <= target (cum is nondecreasing)
            hi = int(np.searchsorted(cum, target, side="right"))   # number of s with cum[s-1] <= target
            stop = max(stop, min(hi, n))
        ranges.append((start, stop))
        start = stop + 1
    return ranges


class SchurStructure:
    """SchurStructure (src/schur_structure.jl:50-166); fields named and valued as in the reference.

    Construct positionally, ``SchurStructure(r_dims, e_dims, couplings, nchunks=1, min_work=4096, order=None,
    retained_pairs=())`` (couplings: per eliminated block, the layout numbers of its retained blocks), or from a
    description with :meth:`from_description` (the public constructor, src/schur_description.jl:97-101).
    """

    def __init__(self, r_dims, e_dims, couplings, nchunks=1, min_work=DEFAULT_MIN_WORK, order=None,
                 retained_pairs=()):
        r_dims = np.asarray(r_dims, dtype=I64)
        e_dims = np.asarray(e_dims, dtype=I64)
        nrb, neb = len(r_dims), len(e_dims)
        cptr, cflat = _as_csr(couplings)
        if len(cptr) - 1 != neb:
            raise DimensionMismatch(f"Number of coupling lists ({len(cptr) - 1}) must match the number of eliminated blocks ({neb})")
        if nchunks < 1:
            raise ValueError(f"nchunks must be at least 1, got {nchunks}")
        if min_work < 0:
            raise ValueError(f"min_work must be non-negative, got {min_work}")
        _check_dims(r_dims)
        _check_dims(e_dims)
        owner = np.repeat(np.arange(neb, dtype=I64), np.diff(cptr))
        if cflat.size:
            bad = np.nonzero((cflat < 1) | (cflat > nrb))[0]
            if bad.size:
                i = int(bad[0])
                raise ValueError(f"Eliminated block {int(owner[i]) + 1} references retained block {int(cflat[i])}, but there are only {nrb} retained blocks")
        # sort each coupling list (stable per eliminated block) and reject repeats
        perm = np.lexsort((cflat, owner))
        entry_retained = cflat[perm]
        if entry_retained.size > 1:
            dup = np.nonzero((np.diff(entry_retained) == 0) & (np.diff(owner) == 0))[0]
            if dup.size:
                i = int(dup[0])
                raise ValueError(f"Retained block {int(entry_retained[i])} is listed more than once in the couplings of eliminated block {int(owner[i]) + 1}")

        self.nr = int(r_dims.sum())
        self.nrblocks, self.neblocks = nrb, neb
        self.r_dims, self.e_dims = r_dims, e_dims
        self.boffsets_r = startindices(r_dims)
        self.boffsets_e = startindices(e_dims)
        self.eoffsets = startindices(e_dims * e_dims)
        self.eliminated_ptr = cptr + 1
        self.entry_retained = entry_retained.astype(I64)
        csz = r_dims[entry_retained - 1] * e_dims[owner] if entry_retained.size else np.zeros(0, I64)
        self.coffsets = startindices(csz) if csz.size else np.zeros(0, I64)
        self.clength = int(csz.sum())
        obs = np.zeros(neb, dtype=I64)
        if entry_retained.size:
            np.add.at(obs, owner, r_dims[entry_retained - 1])
        self.max_coupled_dim = int(obs.max()) if neb else 0
        self.chunks = []
        for (a, b, _d) in runs(e_dims):
            ln = b - a + 1
            csz_ = -(-ln // nchunks)
            grp = []
            for k in range(1, nchunks + 1):
                first = a + (k - 1) * csz_
                last = min(b, first + csz_ - 1)
                grp.append((first, last) if first <= last else (first, first - 1))
            self.chunks.append(grp)
        ntotal = nrb + neb
        self.order = np.arange(1, ntotal + 1, dtype=I64) if order is None else np.asarray(order, dtype=I64)
        if len(self.order) != ntotal:
            raise DimensionMismatch(f"order has length {len(self.order)}, but there are {ntotal} blocks")
        pairs = sorted({(int(i), int(j)) for (i, j) in retained_pairs})
        for (i, j) in pairs:
            if not (1 <= i < j <= nrb):
                raise ValueError(f"Retained pair ({i}, {j}) must satisfy 1 <= i < j <= {nrb}")
        self.retained_pairs = pairs
        self.nchunks, self.min_work = int(nchunks), int(min_work)

    # ---- accessors (src/schur_description.jl:199-214)
    def blockorder(self):
        return self.order.copy()

    def nretained(self):
        return self.nrblocks

    def nblocks(self):
        return self.nrblocks + self.neblocks

    def blocksizes(self):
        return np.concatenate([self.r_dims, self.e_dims])

    @property
    def ncoup(self):
        return len(self.entry_retained)

    @property
    def ne(self):
        return int(self.e_dims.sum())

    def eliminated_storage_length(self):
        return int((self.e_dims ** 2).sum())

    def coupling_storage_length(self):
        return self.clength

    @classmethod
    def from_description(cls, blocksizes, rowblocks, eliminated, nchunks=1, min_work=DEFAULT_MIN_WORK):
        """SchurStructure(blocksizes, rowblocks; eliminated) (src/schur_description.jl:97-101)."""
        d = describe_elimination(blocksizes, rowblocks, eliminated, None)
        return description_structure(d, blocksizes, nchunks, min_work)


def sort_schur_blocks(is_eliminated, block_sizes, couplings):
    """sort_schur_blocks (src/schur_structure.jl:204-228): retained by size descending, eliminated by
    (size descending, smallest coupled block); Julia's sort! is stable, so ties keep caller order. 1-based."""
    el = np.asarray(is_eliminated, dtype=bool)
    bs = np.asarray(block_sizes, dtype=I64)
    n = len(el)
    if len(bs) != n:
        raise DimensionMismatch(f"block_sizes has length {len(bs)}, but is_eliminated has length {n}")
    cptr, cflat = _as_csr(couplings)
    if len(cptr) - 1 != n:
        raise DimensionMismatch(f"couplings has length {len(cptr) - 1}, but is_eliminated has length {n}")
    r_idx = np.nonzero(~el)[0]
    r_idx = r_idx[np.argsort(-bs[r_idx], kind="stable")]
    e_idx = np.nonzero(el)[0]
    big = np.iinfo(np.int64).max
    rmin = np.full(n, big, dtype=I64)
    if cflat.size:
        owner = np.repeat(np.arange(n, dtype=I64), np.diff(cptr))
        np.minimum.at(rmin, owner, cflat)
    e_idx = e_idx[np.lexsort((rmin[e_idx], -bs[e_idx]))]
    return np.concatenate([r_idx, e_idx]).astype(I64) + 1


def expand_block_perm(block_perm, block_sizes):
    """expand_block_perm (src/schur_structure.jl:173-194)."""
    bp = np.asarray(block_perm, dtype=I64)
    bs = np.asarray(block_sizes, dtype=I64)
    if len(bp) != len(bs):
        raise DimensionMismatch(f"block_perm has length {len(bp)}, but there are {len(bs)} block sizes")
    if sorted(bp.tolist()) != list(range(1, len(bs) + 1)):
        raise ValueError(f"block_perm must be a permutation of 1:{len(bs)}")
    bo = startindices(bs)
    return np.concatenate([np.arange(bo[b - 1], bo[b - 1] + bs[b - 1], dtype=I64) for b in bp]) if len(bp) else np.zeros(0, I64)


class EliminationDescription:
    def __init__(self, nretained, order, position, rowblock_eliminated, couplings, retained_pairs):
        self.nretained = nretained
        self.order = order
        self.position = position
        self.rowblock_eliminated = rowblock_eliminated
        self.couplings = couplings          # (ptr0, flat) per eliminated layout block, ascending
        self.retained_pairs = retained_pairs


def describe_elimination(blocksizes, rowblocks, eliminated, rowblocksizes):
    """describe_elimination (src/schur_description.jl:16-89), vectorised. Row blocks as lists or (ptr0, flat)."""
    bs = np.asarray(blocksizes, dtype=I64)
    el = np.asarray(eliminated, dtype=bool)
    n = len(bs)
    rptr, rflat = _as_csr(rowblocks)
    K = len(rptr) - 1
    if len(el) != n:
        raise DimensionMismatch(f"eliminated has length {len(el)}, but there are {n} column blocks")
    if rowblocksizes is not None:
        rowblocksizes = np.asarray(rowblocksizes, dtype=I64)
        if len(rowblocksizes) != K:
            raise DimensionMismatch(f"rowblocksizes has length {len(rowblocksizes)}, but there are {K} row blocks")
    _check_dims(bs)
    nret = int((~el).sum())
    if not (0 < nret < n):
        raise ValueError(f"An elimination needs at least one retained and one eliminated column block; got {nret} retained and {n - nret} eliminated")
    rowid = np.repeat(np.arange(K, dtype=I64), np.diff(rptr))
    if rowblocksizes is not None:
        bad = np.nonzero(rowblocksizes < 1)[0]
        if bad.size:
            k = int(bad[0])
            raise ValueError(f"Row block {k + 1} has {int(rowblocksizes[k])} rows; every row block needs at least 1 row")
    if rflat.size:
        bad = np.nonzero((rflat < 1) | (rflat > n))[0]
        if bad.size:
            i = int(bad[0])
            raise ValueError(f"Row block {int(rowid[i]) + 1} touches column block {int(rflat[i])}, but there are {n} column blocks")
        o = np.lexsort((rflat, rowid))
        dup = np.nonzero((np.diff(rflat[o]) == 0) & (np.diff(rowid[o]) == 0))[0]
        if dup.size:
            i = o[int(dup[0])]
            raise ValueError(f"Row block {int(rowid[i]) + 1} lists column block {int(rflat[i])} more than once")
    iselim = el[rflat - 1] if rflat.size else np.zeros(0, bool)
    nel = np.bincount(rowid[iselim], minlength=K) if rflat.size else np.zeros(K, I64)
    if (nel > 1).any():
        k = int(np.nonzero(nel > 1)[0][0])
        es = rflat[rptr[k]:rptr[k + 1]][el[rflat[rptr[k]:rptr[k + 1]] - 1]]
        raise ValueError(f"Row block {k + 1} touches two eliminated column blocks, {int(es[0])} and {int(es[1])}; a row block may touch at most one")
    rb_elim = np.zeros(K, dtype=I64)
    rb_elim[rowid[iselim]] = rflat[iselim]
    if rowblocksizes is not None:
        hasrow = np.zeros(n, bool)
        hasrow[rflat - 1] = True
        miss = np.nonzero(el & ~hasrow)[0]
        if miss.size:
            raise ValueError(f"Eliminated column block {int(miss[0]) + 1} is in no row block; eliminating a block needs at least one row")
    # couplings in caller numbering: retained entries of row blocks whose eliminated block is e
    sel = (~iselim) & (rb_elim[rowid] != 0) if rflat.size else np.zeros(0, bool)
    ce = rb_elim[rowid[sel]]
    cr = rflat[sel]
    key = np.unique(ce * (n + 1) + cr)
    ce_u, cr_u = key // (n + 1), key % (n + 1)
    cnt = np.bincount(ce_u - 1, minlength=n)
    cptr = np.zeros(n + 1, dtype=I64)
    np.cumsum(cnt, out=cptr[1:])
    order = sort_schur_blocks(el, bs, (cptr, cr_u))
    position = np.empty(n, dtype=I64)
    position[order - 1] = np.arange(1, n + 1, dtype=I64)
    # couplings per eliminated layout block, in layout numbering, sorted
    elay = order[nret:]                                            # caller numbers of eliminated layout blocks
    lens = cnt[elay - 1]
    lptr = np.zeros(len(elay) + 1, dtype=I64)
    np.cumsum(lens, out=lptr[1:])
    src = np.concatenate([np.arange(cptr[b - 1], cptr[b], dtype=I64) for b in elay]) if len(elay) else np.zeros(0, I64)
    vals = position[cr_u[src] - 1] if src.size else np.zeros(0, I64)
    own = np.repeat(np.arange(len(elay), dtype=I64), lens)
    vals = vals[np.lexsort((vals, own))]
    # pairs of retained blocks sharing a row block, layout numbering
    pairs = set()
    if rflat.size:
        ret = ~iselim
        rid = rowid[ret]
        rpos = position[rflat[ret] - 1]
        m = np.bincount(rid, minlength=K)
        if (m > 1).any():
            rp = np.zeros(K + 1, dtype=I64)
            np.cumsum(m, out=rp[1:])
            maxm = int(m.max())
            pl = []
            for a in range(1, maxm):
                rows = np.nonzero(m > a)[0]
                for c in range(a):
                    x = rpos[rp[rows] + a]
                    y = rpos[rp[rows] + c]
                    pl.append(np.stack([np.minimum(x, y), np.maximum(x, y)], 1))
            allp = np.unique(np.concatenate(pl), axis=0)
            pairs = [(int(i), int(j)) for i, j in allp]
    return EliminationDescription(nret, order, position, rb_elim, (lptr, vals), sorted(pairs) if pairs else [])


def description_structure(d, blocksizes, nchunks, min_work):
    """description_structure (src/schur_description.jl:91-95)."""
    bs = np.asarray(blocksizes, dtype=I64)
    r_dims = bs[d.order[:d.nretained] - 1]
    e_dims = bs[d.order[d.nretained:] - 1]
    return SchurStructure(r_dims, e_dims, d.couplings, nchunks=nchunks, min_work=min_work, order=d.order,
                          retained_pairs=d.retained_pairs)


class SchurPattern(dict):
    """build_schur_pattern's named tuple (integer part): attribute access to the dict entries."""
    __getattr__ = dict.__getitem__


def _pairs_per_eliminated(st):
    """All (r1, r2), r1 <= r2, of retained blocks coupled to one eliminated block (with the diagonal)."""
    ptr = st.eliminated_ptr - 1
    k = np.diff(ptr)
    I, J = [], []
    for kk in np.unique(k):
        kk = int(kk)
        if kk == 0:
            continue
        es = np.nonzero(k == kk)[0]
        mat = st.entry_retained[ptr[es][:, None] + np.arange(kk)[None, :]]
        a, b = np.triu_indices(kk)
        I.append(mat[:, a].ravel())
        J.append(mat[:, b].ravel())
    if not I:
        return np.zeros(0, I64), np.zeros(0, I64)
    return np.concatenate(I), np.concatenate(J)


def build_schur_pattern(st, extra_blocks=()):
    """build_schur_pattern (src/schur_structure.jl:241-323): the upper block pattern of the reduced system
    (pairs sharing an eliminated block, retained pairs, extras, the diagonal) and its scalar CSC with full
    blocks. All indices 1-based."""
    nrb = st.nrblocks
    # co-occurrence through eliminated blocks: B'B of the eliminated x retained incidence
    owner = np.repeat(np.arange(st.neblocks, dtype=I64), np.diff(st.eliminated_ptr))
    B = sp.csr_matrix((np.ones(len(owner), dtype=np.int32), (owner, st.entry_retained - 1)), shape=(st.neblocks, nrb))
    P = sp.triu((B.T @ B).tocoo())
    I = [P.row.astype(I64), np.arange(nrb, dtype=I64)]
    J = [P.col.astype(I64), np.arange(nrb, dtype=I64)]
    if st.retained_pairs:
        rp = np.asarray(st.retained_pairs, dtype=I64) - 1
        I.append(rp[:, 0]); J.append(rp[:, 1])
    if len(extra_blocks):
        eb = np.asarray(list(extra_blocks), dtype=I64).reshape(-1, 2)
        bad = (eb < 1) | (eb > nrb)
        if bad.any():
            i, j = eb[np.nonzero(bad.any(1))[0][0]]
            raise ValueError(f"Extra block pair ({i}, {j}) is outside the {nrb} retained blocks")
        I.append(np.minimum(eb[:, 0], eb[:, 1]) - 1); J.append(np.maximum(eb[:, 0], eb[:, 1]) - 1)
    I = np.concatenate(I); J = np.concatenate(J)
    key = np.unique(J * nrb + I)            # CSC order: column, then row
    bcol, brow = key // nrb, key % nrb
    block_colptr = np.zeros(nrb + 1, dtype=I64)
    np.cumsum(np.bincount(bcol, minlength=nrb), out=block_colptr[1:])
    block_rowval = brow + 1
    rd = st.r_dims
    h = rd[brow]
    csum = np.cumsum(h)
    colstart = block_colptr[:-1]
    col_total = np.zeros(nrb, dtype=I64)
    np.add.at(col_total, bcol, h)
    excl = csum - h
    block_row_offsets = excl - (excl[colstart[bcol]] if len(bcol) else 0)
    col_heights = col_total
    # scalar CSC: each scalar column of block column r2 has col_heights[r2] rows (full blocks)
    nr = st.nr
    colh_scalar = np.repeat(col_heights, rd)
    S_colptr = np.ones(nr + 1, dtype=I64)
    np.cumsum(colh_scalar, out=S_colptr[1:])
    S_colptr[1:] += 1
    # row values of one scalar column of block column r2: concatenated row ranges of its blocks
    rows_per_block = [np.arange(st.boffsets_r[r], st.boffsets_r[r] + rd[r], dtype=I64) for r in range(nrb)]
    colrows = []
    for r2 in range(nrb):
        blks = brow[block_colptr[r2]:block_colptr[r2 + 1]]
        colrows.append(np.concatenate([rows_per_block[r1] for r1 in blks]))
    S_rowval = np.concatenate([np.tile(colrows[r2], rd[r2]) for r2 in range(nrb)]) if nrb else np.zeros(0, I64)
    # diag_inds: the stored diagonal entry of each scalar column
    diag_inds = np.empty(nr, dtype=I64)
    for r2 in range(nrb):
        ro = st.boffsets_r[r2]
        idx = np.searchsorted(block_rowval[block_colptr[r2]:block_colptr[r2 + 1]], r2 + 1) + block_colptr[r2]
        off = block_row_offsets[idx]
        cols = np.arange(rd[r2], dtype=I64)
        diag_inds[ro - 1 + cols] = S_colptr[ro - 1 + cols] + off + cols
    return SchurPattern(block_colptr=block_colptr + 1, block_rowval=block_rowval, block_row_offsets=block_row_offsets,
                        col_heights=col_heights, S_colptr=S_colptr, S_rowval=S_rowval, nnz=int(S_colptr[-1] - 1),
                        diag_inds=diag_inds, nr=nr, nnz_blocks=len(key))


class SchurReductionWorkspaceIndex:
    """The integer part of SchurReductionWorkspace (src/schur_reduction.jl:51-92): retained-block-major index of
    the couplings and the phase-2 column chunks."""

    def __init__(self, st, nchunks):
        ncoup = st.ncoup
        owner = np.repeat(np.arange(1, st.neblocks + 1, dtype=I64), np.diff(st.eliminated_ptr))
        self.entry_eliminated = owner
        counts = np.bincount(st.entry_retained - 1, minlength=st.nrblocks) if ncoup else np.zeros(st.nrblocks, I64)
        self.retained_ptr = np.ones(st.nrblocks + 1, dtype=I64)
        np.cumsum(counts, out=self.retained_ptr[1:])
        self.retained_ptr[1:] += 1
        # ascending c within each retained block (stable sort by retained block)
        self.retained_entry = (np.argsort(st.entry_retained, kind="stable") + 1).astype(I64)
        self.col_chunks = partition_columns(self.retained_ptr, max(nchunks, 1))


class SparseBlockLinearSystem:
    """Block storage of SparseBlockLinearSystem (src/SparseBlockLinearSystem.jl:11-64): blocks column-major,
    contiguous in `data`, ordered by block row then block column; `start` gives each stored block's 1-based start.
    Built from a pattern given transposed (TransposedPattern): entries (j, i) for stored blocks (i, j)."""

    def __init__(self, pattern_t_colptr, pattern_t_rowval, rowblocksizes, colblocksizes=None):
        rbs = np.asarray(rowblocksizes, dtype=I64)
        cbs = rbs if colblocksizes is None else np.asarray(colblocksizes, dtype=I64)
        self.rowblocksizes, self.columnblocksizes = rbs, cbs
        cp = np.asarray(pattern_t_colptr, dtype=I64)   # 1-based colptr over row blocks
        rv = np.asarray(pattern_t_rowval, dtype=I64)   # block columns, 1-based
        nb = len(rv)
        brow = np.repeat(np.arange(1, len(rbs) + 1, dtype=I64), np.diff(cp))
        bcol = rv
        sz = rbs[brow - 1] * cbs[bcol - 1]
        self.start = startindices(sz) if nb else np.zeros(0, I64)
        self.brow, self.bcol = brow, bcol
        self.nnzvals = int(sz.sum())
        self.m, self.n = int(rbs.sum()), int(cbs.sum())
        # the non-transposed index: per block column, stored blocks ordered by block row
        o = np.lexsort((brow, bcol))
        self.ind_colptr = np.ones(len(cbs) + 1, dtype=I64)
        np.cumsum(np.bincount(bcol - 1, minlength=len(cbs)), out=self.ind_colptr[1:])
        self.ind_colptr[1:] += 1
        self.ind_rowval, self.ind_start = brow[o], self.start[o]
        self.it_colptr, self.it_rowval, self.it_start = cp, rv, self.start

    @classmethod
    def from_blocks(cls, blockcoords, rowblocksizes, colblocksizes=None):
        """SparseBlockLinearSystem{T}(blockcoords, rowblocksizes, colblocksizes) (src/SparseBlockLinearSystem.jl:70-77)."""
        bc = sorted({(int(i), int(j)) for (i, j) in blockcoords})
        nrow = len(rowblocksizes)
        cp = np.ones(nrow + 1, dtype=I64)
        counts = np.zeros(nrow, dtype=I64)
        for (i, _j) in bc:
            counts[i - 1] += 1
        np.cumsum(counts, out=cp[1:]); cp[1:] += 1
        rv = np.array([j for (i, j) in sorted(bc, key=lambda t: (t[0], t[1]))], dtype=I64)
        return cls(cp, rv, rowblocksizes, colblocksizes)

    def block_start(self, i, j):
        lo, hi = self.it_colptr[i - 1] - 1, self.it_colptr[i] - 1
        k = np.searchsorted(self.it_rowval[lo:hi], j)
        if k < hi - lo and self.it_rowval[lo + k] == j:
            return int(self.it_start[lo + k])
        return 0


def hessian_retained_storage(st, pattern):
    """The SparseBlockLinearSystem of a sparse SchurBlockHessian's R: TransposedPattern(block_pattern)
    (src/schur_hessian.jl:279-291), i.e. the lower blocks (r2, r1) for every upper pattern block (r1, r2)."""
    return SparseBlockLinearSystem(pattern.block_colptr, pattern.block_rowval, st.r_dims, st.r_dims)


def makesparseindices(bsm, symmetrify=False):
    """makesparseindices (src/SparseBlockLinearSystem.jl:233-282): the CSC (colptr, rowval, nzval) whose nzval
    are 1-based indices into bsm's data; symmetrify mirrors the strictly-upper blocks of a lower storage."""
    rbs, cbs = bsm.rowblocksizes, bsm.columnblocksizes
    if symmetrify and not np.array_equal(rbs, cbs):
        raise ValueError("SparseBlockLinearSystem must be block-square for symmetrification")
    startrow = np.ones(len(rbs) + 1, dtype=I64)
    np.cumsum(rbs, out=startrow[1:]); startrow[1:] += 1
    rows_out, idx_out, colptr = [], [], [1]
    for c in range(1, len(cbs) + 1):
        cbsz = int(cbs[c - 1])
        lo, hi = bsm.ind_colptr[c - 1] - 1, bsm.ind_colptr[c] - 1
        lower_rows, lower_start = bsm.ind_rowval[lo:hi], bsm.ind_start[lo:hi]
        if symmetrify:
            ulo, uhi = bsm.it_colptr[c - 1] - 1, bsm.it_colptr[c] - 1
            if len(lower_rows) and lower_rows[0] == c:
                uhi -= 1     # the diagonal block is emitted with the lower ones
            upper_rows, upper_start = bsm.it_rowval[ulo:uhi], bsm.it_start[ulo:uhi]
        else:
            upper_rows = upper_start = np.zeros(0, I64)
        for ic in range(cbsz):
            n = 0
            for row, st0 in zip(upper_rows, upper_start):     # mirrored: element (ic, i) of block (c, row)
                cc = int(rbs[row - 1])
                rows_out.append(startrow[row - 1] + np.arange(cc, dtype=I64))
                idx_out.append(st0 + ic + cbsz * np.arange(cc, dtype=I64))
                n += cc
            for row, st0 in zip(lower_rows, lower_start):     # element (i, ic) of block (row, c)
                cc = int(rbs[row - 1])
                rows_out.append(startrow[row - 1] + np.arange(cc, dtype=I64))
                idx_out.append(st0 + ic * cc + np.arange(cc, dtype=I64))
                n += cc
            colptr.append(colptr[-1] + n)
    rowval = np.concatenate(rows_out) if rows_out else np.zeros(0, I64)
    nzval = np.concatenate(idx_out) if idx_out else np.zeros(0, I64)
    return np.asarray(colptr, dtype=I64), rowval, nzval


def sparse_with_map(bsm, symmetric=False):
    """sparse_with_map (src/SparseBlockLinearSystem.jl:343-354): the CSC pattern and its scatter map."""
    colptr, rowval, nzidx = makesparseindices(bsm, symmetric)
    return colptr, rowval, nzidx.copy()


def build_r_to_s_map(R, pattern, boffsets_r):
    """build_r_to_s_map(R_sys, pattern, boffsets_r) (src/schur_structure.jl:406-431): for each entry of R's data,
    its 1-based index in S.nzval (0 for the strict lower triangle of diagonal blocks). Vectorized over the stored
    blocks of R, grouped by block shape."""
    out = np.zeros(R.nnzvals, dtype=I64)
    brow, bcol, ds = np.asarray(R.brow, I64), np.asarray(R.bcol, I64), np.asarray(R.start, I64)
    keep = brow >= bcol
    brow, bcol, ds = brow[keep], bcol[keep], ds[keep]
    if len(brow) == 0:
        return out
    nrb = len(pattern.block_colptr) - 1
    bc = np.asarray(pattern.block_colptr, I64)
    rv = np.asarray(pattern.block_rowval, I64)
    col_of_slot = np.repeat(np.arange(1, nrb + 1, dtype=I64), np.diff(bc))
    keys = col_of_slot * (nrb + 1) + rv                       # ascending: slots are column-major, rows sorted
    want = brow * (nrb + 1) + bcol                            # block (bcol, brow) of the upper pattern
    k = np.searchsorted(keys, want)
    bad = (k >= len(keys)) | (keys[np.minimum(k, len(keys) - 1)] != want)
    if bad.any():
        i = int(np.argmax(bad))
        raise ValueError(f"Block ({int(bcol[i])}, {int(brow[i])}) of R is not present in the pattern of S")
    base = np.asarray(pattern.S_colptr, I64)[np.asarray(boffsets_r, I64)[brow - 1] - 1] + \
        np.asarray(pattern.block_row_offsets, I64)[k]
    stride = np.asarray(pattern.col_heights, I64)[brow - 1]
    drow = np.asarray(R.rowblocksizes, I64)[brow - 1]
    dcol = np.asarray(R.columnblocksizes, I64)[bcol - 1]
    for dr, dc in set(zip(drow.tolist(), dcol.tolist())):
        col, row = np.meshgrid(np.arange(1, dc + 1, dtype=I64), np.arange(1, dr + 1, dtype=I64), indexing="ij")
        col, row = col.ravel()[None, :], row.ravel()[None, :]
        idx = np.flatnonzero((drow == dr) & (dcol == dc))
        for c0 in range(0, len(idx), 1 << 16):      # bounded temporaries
            m = idx[c0:c0 + (1 << 16)]
            pos = ds[m][:, None] + (col - 1) * dr + row - 1
            diag = (brow[m] == bcol[m])[:, None]
            upper = base[m][:, None] + (col - 1) * stride[m][:, None] + row - 1   # diagonal block: (row, col)
            lower = base[m][:, None] + (row - 1) * stride[m][:, None] + col - 1   # lower block (brow, bcol): transposed
            val = np.where(diag, upper, lower)
            sel = ~diag | (row <= col)
            out[pos[sel] - 1] = val[sel]
    return out


def reduced_symmetric_scatter(st, pattern):
    """The full symmetric CSC of the reduced system and the scatter map into S.nzval of the upper-block pattern:
    entry (i, j) of the full matrix reads S.nzval at the upper element (min, max) of its block pair, diagonal
    blocks mirrored from their upper triangle. Applying it is copyto!(csc, S, scatter_map) on S's values
    (src/SparseBlockLinearSystem.jl:356-362). Returns (colptr, rowval, scatter_map), 1-based."""
    nrb, rd, bo = st.nrblocks, st.r_dims, st.boffsets_r
    bc, brv, bro, ch = pattern.block_colptr - 1, pattern.block_rowval - 1, pattern.block_row_offsets, pattern.col_heights
    Scp = pattern.S_colptr
    # block rows of each full block column c: upper blocks (r, c) r <= c, then lower blocks (c, r) r > c
    bcol_of = np.repeat(np.arange(nrb), np.diff(bc))
    lower_by_col = [[] for _ in range(nrb)]     # for column c, list of (r, slot) with block (c, r) stored, r > c
    for slot in range(len(brv)):
        r1, r2 = brv[slot], bcol_of[slot]
        if r1 < r2:
            lower_by_col[r1].append((r2, slot))
    colptr = [1]
    rows, maps = [], []
    for c in range(nrb):
        blocks = [(int(brv[s]), s, False) for s in range(bc[c], bc[c + 1])] + [(r, s, True) for (r, s) in lower_by_col[c]]
        blocks.sort()
        for j in range(rd[c]):
            for (r, s, lower) in blocks:
                d = rd[r]
                i = np.arange(d, dtype=I64)
                rows.append(bo[r] + i)
                if not lower:   # block (r, c) stored upper in column c: element (i, j)
                    if r == c:
                        a, b = np.minimum(i, j), np.maximum(i, j)
                        maps.append(Scp[bo[c] - 1 + b] + bro[s] + a)
                    else:
                        maps.append(Scp[bo[c] - 1 + j] + bro[s] + i)
                else:           # block (r, c), r > c, is the transpose of stored (c, r) in column r: element (j, i)
                    maps.append(Scp[bo[r] - 1 + i] + bro[s] + j)
            colptr.append(colptr[-1] + sum(rd[r] for (r, _s, _l) in blocks))
    return np.asarray(colptr, dtype=I64), np.concatenate(rows).astype(I64), np.concatenate(maps).astype(I64)


class SchurJacobianStructure:
    """SchurJacobianStructure (src/schur_jacobian.jl:30-160): one dense table per eliminated block
    ``[J_r | J_e]`` (plus its Float64 residual column kept apart), then one table per free group.

    Positional: ``SchurJacobianStructure(schur, e_rows, free_rows=(), free_retained=(), nchunks=None)``.
    From a description: :meth:`from_description` (src/schur_description.jl:137-189)."""

    def __init__(self, schur, e_rows, free_rows=(), free_retained=(), nchunks=None, rowblocks=None):
        st = schur
        ne, nrb = st.neblocks, st.nrblocks
        e_rows = np.asarray(e_rows, dtype=I64)
        free_rows = np.asarray(free_rows, dtype=I64)
        nfree = len(free_rows)
        if nchunks is None:
            nchunks = max([len(c) for c in st.chunks] + [1])
        if len(e_rows) != ne:
            raise DimensionMismatch(f"Number of eliminated row counts ({len(e_rows)}) must match the number of eliminated blocks ({ne})")
        if len(free_retained) != nfree:
            raise DimensionMismatch(f"Number of free groups' retained lists ({len(free_retained)}) must match the number of free row counts ({nfree})")
        if nchunks < 1:
            raise ValueError(f"nchunks must be at least 1, got {nchunks}")
        bad = np.nonzero(e_rows < 1)[0]
        if bad.size:
            raise ValueError(f"Eliminated block {int(bad[0]) + 1} has {int(e_rows[bad[0]])} rows; every eliminated block needs at least 1 row")
        bad = np.nonzero(free_rows < 1)[0]
        if bad.size:
            raise ValueError(f"Free group {int(bad[0]) + 1} has {int(free_rows[bad[0]])} rows; every free group needs at least 1 row")
        for g, rs in enumerate(free_retained):
            seen = set()
            for r in rs:
                if not (1 <= r <= nrb):
                    raise ValueError(f"Free group {g + 1} references retained block {r}, but there are only {nrb} retained blocks")
                if r in seen:
                    raise ValueError(f"Retained block {r} is listed more than once in free group {g + 1}")
                seen.add(r)
        ntables = ne + nfree
        sorted_free = [np.sort(np.asarray(rs, dtype=I64)) for rs in free_retained]
        # entries: the coupling entries of the Schur structure, then the free groups' (sorted)
        k_e = np.diff(st.eliminated_ptr)
        k_f = np.asarray([len(x) for x in sorted_free], dtype=I64)
        counts = np.concatenate([k_e, k_f])
        self.table_ptr = np.ones(ntables + 1, dtype=I64)
        np.cumsum(counts, out=self.table_ptr[1:]); self.table_ptr[1:] += 1
        self.entry_retained = np.concatenate([st.entry_retained] + sorted_free).astype(I64) if ntables else np.zeros(0, I64)
        self.entry_table = np.repeat(np.arange(1, ntables + 1, dtype=I64), counts)
        widths = st.r_dims[self.entry_retained - 1] if len(self.entry_retained) else np.zeros(0, I64)
        # first column of each entry inside its table: exclusive prefix of widths within the table
        cw = np.cumsum(widths)
        first_of_table = np.repeat(self.table_ptr[:-1] - 1, counts)
        base = np.where(first_of_table > 0, cw[np.maximum(first_of_table - 1, 0)], 0) if len(widths) else np.zeros(0, I64)
        self.entry_col = (cw - widths - base + 1).astype(I64) if len(widths) else np.zeros(0, I64)
        retained_cols = np.zeros(ntables, dtype=I64)
        if len(widths):
            np.add.at(retained_cols, self.entry_table - 1, widths)
        de = np.concatenate([st.e_dims, np.zeros(nfree, dtype=I64)])
        self.table_top = de.copy()
        self.table_nrows = np.concatenate([e_rows + st.e_dims, free_rows]).astype(I64)
        self.table_ncols = (retained_cols + de + 1).astype(I64)
        self.table_offsets = np.ones(ntables + 1, dtype=I64)
        np.cumsum(self.table_nrows * self.table_ncols, out=self.table_offsets[1:]); self.table_offsets[1:] += 1
        self.raw_offsets = np.ones(ntables + 1, dtype=I64)
        np.cumsum(self.table_nrows * (self.table_ncols - 1), out=self.raw_offsets[1:]); self.raw_offsets[1:] += 1
        self.residual_offsets = np.ones(ntables + 1, dtype=I64)
        np.cumsum(self.table_nrows, out=self.residual_offsets[1:]); self.residual_offsets[1:] += 1
        self.reduced_offsets = np.ones(ntables + 1, dtype=I64)
        np.cumsum(self.table_nrows - self.table_top, out=self.reduced_offsets[1:]); self.reduced_offsets[1:] += 1
        cnt = np.bincount(self.entry_retained - 1, minlength=nrb) if len(self.entry_retained) else np.zeros(nrb, I64)
        self.retained_ptr = np.ones(nrb + 1, dtype=I64)
        np.cumsum(cnt, out=self.retained_ptr[1:]); self.retained_ptr[1:] += 1
        self.retained_entry = (np.argsort(self.entry_retained, kind="stable") + 1).astype(I64)
        self.doffsets = np.ones(nrb + 1, dtype=I64)
        np.cumsum(st.r_dims ** 2, out=self.doffsets[1:]); self.doffsets[1:] += 1
        self.col_chunks = partition_columns(self.retained_ptr, nchunks)
        self.schur, self.e_rows, self.free_rows, self.ntables = st, e_rows, free_rows, ntables
        self.free_retained = sorted_free
        self.rowblocks = rowblocks

    def raw_length(self):
        return int(self.raw_offsets[-1] - 1)

    def residual_length(self):
        return int(self.residual_offsets[-1] - 1)

    def reduced_nrows(self):
        return int(self.reduced_offsets[-1] - 1)

    def coupling_entry(self, t, r):
        """coupling_entry (src/schur_jacobian.jl:181-187)."""
        if not (1 <= t <= self.ntables):
            raise IndexError(t)
        lo, hi = self.table_ptr[t - 1] - 1, self.table_ptr[t] - 1
        k = np.searchsorted(self.entry_retained[lo:hi], r)
        if not (k < hi - lo and self.entry_retained[lo + k] == r):
            raise ValueError(f"Retained block {r} is not coupled to table {t}")
        return int(lo + k + 1)

    @classmethod
    def from_description(cls, blocksizes, rowblocks, rowblocksizes, eliminated, nchunks=1, min_work=DEFAULT_MIN_WORK):
        """SchurJacobianStructure(blocksizes, rowblocks, rowblocksizes; eliminated) (src/schur_description.jl:137-189):
        each row block's rows go to its eliminated block's table, or to the table of its free group (row blocks
        with the same set of retained blocks, numbered by first appearance); rows in row-block order."""
        d = describe_elimination(blocksizes, rowblocks, eliminated, rowblocksizes)
        st = description_structure(d, blocksizes, nchunks, min_work)
        nret = d.nretained
        n = len(d.order)
        ne = n - nret
        rptr, rflat = _as_csr(rowblocks)
        K = len(rptr) - 1
        m = np.asarray(rowblocksizes, dtype=I64)
        row_table = np.empty(K, dtype=I64)
        row_first = np.empty(K, dtype=I64)
        e_rows = np.zeros(ne, dtype=I64)
        group, free_retained, free_rows = {}, [], []
        row_blocks = d.position[rflat - 1] if rflat.size else np.zeros(0, I64)
        for k in range(K):
            seg = np.sort(row_blocks[rptr[k]:rptr[k + 1]])
            row_blocks[rptr[k]:rptr[k + 1]] = seg
            e = int(d.rowblock_eliminated[k])
            if e != 0:
                t = int(d.position[e - 1]) - nret
                row_table[k] = t
                row_first[k] = e_rows[t - 1] + 1
                e_rows[t - 1] += m[k]
            else:
                key = tuple(int(x) for x in seg)
                g = group.get(key, 0)
                if g == 0:
                    free_retained.append(list(key)); free_rows.append(0)
                    g = group[key] = len(free_rows)
                row_table[k] = ne + g
                row_first[k] = free_rows[g - 1] + 1
                free_rows[g - 1] += int(m[k])
        rb = dict(table=row_table, first=row_first, nrows=m.copy(), ptr=rptr + 1, blocks=row_blocks)
        return cls(st, e_rows, free_rows=free_rows, free_retained=free_retained, nchunks=nchunks, rowblocks=rb)

    def blockentry(self, k, j):
        """blockentry (src/schur_jacobian.jl:330-337): a retained block's entry in row block k's table (> 0) or
        minus the eliminated block (< 0)."""
        rb = self.rowblocks
        lo, hi = rb["ptr"][k - 1] - 1, rb["ptr"][k] - 1
        blocks = rb["blocks"][lo:hi]
        idx = np.searchsorted(blocks, j)
        if not (idx < hi - lo and blocks[idx] == j):
            raise ValueError(f"Row block {k} does not touch column block {j}")
        nr = self.schur.nrblocks
        return self.coupling_entry(int(rb["table"][k - 1]), j) if j <= nr else -(j - nr)
