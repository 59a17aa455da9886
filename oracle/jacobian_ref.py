"""Quantities of the full problem from the QR tables, and J~ written out (TEST INFRASTRUCTURE ONLY): the checker of
the device's bls_jtr_*, bls_jtj_diag_*, bls_jtj_quadform_*, bls_build_reduced_jacobian_pattern and
bls_reduced_jacobian_*.

Loop-by-loop restatement of src/qr_reduced.jl:529-645 (sum_retained_columns!, foreach_eliminated_column, jtr!,
jtj_diag!, jtj_quadform) over a SchurJacobianStructure's tables (raw or factored, column-major, M rows each) and
their Float64 residual, accumulated in Float64.

build_reduced_jacobian_pattern and reduced_jacobian (src/qr_reduced.jl:247-370) restated the same way.

fill_tables restates the assignments block(J, k, blockentry(jst, k, j)) .= J_kj and residual_block(J, k) .= r_k
(src/schur_jacobian.jl:264-387) for every row block k: the checker of the device's bls_jfill_*.

fill_tables is pinned (tests/test_oracle_qr.py) against the tables that the problem generator writes by its own
vectorised indexing (bitwise), and against the dense J of the row blocks (J'J and J'r of the filled tables).

Parity pinning (tests/test_conditioning.py): against the dense J and r of the same tables, J'r, diag(J'J) and
||J v||^2 by numpy; J~ against the C oracle's reduced_mul / reduced_tmul on the same factored tables (J~ v and J~'w
for random v, w), its pattern against the CSC invariants (ascending rows, nnz = the sum of d_r (M - top) over the
entries) and its dense form against the sparse one, entry for entry (test/schur_qr.jl:180-270).
"""
import numpy as np


def _eliminated_columns(jst, t):
    """(1-based first raw column of table t's eliminated block, its dim): the last d_e of its table_ncols - 1 raw
    columns ((C - de - 2 + k) M from the table's start, k = 1..d_e, :562-567); free tables have none."""
    st = jst.schur
    if t >= st.neblocks:
        return 0, 0
    de = int(st.e_dims[t])
    return jst.table_ncols[t] - de, de


def _columns(jst, D, R, colfun, out):
    st = jst.schur
    for r in range(st.nrblocks):                               # sum_retained_columns! (:529-551)
        ro, dr = st.boffsets_r[r] - 1, int(st.r_dims[r])
        out[ro:ro + dr] = 0.0
        for k in range(jst.retained_ptr[r] - 1, jst.retained_ptr[r + 1] - 1):
            u = jst.retained_entry[k] - 1
            t = jst.entry_table[u] - 1
            M = jst.table_nrows[t]
            cb = jst.raw_offsets[t] - 1 + (jst.entry_col[u] - 1) * M
            rb = jst.residual_offsets[t] - 1
            for j in range(dr):
                out[ro + j] += colfun(D, cb + j * M, R, rb, M)
    for e in range(st.neblocks):                               # foreach_eliminated_column (:553-570)
        M = jst.table_nrows[e]
        c1, de = _eliminated_columns(jst, e)
        eo = st.nr + st.boffsets_e[e] - 1
        for k in range(de):
            out[eo + k] = colfun(D, jst.raw_offsets[e] - 1 + (c1 - 1 + k) * M, R, jst.residual_offsets[e] - 1, M)
    return out


def _dot_run64(a, ia, b, ib, n):
    acc = 0.0
    for i in range(n):
        acc += float(a[ia + i]) * float(b[ib + i])
    return acc


def jtr(jst, D, R):
    """jtr! (:574-583): J'r."""
    n = jst.schur.nr + int(np.sum(jst.schur.e_dims))
    return _columns(jst, D, R, lambda D_, cb, R_, rb, M: _dot_run64(D_, cb, R_, rb, M), np.zeros(n))


def jtj_diag(jst, D):
    """jtj_diag! (:585-595): diag(J'J)."""
    n = jst.schur.nr + int(np.sum(jst.schur.e_dims))
    return _columns(jst, D, None, lambda D_, cb, R_, rb, M: _dot_run64(D_, cb, D_, cb, M), np.zeros(n))


def jtj_quadform(jst, D, v):
    """jtj_quadform (:597-645): ||J v||^2, per table and row."""
    st = jst.schur
    total = 0.0
    for t in range(jst.ntables):
        M = jst.table_nrows[t]
        off = jst.raw_offsets[t] - 1
        c1, de = _eliminated_columns(jst, t)
        eo = st.nr + st.boffsets_e[t] - 1 if de else 0
        part = 0.0
        for i in range(M):
            yi = 0.0
            for u in range(jst.table_ptr[t] - 1, jst.table_ptr[t + 1] - 1):
                r = jst.entry_retained[u] - 1
                ro = st.boffsets_r[r] - 1
                cb = off + (jst.entry_col[u] - 1) * M + i
                for j in range(int(st.r_dims[r])):
                    yi += float(D[cb + j * M]) * v[ro + j]
            cb = off + (c1 - 1) * M + i
            for k in range(de):
                yi += float(D[cb + k * M]) * v[eo + k]
            part += yi * yi
        total += part
    return total


def build_reduced_jacobian_pattern(jst):
    """src/qr_reduced.jl:289-310: (colptr, rowval), 1-based; each column of retained block r holds the reduced rows
    of every table coupled to r, tables in retained_entry order."""
    st = jst.schur
    colptr = [1]
    rowval = []
    for r in range(st.nrblocks):
        for _ in range(int(st.r_dims[r])):
            for k in range(jst.retained_ptr[r] - 1, jst.retained_ptr[r + 1] - 1):
                t = jst.entry_table[jst.retained_entry[k] - 1] - 1
                rowval.extend(jst.reduced_offsets[t] + i for i in range(jst.table_nrows[t] - jst.table_top[t]))
            colptr.append(len(rowval) + 1)
    return np.array(colptr, dtype=np.int64), np.array(rowval, dtype=np.int64)


def reduced_jacobian_sparse(jst, F, colptr):
    """src/qr_reduced.jl:334-370: nzval of J~ in the pattern, from the factored tables F."""
    st = jst.schur
    nz = np.zeros(colptr[-1] - 1, dtype=F.dtype)
    for r in range(st.nrblocks):
        ro, dr = st.boffsets_r[r], int(st.r_dims[r])
        p = 0
        for k in range(jst.retained_ptr[r] - 1, jst.retained_ptr[r + 1] - 1):
            u = jst.retained_entry[k] - 1
            t = jst.entry_table[u] - 1
            M, top = jst.table_nrows[t], jst.table_top[t]
            n = M - top
            cb = jst.raw_offsets[t] - 1 + (jst.entry_col[u] - 1) * M + top
            for j in range(dr):
                q = colptr[ro - 1 + j] - 1 + p
                nz[q:q + n] = F[cb + j * M:cb + j * M + n]
            p += n
    return nz


def reduced_jacobian_dense(jst, F):
    """src/qr_reduced.jl:247-270: J~ as a dense reduced_nrows x nr array."""
    st = jst.schur
    m = int(jst.reduced_offsets[-1] - 1)
    Jt = np.zeros((m, st.nr), dtype=F.dtype)
    for t in range(jst.ntables):
        M, top = jst.table_nrows[t], jst.table_top[t]
        off, yo = jst.raw_offsets[t] - 1, jst.reduced_offsets[t] - 1
        for u in range(jst.table_ptr[t] - 1, jst.table_ptr[t + 1] - 1):
            r = jst.entry_retained[u] - 1
            ro = st.boffsets_r[r] - 1
            cb = off + (jst.entry_col[u] - 1) * M + top
            for j in range(int(st.r_dims[r])):
                Jt[yo:yo + M - top, ro + j] = F[cb + j * M:cb + j * M + M - top]
    return Jt


def fill_tables(jst, vals, res, raw, r):
    """Write row block k's blocks and residual rows into the raw tables and the residual, k = 1, 2, ...: block j
    of the row block (its layout column blocks ascending) is an m_k x d column-major run of ``vals``; the run of row
    block k follows that of k - 1. blockentry (:330-337) gives a retained block's entry in the table (its columns
    from entry_col) or minus the eliminated block (columns table_ncols - d_e, ...); the rows start at row_first."""
    rb, st = jst.rowblocks, jst.schur
    src = rsrc = 0
    for k in range(1, len(rb["table"]) + 1):
        t, first, m = int(rb["table"][k - 1]), int(rb["first"][k - 1]), int(rb["nrows"][k - 1])
        M = int(jst.table_nrows[t - 1])
        base = int(jst.raw_offsets[t - 1]) - 1
        for j in rb["blocks"][rb["ptr"][k - 1] - 1: rb["ptr"][k] - 1]:
            entry = jst.blockentry(k, int(j))
            if entry > 0:
                col = int(jst.entry_col[entry - 1])
                d = int(st.r_dims[jst.entry_retained[entry - 1] - 1])
            else:
                d = int(st.e_dims[-entry - 1])
                col = int(jst.table_ncols[-entry - 1]) - d
            for c in range(d):
                for i in range(m):
                    raw[base + (col - 1 + c) * M + first - 1 + i] = vals[src]
                    src += 1
        ro = int(jst.residual_offsets[t - 1]) - 1
        for i in range(m):
            r[ro + first - 1 + i] = res[rsrc]
            rsrc += 1
    return src, rsrc
