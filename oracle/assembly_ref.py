"""Loop-by-loop restatement of addblock! (TEST INFRASTRUCTURE ONLY): filling a Schur Hessian block by block.

Plain Python over numpy blocks, written from src/schur_assembly.jl:140-246: the checker of the device's batched
assembly (bls_asm_hessian_f64). A row block k with Jacobian J_k over blocks (i_1, ..., i_p) (its column order) and
residual r_k calls addblock!(H, i_a, i_b, B, rows_a, rows_b) with B = J_k' J_k for every pair a <= b in that order,
and adds J_k' r_k to b; H starts at zero. Layout block numbers are 1-based: retained 1..nrb, eliminated nrb + e.

Also dot(x, H, y) over that storage (schur_bilinear, src/linalg.jl:145-260), term by term, and its diagonal
(getdiag!, src/schur_linear_systems.jl:239-270).

Parity pinning: the reference docstring's own example (src/schur_assembly.jl:126-137: a point and a camera, B =
ones(9, 9), sums (9, 18, 36)), and H assembled here against the dense J'J of the stacked row blocks, read back
through the storage layouts; bilinear against x' (J'J) y; getdiag against diag(J'J), and zeros for a retained block
whose diagonal block a sparse R does not store (tests/test_assembly.py, tests/test_conditioning.py).
"""
import numpy as np


class HessianStorage:
    """E_data, C_data and R (dense nr x nr, or the stored lower blocks of a SparseBlockLinearSystem given as a block
    CSC of block rows >= the column with 1-based data offsets) of a SchurStructure, zero-initialised."""

    def __init__(self, st, r_colptr=None, r_rowval=None, r_offsets=None):
        self.st = st
        self.E = np.zeros(int(np.sum(np.asarray(st.e_dims) ** 2)))
        rd = np.asarray(st.r_dims)
        er = np.asarray(st.entry_retained) - 1
        ep = np.asarray(st.eliminated_ptr) - 1
        nc = int(ep[-1]) if len(ep) else 0
        ed = np.repeat(np.asarray(st.e_dims), np.diff(ep)) if nc else np.zeros(0, np.int64)
        self.C = np.zeros(int(np.sum(ed * rd[er])) if nc else 0)
        self.nr = int(np.sum(rd))
        self.dense = r_colptr is None
        if self.dense:
            self.R = np.zeros((self.nr, self.nr))
        else:
            self.r_colptr, self.r_rowval, self.r_offsets = r_colptr, r_rowval, r_offsets
            last = [r_offsets[k] - 1 + rd[r_rowval[k] - 1] * rd[q] for q in range(len(rd))
                    for k in range(r_colptr[q] - 1, r_colptr[q + 1] - 1)]
            self.R = np.zeros(int(max(last)) if last else 0)
        self.b_r = np.zeros(self.nr)
        self.b_e = np.zeros(int(np.sum(st.e_dims)))

    def r_block_offset(self, i, j):
        """0-based start of stored lower block (i, j), i >= j, 1-based block numbers (indicestransposed[j, i])."""
        for k in range(self.r_colptr[j - 1] - 1, self.r_colptr[j] - 1):
            if self.r_rowval[k] == i:
                return self.r_offsets[k] - 1
        raise ValueError(f"Blocks {i} and {j} share no row block: the system has no storage for the pair")


def find_coupling(st, e, r):
    """src/schur_assembly.jl:94-98: coupling entry (1-based) of retained r in eliminated e, 0 if none."""
    lo, hi = st.eliminated_ptr[e - 1], st.eliminated_ptr[e] - 1
    for c in range(lo, hi + 1):
        if st.entry_retained[c - 1] == r:
            return c
    return 0


def addblock(H, i, j, B, rows, cols):
    """src/schur_assembly.jl:140-159 with check_addblock's storage test (:161-181) and the add_* helpers
    (:189-246). rows / cols: 0-based index lists into B."""
    st = H.st
    nr = st.nrblocks
    if i == j:
        if i <= nr:                                        # add_retained_diagonal! (:189-201)
            n = len(rows)
            if H.dense:
                ro = st.boffsets_r[i - 1] - 1
                for col in range(n):
                    for row in range(n):
                        H.R[ro + row, ro + col] += B[rows[row], rows[col]]
            else:
                off = H.r_block_offset(i, i)
                for col in range(n):
                    for row in range(n):
                        H.R[off + col * n + row] += B[rows[row], rows[col]]
        else:                                              # add_eliminated! (:226-234)
            idx = st.eoffsets[i - nr - 1] - 1
            n = len(rows)
            for col in range(n):
                for row in range(n):
                    H.E[idx] += B[rows[row], rows[col]]
                    idx += 1
    elif i <= nr and j <= nr:                              # add_retained_pair! (:206-223)
        if H.dense:
            a, ra, b, rb = (i, rows, j, cols) if i < j else (j, cols, i, rows)
            ro_a, ro_b = st.boffsets_r[a - 1] - 1, st.boffsets_r[b - 1] - 1
            for col in range(len(rb)):
                for row in range(len(ra)):
                    H.R[ro_a + row, ro_b + col] += B[ra[row], rb[col]]
        else:
            if i > j:
                hi, rh, lo, rl = i, rows, j, cols
            else:
                hi, rh, lo, rl = j, cols, i, rows
            off = H.r_block_offset(hi, lo)
            for col in range(len(rl)):
                for row in range(len(rh)):
                    H.R[off + col * len(rh) + row] += B[rh[row], rl[col]]
    elif i > nr and j > nr:
        raise ValueError(f"Blocks {i} and {j} are both eliminated: the system has no storage for the pair")
    else:                                                  # add_coupling! (:238-246)
        if i <= nr:
            e, r, rr, re = j - nr, i, rows, cols
        else:
            e, r, rr, re = i - nr, j, cols, rows
        c = find_coupling(st, e, r)
        if c == 0:
            raise ValueError(f"Blocks {i} and {j} share no row block: the system has no storage for the pair")
        co = st.coffsets[c - 1] - 1
        de = len(re)
        for col in range(de):
            for row in range(len(rr)):
                H.C[co + row * de + col] += B[rr[row], re[col]]


def assemble(H, rowblocks, rb_rows, J, r):
    """Every pair of blocks of every row block, in row-block order (the device's order); b += J_k' r_k.
    rowblocks: lists of 1-based layout block numbers; J, r: the row blocks one after another (column-major)."""
    st = H.st
    nrb = st.nrblocks
    dims = list(st.r_dims) + list(st.e_dims)
    joff = roff = 0
    for blocks, m in zip(rowblocks, rb_rows):
        W = sum(int(dims[b - 1]) for b in blocks)
        Jk = np.asarray(J[joff:joff + m * W]).reshape(W, m).T
        rk = np.asarray(r[roff:roff + m])
        B = Jk.T @ Jk
        g = Jk.T @ rk
        cols, c0 = [], 0
        for b in blocks:
            cols.append(list(range(c0, c0 + int(dims[b - 1]))))
            c0 += int(dims[b - 1])
        for a2 in range(len(blocks)):
            for a1 in range(a2 + 1):
                addblock(H, blocks[a1], blocks[a2], B, cols[a1], cols[a2])
        for bl, cl in zip(blocks, cols):
            if bl <= nrb:
                o = st.boffsets_r[bl - 1] - 1
                H.b_r[o:o + len(cl)] += g[cl]
            else:
                o = st.boffsets_e[bl - nrb - 1] - 1
                H.b_e[o:o + len(cl)] += g[cl]
        joff += m * W
        roff += m
    return H


def bilinear(H, x, y):
    """schur_bilinear (src/linalg.jl:145-260): x' H y from the stored triangle of a dense R (:153-161) or the stored
    blocks of a sparse R (:162-190), then x_e' E y_e and the coupling terms per eliminated block (:208-260)."""
    st = H.st
    nr = H.nr
    xr, xe, yr, ye = x[:nr], x[nr:], y[:nr], y[nr:]
    total = 0.0
    if H.dense:
        for j in range(nr):
            for i in range(j):
                total += H.R[i, j] * xr[i] * yr[j] + H.R[i, j] * xr[j] * yr[i]
            total += H.R[j, j] * xr[j] * yr[j]
    else:
        rd, bo = st.r_dims, np.asarray(st.boffsets_r) - 1
        for q in range(st.nrblocks):
            for k in range(H.r_colptr[q] - 1, H.r_colptr[q + 1] - 1):
                p = H.r_rowval[k] - 1
                off, drow, dcol = H.r_offsets[k] - 1, int(rd[p]), int(rd[q])
                for jc in range(dcol):
                    bx = by = 0.0
                    for ir in range(drow):
                        bv = H.R[off + jc * drow + ir]
                        bx += bv * xr[bo[p] + ir]
                        by += bv * yr[bo[p] + ir]
                    total += bx * yr[bo[q] + jc]
                    if p != q:
                        total += by * xr[bo[q] + jc]
    for e in range(st.neblocks):
        d = int(st.e_dims[e])
        eo, be = st.eoffsets[e] - 1, st.boffsets_e[e] - 1
        for col in range(d):
            for row in range(d):
                total += H.E[eo + col * d + row] * xe[be + row] * ye[be + col]
        for c in range(st.eliminated_ptr[e] - 1, st.eliminated_ptr[e + 1] - 1):
            r = st.entry_retained[c] - 1
            dr, ro, co = int(st.r_dims[r]), st.boffsets_r[r] - 1, st.coffsets[c] - 1
            for ir in range(dr):
                cx = sum(H.C[co + ir * d + i] * xe[be + i] for i in range(d))
                cy = sum(H.C[co + ir * d + i] * ye[be + i] for i in range(d))
                total += xr[ro + ir] * cy + yr[ro + ir] * cx
    return total



def getdiag(H):
    """getdiag! (src/schur_linear_systems.jl:239-270): [diag(R); diag(E_1); ...]; diag_retained! reads the diagonal
    block of a sparse R, zeros where it is not stored."""
    st = H.st
    d = np.zeros(H.nr + int(np.sum(st.e_dims)))
    if H.dense:
        d[:H.nr] = np.diag(H.R)
    else:
        for b in range(st.nrblocks):
            db, bo = int(st.r_dims[b]), st.boffsets_r[b] - 1
            try:
                off = H.r_block_offset(b + 1, b + 1)
            except ValueError:
                continue
            for i in range(db):
                d[bo + i] = H.R[off + i * (db + 1)]
    for e in range(st.neblocks):
        eo, de = st.eoffsets[e] - 1, int(st.e_dims[e])
        o = H.nr + st.boffsets_e[e] - 1
        for i in range(de):
            d[o + i] = H.E[eo + i * (de + 1)]
    return d
