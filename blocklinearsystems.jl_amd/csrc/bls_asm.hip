// Assembly of a Schur Hessian from row blocks on gfx950: the batched form of addblock!
// (src/schur_assembly.jl:140-246) behind the C ABI of include/bls_amd.h.
//
// A row block k (observation rows touching some layout blocks, in the column order of its Jacobian J_k) contributes
// J_k' J_k to every pair of its blocks and J_k' r_k to each block. One read of J_k forms every pair block into a
// contribution buffer -- by 16 lanes and shuffles for a row block of at most 16 columns and 8 rows, else as a table
// of the Gram pass (bls_gram.h) on the FP64 matrix instruction -- and every stored block of E_data, C_data and R, and
// each block of b_r and b_e, is then the sum of its contributions in ascending row-block order, the order in which
// addblock! calls made in row-block order add them (blocks with few contributions several per wavefront, the others
// by runs over a workgroup).
// No atomics; bitwise reproducible. A pair's contribution is read in the orientation its destination stores
// (add_retained_diagonal!, add_retained_pair!, add_eliminated!, add_coupling!, src/schur_assembly.jl:189-246).
//
// The handle's storage layout also serves the queries that read a filled Hessian: dot(x, H, y) (src/linalg.jl:133-260,
// per-block partials and a fixed-order sum) and getdiag! (src/schur_linear_systems.jl:239-270).
#include <algorithm>
#include <type_traits>
#include <cmath>
#include <memory>

#include "bls_internal.h"

using namespace bls;

namespace {
#include "bls_gram.h"
#include "bls_fsum.h"

constexpr int TINY_W = 16, TINY_M = 8;   // row blocks of at most 16 columns and 8 rows take the tiny path

// Tiny row blocks (observations: 2 rows over a 3- and a 6-column block): `slot` lanes per row block (the widest tiny
// row block, at most 16; 64 / slot row blocks per wavefront, the lanes left over idle), lane c holding column c of J_k. Column a is passed round by shuffle and lane b (>= a) forms element (a, b) of J_k'J_k, written
// where the Gram pass writes it: pair block (e1, e2) of the row block's entries at Q[e2] + d2 P[e1], row-major, so
// (a, b) -> goff + Q[e2] + (b - P[e2]) + d2 a (the columns of a row block are its entries' in order, from 0); and
// lane c forms (J_k' r_k)[c] at eoff + c.
// A row block of two entries, one eliminated and one retained, whose coupling block has no other contribution
// (cw[t] = 2 o + [entry 0 is the eliminated one], o the block's offset in C_data; else -1) writes that pair block
// straight into C_data (d_e x d_r column-major at o) as 0 + value, the ordered sum of its one contribution.
template <class T>
__global__ __launch_bounds__(256) void k_asm_tiny(const RGTab *__restrict__ tabs, int64_t cnt, int slot,
                                                  const int32_t *__restrict__ ecd, const T *__restrict__ F,
                                                  const double *__restrict__ Fr, double *__restrict__ G,
                                                  double *__restrict__ gp, double *__restrict__ Cd,
                                                  const int64_t *__restrict__ cw) {
    const int wl = threadIdx.x & 63, per = 64 / slot, grp = wl / slot;
    const int l = wl - grp * slot, base = grp * slot;   // idle lanes (grp == per) join the shuffles, write nothing
    const int64_t t = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * per + grp;
    const bool valid = grp < per && t < cnt;
    const RGTab g = rg_load(tabs + (valid ? t : 0));
    const int m = valid ? g.n : 0, W = valid ? g.W : 0;
    const int64_t cwt = valid ? cw[t] : -1;
    // a coupling written in place: entry 0's columns are [0, d0), entry 1's [d0, W)
    const int d0 = cwt >= 0 ? (ecd[g.u0] & 0xffff) : W, efirst = (int)(cwt & 1);
    const int64_t cbase = cwt >> 1;
    double x[TINY_M];
#pragma unroll
    for (int i = 0; i < TINY_M; i++) x[i] = (l < W && i < m) ? (double)F[g.raw_off + (int64_t)l * m + i] : 0.0;
    // the entry holding column l: its offset Q[e] + (l - P[e]) in the row block's pair blocks, and its dim
    int qb = 0, db = 1;
    if (l < W) {
        int q = 0;
        for (int e = 0; e < g.ne; e++) {
            const int w = ecd[g.u0 + e], cst = w >> 16, d = w & 0xffff;
            if (l >= cst && l < cst + d) {
                qb = q + (l - cst);
                db = d;
            }
            q += d * (cst + d);   // Q[e + 1] = Q[e] + d_e P[e + 1]
        }
        double acc = 0.0;
#pragma unroll
        for (int i = 0; i < TINY_M; i++)
            if (i < m) acc += x[i] * Fr[g.res_off + i];
        gp[g.eoff + l] = acc;
    }
    // the wavefront's widest and tallest row block bound the shuffles (uniform across the wavefront)
    int wmax = W, mmax = m;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        wmax = max(wmax, __shfl_xor(wmax, o, 64));
        mmax = max(mmax, __shfl_xor(mmax, o, 64));
    }
    for (int a = 0; a < wmax; a++) {
        double xa[TINY_M];
#pragma unroll
        for (int i = 0; i < TINY_M; i++) xa[i] = i < mmax ? __shfl(x[i], base + a, 64) : 0.0;
        if (a < W && a <= l && l < W) {
            double acc = 0.0;
#pragma unroll
            for (int i = 0; i < TINY_M; i++)
                if (i < mmax) acc += xa[i] * x[i];
            if (a < d0 && l >= d0)   // element (a, l) of the coupling pair block (only when d0 < W: cw >= 0)
                Cd[efirst ? cbase + (int64_t)(l - d0) * d0 + a : cbase + (int64_t)a * (W - d0) + (l - d0)] = 0.0 + acc;
            else
                G[g.goff + qb + (int64_t)db * a] = acc;
        }
    }
}

// Destination blocks with few contributions (at most LIGHT_MAX): `slot` lanes per block (the largest block of the
// launch, at most 64), 64 / slot whole blocks per wavefront (the lanes left over stay idle); lane w of a block sums entry w (and
// w + slot, ...) over the block's contributions in order (as k_red_gram_sum, without runs).
constexpr int LIGHT_MAX = 32;
template <class T>
__global__ __launch_bounds__(256) void k_red_gram_sum_light(int64_t nblk, int slot, const RGBlock *__restrict__ blocks,
                                                            const int64_t *__restrict__ ptr,
                                                            const int64_t *__restrict__ contrib,
                                                            const double *__restrict__ G, T *__restrict__ S) {
    const int lane = threadIdx.x & 63, per = 64 / slot;
    if (lane >= per * slot) return;   // lanes past the wavefront's last whole slot; no synchronisation below
    const int64_t b = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * per + lane / slot;
    if (b >= nblk) return;
    const RGBlock bk = blocks[b];
    const int sz = bk.d1 * bk.d2;
    const int64_t cb = ptr[b], ce = ptr[b + 1];
    const bool tr = bk.flags & RG_TRANSPOSED, full = bk.flags & RG_FULL;
    for (int w = lane % slot; w < sz; w += slot) {
        const int row = w / bk.d2, col = w - row * bk.d2;
        const bool lower = bk.diag && row > col;
        double acc = 0.0;
        if (!lower) {
            const int wt = col * bk.d1 + row;
            auto at = [&](int64_t ci) { return ci >= 0 ? ci + w : ~ci + wt; };
            int64_t c = cb;
            for (; c + 4 <= ce; c += 4) {   // four contributions' loads in flight, added in order
                const int64_t i0 = contrib[c], i1 = contrib[c + 1], i2 = contrib[c + 2], i3 = contrib[c + 3];
                const double v0 = G[at(i0)], v1 = G[at(i1)], v2 = G[at(i2)], v3 = G[at(i3)];
                acc += v0;
                acc += v1;
                acc += v2;
                acc += v3;
            }
            for (; c < ce; c++) acc += G[at(contrib[c])];
        }
        const int64_t at = tr ? bk.dst + (int64_t)row * bk.cstride + col : bk.dst + (int64_t)col * bk.cstride + row;
        if (!(lower && full)) S[at] = (T)acc;
        if (full && bk.diag && row < col) S[bk.dst + (int64_t)row * bk.cstride + col] = (T)acc;
    }
}

// dot(x, H, y) (src/linalg.jl:133-260), x and y laid out [x_r; x_e]. Partials in a fixed order, then one sum.
struct HDotArgs {
    int64_t nr, nrb, neb, nrblk;
    const int64_t *eoff, *boff_e, *cptr, *coff, *boff_r;   // 0-based
    const int32_t *de, *c_r, *rd;
    const RGBlock *rblk;                                    // the stored blocks of R (dst, column stride, dims)
    const int32_t *rp, *rq;                                 // their block row and column
    int dense;
    const int64_t *rdg;                                     // per retained block: its diagonal block in rblk, or -1
    const int64_t *cro;                                     // per coupling: its retained block's offset in x_r
    const int32_t *clen;                                    // per coupling: d_e d_r, its block's length in C_data
};

// HD_LANES lanes per eliminated block (several per wavefront, 256 / HD_LANES per workgroup): x_e' E y_e + sum over its couplings c and
// rows ir of x_r[ir] (c_ir . y_e) + y_r[ir] (c_ir . x_e), c_ir column ir of C_er (schur_bilinear_eliminated,
// src/linalg.jl:208-260), as one sum of per-element terms: lane q takes element q (and q + HD_LANES, ...) of E_e and of the
// couplings' blocks, which lie one after another in C_data, so the lanes read them as one contiguous run; the lanes'
// sums are then added in a fixed tree within the group
constexpr int HD_LANES = 16;
template <class T>
__global__ __launch_bounds__(256) void k_hdot_e(HDotArgs a, const double *__restrict__ x, const double *__restrict__ y,
                                                const T *__restrict__ E, const T *__restrict__ C,
                                                double *__restrict__ part) {
    const int64_t e = (int64_t)blockIdx.x * (256 / HD_LANES) + threadIdx.x / HD_LANES;
    const int lane = threadIdx.x % HD_LANES;
    double acc = 0.0;
    if (e < a.neb) {
        const int d = a.de[e];
        const double *xe = x + a.nr + a.boff_e[e], *ye = y + a.nr + a.boff_e[e];
        const T *Ee = E + a.eoff[e];
        for (int w = lane; w < d * d; w += HD_LANES) {
            const int col = w / d, row = w - col * d;
            acc += (double)Ee[w] * xe[row] * ye[col];
        }
        const int64_t c0 = a.cptr[e], c1 = a.cptr[e + 1];
        if (c1 > c0) {
            const T *span = C + a.coff[c0];
            int64_t c = c0, start = 0;   // the coupling holding element q, and where its block starts in the span
            int len = a.clen[c0];
            for (int64_t q = lane;; q += HD_LANES) {
                while (c < c1 && q >= start + len) {
                    start += len;
                    if (++c < c1) len = a.clen[c];
                }
                if (c >= c1) break;
                const int w = (int)(q - start), ir = w / d, i = w - ir * d;
                const int64_t ro = a.cro[c];
                acc += (double)span[q] * (x[ro + ir] * ye[i] + y[ro + ir] * xe[i]);
            }
        }
    }
#pragma unroll
    for (int o = HD_LANES / 2; o >= 1; o >>= 1) acc += __shfl_xor(acc, o, 64);   // within the group
    if (lane == 0 && e < a.neb) part[e] = acc;
}

// the same sum for layouts of uniform blocks (d_e = DE) with the default offsets whose landmarks couple contiguous
// runs of retained blocks: a landmark's couplings are one run of values, value (i, q) — component i, row q of its
// window's scalar retained rows — at DE q + i, against x_r and y_r at the window's first row + q. 32 lanes per
// eliminated block, a lane per window row (coalesced over C, x_r and y_r), E_e by the lanes w < DE^2, the lanes'
// sums added in a fixed tree
template <int DE, class T>
__global__ __launch_bounds__(256) void k_hdot_win(int64_t neb, int64_t nr, int dr, const int64_t *__restrict__ cptr,
                                                  const int64_t *__restrict__ cro, const double *__restrict__ x,
                                                  const double *__restrict__ y, const T *__restrict__ E,
                                                  const T *__restrict__ C, double *__restrict__ part) {
    const int64_t e = (int64_t)blockIdx.x * 8 + threadIdx.x / 32;
    const int lane = threadIdx.x % 32;
    double acc = 0.0;
    if (e < neb) {
        double xe[DE], ye[DE];
#pragma unroll
        for (int i = 0; i < DE; i++) {
            xe[i] = x[nr + e * DE + i];
            ye[i] = y[nr + e * DE + i];
        }
        if (lane < DE * DE) {
            const int col = lane / DE, row = lane - col * DE;
            acc += (double)E[e * DE * DE + lane] * xe[row] * ye[col];
        }
        const int64_t c0 = cptr[e], nq = (cptr[e + 1] - c0) * dr;
        if (nq > 0) {
            const int64_t xb = cro[c0];
            const T *cp = C + c0 * DE * dr;
            for (int64_t q = lane; q < nq; q += 32) {
                const double xr = x[xb + q], yr = y[xb + q];
#pragma unroll
                for (int i = 0; i < DE; i++) acc += (double)cp[DE * q + i] * (xr * ye[i] + yr * xe[i]);
            }
        }
    }
#pragma unroll
    for (int o = 16; o >= 1; o >>= 1) acc += __shfl_xor(acc, o, 32);
    if (lane == 0 && e < neb) part[e] = acc;
}

// one lane per stored block of R (src/linalg.jl:153-190): a dense R is read in its upper triangle (each
// off-diagonal entry counted both ways, a diagonal block by its upper triangle); a sparse R's blocks whole, an
// off-diagonal block both ways
template <class T>
__global__ __launch_bounds__(256) void k_hdot_r(HDotArgs a, const double *__restrict__ x, const double *__restrict__ y,
                                                const T *__restrict__ R, double *__restrict__ part) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= a.nrblk) return;
    const RGBlock bk = a.rblk[b];
    const int p = a.rp[b], q = a.rq[b];
    const double *xp = x + a.boff_r[p], *yp = y + a.boff_r[p], *xq = x + a.boff_r[q], *yq = y + a.boff_r[q];
    const T *B = R + bk.dst;
    double acc = 0.0;
    for (int j = 0; j < bk.d2; j++)
        for (int i = 0; i < bk.d1; i++) {
            const double v = (double)B[(int64_t)j * bk.cstride + i];
            if (p != q) acc += v * (xp[i] * yq[j] + yp[i] * xq[j]);
            else if (!a.dense) acc += v * xp[i] * yp[j];
            else if (i < j) acc += v * (xp[i] * yp[j] + xp[j] * yp[i]);
            else if (i == j) acc += v * xp[i] * yp[i];
        }
    part[a.neb + b] = acc;
}

// getdiag! (src/schur_linear_systems.jl:239-270): d = [diag(R); diag(E_1); ...], one thread per block, retained
// blocks first; a retained block whose diagonal block a sparse R does not store has a zero diagonal
template <class T>
__global__ __launch_bounds__(256) void k_hdiag(HDotArgs a, const T *__restrict__ E, const T *__restrict__ R,
                                               double *__restrict__ d) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b < a.nrb) {
        const int n = a.rd[b];
        const int64_t k = a.rdg[b];
        double *o = d + a.boff_r[b];
        if (k >= 0) {
            const RGBlock bk = a.rblk[k];
            for (int i = 0; i < n; i++) o[i] = (double)R[bk.dst + (int64_t)i * bk.cstride + i];
        } else {
            for (int i = 0; i < n; i++) o[i] = 0.0;
        }
    } else if (b < a.nrb + a.neb) {
        const int64_t e = b - a.nrb;
        const int n = a.de[e];
        const T *Ee = E + a.eoff[e];
        double *o = d + a.nr + a.boff_e[e];
        for (int i = 0; i < n; i++) o[i] = (double)Ee[i * (n + 1)];
    }
}
}  // namespace

// ---------------------------------------------------------------------------------------------------------
// The handle: the plan of one fill pattern
// ---------------------------------------------------------------------------------------------------------
struct bls_asm {
    int64_t nr = 0, nrb = 0, neb = 0, ncoup = 0, nrowblocks = 0;
    int64_t E_len = 0, C_len = 0, R_len = 0;
    bool dense = false;
    RGPlan rg;
    DevArray<RGTab> d_tabs, d_tiny;
    DevArray<int64_t> d_tiny_cw;   // per tiny row block: its coupling written in place (k_asm_tiny), or -1
    int64_t ntiny = 0;
    int tiny_slot = 16;   // lanes per tiny row block: the widest of them
    DevArray<int32_t> d_ecd;
    int64_t glen = 0, gplen = 0;
    DevArray<double> d_G, d_gp;
    // destinations: E, C, R, b_r, b_e, each a block list with its contribution lists
    static constexpr int ND = 5;
    DevArray<RGBlock> d_blocks[ND];
    DevArray<int64_t> d_ptr[ND], d_contrib[ND];
    int64_t nblk[ND] = {0, 0, 0, 0, 0};
    int64_t nlight[ND] = {0, 0, 0, 0, 0};   // blocks [0, nlight) have at most LIGHT_MAX contributions
    int slot[ND] = {1, 1, 1, 1, 1};
    // dot(x, H, y): the storage layout
    DevArray<int64_t> d_eoff, d_boff_e, d_cptr, d_coff, d_boff_r;
    DevArray<int32_t> d_de, d_c_r, d_rd, d_rp, d_rq;
    DevArray<RGBlock> d_rblk;
    DevArray<int64_t> d_rdg;   // per retained block: its diagonal block in d_rblk, or -1 (getdiag!)
    DevArray<int64_t> d_cro;   // per coupling: its retained block's offset in x_r (dot)
    DevArray<int32_t> d_clen;  // per coupling: d_e d_r (dot)
    DevArray<double> d_part;
    int64_t nrblk = 0;
    // d_e and d_r when every block has one size (d_e <= 4), the storage offsets are the default ones and every
    // landmark couples a contiguous run of retained blocks (k_hdot_win), else 0
    int win_de = 0, win_dr = 0;
};

namespace {

struct Dest {   // one destination block while planning
    RGBlock blk;
    std::vector<int64_t> contrib;
    bool direct = false;   // a coupling block written in place by k_asm_tiny (no ordered sum)
};

std::string pair_name(int64_t i, int64_t j) {
    return "Blocks " + std::to_string(i) + " and " + std::to_string(j);
}

}  // namespace

extern "C" int bls_asm_create(int64_t nrblocks, int64_t neblocks, const int64_t *r_dims, const int64_t *e_dims,
                              const int64_t *boffsets_r, const int64_t *boffsets_e, const int64_t *eoffsets,
                              const int64_t *eliminated_ptr, const int64_t *entry_retained, const int64_t *coffsets,
                              const int64_t *r_colptr, const int64_t *r_rowval, const int64_t *r_offsets,
                              int64_t nrowblocks, const int64_t *rb_ptr, const int64_t *rb_blocks,
                              const int64_t *rb_rows, bls_asm **out) {
    if (!out || nrblocks < 0 || neblocks < 0 || nrowblocks < 0 || (nrblocks && (!r_dims || !boffsets_r)) ||
        (neblocks && (!e_dims || !boffsets_e || !eoffsets || !eliminated_ptr || !coffsets)) || !rb_ptr ||
        (nrowblocks && (!rb_rows || !rb_blocks)) || (r_colptr && (!r_rowval || !r_offsets)))
        return fail(BLS_ERR_ARG, "bls_asm_create: bad arguments");
    *out = nullptr;
    const int64_t nrb = nrblocks, neb = neblocks, nb = nrb + neb;
    const int64_t ncoup = neb ? eliminated_ptr[neb] - 1 : 0;
    if (ncoup && !entry_retained) return fail(BLS_ERR_ARG, "bls_asm_create: bad arguments");
    std::vector<int64_t> dim(nb), boff(nb);
    int64_t nr = 0;
    for (int64_t r = 0; r < nrb; r++) {
        dim[r] = r_dims[r];
        boff[r] = boffsets_r[r] - 1;
        nr = std::max(nr, boff[r] + dim[r]);
        if (dim[r] < 1 || dim[r] > 16) return fail(BLS_ERR_UNSUPPORTED, "retained block dimension outside 1..16");
    }
    for (int64_t e = 0; e < neb; e++) {
        dim[nrb + e] = e_dims[e];
        boff[nrb + e] = boffsets_e[e] - 1;
        if (dim[nrb + e] < 1 || dim[nrb + e] > 16) return fail(BLS_ERR_UNSUPPORTED, "eliminated block dimension outside 1..16");
    }
    auto h = new bls_asm();
    std::unique_ptr<bls_asm> guard(h);
    h->nr = nr;
    h->nrb = nrb;
    h->neb = neb;
    h->ncoup = ncoup;
    h->nrowblocks = nrowblocks;
    h->dense = r_colptr == nullptr;

    // destination blocks: E (full), couplings, R (dense upper or sparse lower; full diagonal blocks), b_r, b_e
    std::vector<Dest> dE(neb), dC(ncoup), dR, dbr(nrb), dbe(neb);
    for (int64_t e = 0; e < neb; e++) {
        const int d = (int)dim[nrb + e];
        dE[e].blk = RGBlock{eoffsets[e] - 1, d, d, d, 1, RG_FULL};
        h->E_len = std::max(h->E_len, eoffsets[e] - 1 + (int64_t)d * d);
        dbe[e].blk = RGBlock{boff[nrb + e], 1, 1, d, 0, 0};
        for (int64_t c = eliminated_ptr[e] - 1; c < eliminated_ptr[e + 1] - 1; c++) {
            const int64_t r = entry_retained[c] - 1;
            if (r < 0 || r >= nrb) return fail(BLS_ERR_ARG, "bls_asm_create: entry_retained out of range");
            dC[c].blk = RGBlock{coffsets[c] - 1, d, d, (int32_t)dim[r], 0, 0};   // d_e x d_r column-major
            h->C_len = std::max(h->C_len, coffsets[c] - 1 + (int64_t)d * dim[r]);
            if (c + 1 < eliminated_ptr[e + 1] - 1 && coffsets[c + 1] != coffsets[c] + (int64_t)d * dim[r])
                return fail(BLS_ERR_ARG, "bls_asm_create: the coupling blocks of eliminated block " + std::to_string(e + 1) +
                                             " are not one after another in C_data (coffsets)");
        }
    }
    for (int64_t r = 0; r < nrb; r++) dbr[r].blk = RGBlock{boff[r], 1, 1, (int32_t)dim[r], 0, 0};
    {
        const int64_t de = neb ? e_dims[0] : 0, dr = nrb ? r_dims[0] : 0;
        bool ok = de >= 1 && de <= 4 && dr >= 1;
        for (int64_t r = 0; ok && r < nrb; r++) ok = r_dims[r] == dr && boffsets_r[r] - 1 == r * dr;
        for (int64_t e = 0; ok && e < neb; e++) {
            ok = e_dims[e] == de && eoffsets[e] - 1 == e * de * de && boffsets_e[e] - 1 == e * de;
            for (int64_t c = eliminated_ptr[e]; ok && c < eliminated_ptr[e + 1] - 1; c++)   // 1-based: c - 1 and c
                ok = entry_retained[c] == entry_retained[c - 1] + 1;
        }
        for (int64_t c = 0; ok && c < ncoup; c++) ok = coffsets[c] - 1 == c * de * dr;
        if (ok) {
            h->win_de = (int)de;
            h->win_dr = (int)dr;
        }
    }
    // R: index of the destination of retained pair (p, q)
    std::vector<int64_t> rcol;   // sparse: block column start in dR
    if (h->dense) {
        dR.resize((size_t)(nrb * (nrb + 1) / 2));
        for (int64_t q = 0; q < nrb; q++)
            for (int64_t p = 0; p <= q; p++)   // upper block (p, q): element (p row, q col) at (bo_q + col) nr + bo_p + row
                dR[q * (q + 1) / 2 + p].blk =
                    RGBlock{boff[p] + boff[q] * nr, nr, (int32_t)dim[p], (int32_t)dim[q], p == q, p == q ? RG_FULL : 0};
        h->R_len = nr * nr;
    } else {
        rcol.assign(nrb + 1, 0);
        for (int64_t q = 0; q < nrb; q++) {
            rcol[q] = (int64_t)dR.size();
            for (int64_t k = r_colptr[q] - 1; k < r_colptr[q + 1] - 1; k++) {
                const int64_t p = r_rowval[k] - 1;   // lower block (p, q), p >= q: d_p x d_q column-major
                if (p < q || p >= nrb) return fail(BLS_ERR_ARG, "bls_asm_create: R block row outside q..nrblocks");
                if (k > r_colptr[q] - 1 && p <= r_rowval[k - 1] - 1)
                    return fail(BLS_ERR_ARG, "bls_asm_create: R block rows not ascending");
                Dest d;
                d.blk = RGBlock{r_offsets[k] - 1, dim[p], (int32_t)dim[p], (int32_t)dim[q], p == q, p == q ? RG_FULL : 0};
                h->R_len = std::max(h->R_len, r_offsets[k] - 1 + dim[p] * dim[q]);
                dR.push_back(std::move(d));
            }
        }
        rcol[nrb] = (int64_t)dR.size();
    }
    auto coupling = [&](int64_t e, int64_t r) -> int64_t {   // find_coupling (src/schur_assembly.jl:94-98)
        const int64_t *lo = entry_retained + eliminated_ptr[e] - 1, *hi = entry_retained + eliminated_ptr[e + 1] - 1;
        const int64_t *it = std::lower_bound(lo, hi, r + 1);
        return (it != hi && *it == r + 1) ? (int64_t)(it - entry_retained) : -1;
    };
    auto rblock = [&](int64_t p, int64_t q) -> int64_t {   // stored R block of retained pair p != q or p == q
        if (h->dense) {
            const int64_t a = std::min(p, q), b = std::max(p, q);
            return b * (b + 1) / 2 + a;
        }
        const int64_t a = std::max(p, q), b = std::min(p, q);   // lower block (a, b)
        const int64_t *lo = r_rowval + r_colptr[b] - 1, *hi = r_rowval + r_colptr[b + 1] - 1;
        const int64_t *it = std::lower_bound(lo, hi, a + 1);
        return (it != hi && *it == a + 1) ? rcol[b] + (int64_t)(it - lo) : -1;
    };

    // row blocks: records of the Gram pass, entry words, and each pair's destination
    std::vector<RGTab> recs, tiny;
    std::vector<int32_t> ecd;
    struct Direct { int64_t tiny, c; int efirst; };   // tiny two-entry row blocks over a coupling
    std::vector<Direct> direct;
    int64_t joff = 0, roff = 0, goff = 0, eoff = 0;
    for (int64_t k = 0; k < nrowblocks; k++) {
        const int64_t u0 = rb_ptr[k] - 1, u1 = rb_ptr[k + 1] - 1, ne = u1 - u0;
        const int64_t m = rb_rows[k];
        if (ne < 0 || m < 0) return fail(BLS_ERR_ARG, "bls_asm_create: row block " + std::to_string(k + 1) + " is malformed");
        if (ne > RG_MAXE)
            return fail(BLS_ERR_UNSUPPORTED, "row block " + std::to_string(k + 1) + " touches " + std::to_string(ne) +
                                                 " blocks; the device assembly takes up to 64");
        std::vector<int64_t> blk(ne), col(ne);
        int64_t W = 0;
        for (int64_t i = 0; i < ne; i++) {
            blk[i] = rb_blocks[u0 + i] - 1;
            if (blk[i] < 0 || blk[i] >= nb)
                return fail(BLS_ERR_ARG, "row block " + std::to_string(k + 1) + ": block " +
                                             std::to_string(rb_blocks[u0 + i]) + " out of range");
            for (int64_t j = 0; j < i; j++)
                if (blk[j] == blk[i])
                    return fail(BLS_ERR_ARG, "row block " + std::to_string(k + 1) + " lists block " +
                                                 std::to_string(blk[i] + 1) + " twice");
            col[i] = W;
            W += dim[blk[i]];
        }
        if (W > RG_MAXW)
            return fail(BLS_ERR_UNSUPPORTED, "row block " + std::to_string(k + 1) +
                                                 " spans more than 1024 columns; the device assembly takes up to 1024");
        if (ne && m) {
            RGTab g;
            g.raw_off = joff;
            g.res_off = roff;
            g.goff = goff;
            g.eoff = eoff;
            g.M = (int32_t)m;
            g.top = 0;
            g.n = (int32_t)m;
            g.u0 = (int32_t)ecd.size();
            g.ne = (int32_t)ne;
            g.c0 = 0;
            g.W = (int32_t)W;
            g.pad = 0;
            (W <= TINY_W && m <= TINY_M ? tiny : recs).push_back(g);
            for (int64_t i = 0; i < ne; i++) ecd.push_back((int32_t)((col[i] << 16) | dim[blk[i]]));
            // pairs i1 <= i2 in column order, d1 x d2 row-major at goff + (running offset): the Gram pass's order
            int64_t off = goff;
            for (int64_t i2 = 0; i2 < ne; i2++)
                for (int64_t i1 = 0; i1 <= i2; i1++) {
                    const int64_t p = blk[i1], q = blk[i2];
                    const int64_t sz = dim[p] * dim[q];
                    if (i1 == i2) {
                        if (p >= nrb) dE[p - nrb].contrib.push_back(off);
                        else {
                            const int64_t b = rblock(p, p);
                            if (b < 0)
                                return fail(BLS_ERR_ARG, pair_name(p + 1, p + 1) + ": R stores no diagonal block");
                            dR[b].contrib.push_back(off);
                        }
                    } else if (p >= nrb && q >= nrb) {
                        return fail(BLS_ERR_ARG, pair_name(p + 1, q + 1) + " are both eliminated: the system has no storage for the pair");
                    } else if (p >= nrb || q >= nrb) {   // coupling (e, r): stored d_e x d_r, natural orientation (e, r)
                        const int64_t e = (p >= nrb ? p : q) - nrb, r = p >= nrb ? q : p;
                        const int64_t c = coupling(e, r);
                        if (c < 0)
                            return fail(BLS_ERR_ARG, pair_name(p + 1, q + 1) + " share no row block: the system has no storage for the pair");
                        dC[c].contrib.push_back(p >= nrb ? off : ~off);
                        if (ne == 2 && W <= TINY_W && m <= TINY_M)
                            direct.push_back(Direct{(int64_t)tiny.size() - 1, c, p >= nrb ? 1 : 0});
                    } else {   // retained pair: dense stores (min, max) upper, sparse (max, min) lower
                        const int64_t b = rblock(p, q);
                        if (b < 0)
                            return fail(BLS_ERR_ARG, pair_name(p + 1, q + 1) + " share no row block: the system has no storage for the pair");
                        const bool natural = h->dense ? p < q : p > q;
                        dR[b].contrib.push_back(natural ? off : ~off);
                    }
                    off += sz;
                }
            for (int64_t i = 0; i < ne; i++) {
                const int64_t p = blk[i];
                (p >= nrb ? dbe[p - nrb] : dbr[p]).contrib.push_back(eoff + col[i]);
            }
            goff = off;
            eoff += W;
        }
        joff += m * W;
        roff += m;
    }
    std::vector<int64_t> tcw(tiny.size(), -1);
    for (const Direct &d : direct)
        if (dC[d.c].contrib.size() == 1) {
            tcw[d.tiny] = 2 * (coffsets[d.c] - 1) + d.efirst;
            dC[d.c].direct = true;
        }
    std::vector<RGTab> tabs;
    int rc;
    if ((rc = rg_classify(recs, tabs, h->rg))) return rc;
    if ((rc = h->d_tabs.upload(tabs)) || (rc = h->d_tiny.upload(tiny)) || (rc = h->d_ecd.upload(ecd)) ||
        (rc = h->d_tiny_cw.upload(tcw)))
        return rc;
    h->ntiny = (int64_t)tiny.size();
    h->tiny_slot = 1;
    for (const RGTab &g : tiny) h->tiny_slot = std::max(h->tiny_slot, (int)g.W);
    h->rg.tabs = h->d_tabs.ptr;
    h->rg.ecd = h->d_ecd.ptr;
    if ((rc = h->d_G.alloc((size_t)std::max<int64_t>(1, goff))) || (rc = h->d_gp.alloc((size_t)std::max<int64_t>(1, eoff))))
        return rc;
    h->glen = goff;
    h->gplen = eoff;
    {   // the storage layout for dot(x, H, y)
        std::vector<int64_t> eo(neb), be(neb), cp(neb + 1, 0), co(ncoup), br(nrb);
        std::vector<int32_t> dd(neb), cr(ncoup), rdv(nrb), rp, rq;
        std::vector<int64_t> rdg(nrb, -1);
        std::vector<RGBlock> rb;
        for (int64_t e = 0; e < neb; e++) {
            eo[e] = eoffsets[e] - 1;
            be[e] = boff[nrb + e];
            dd[e] = (int32_t)dim[nrb + e];
            cp[e + 1] = eliminated_ptr[e + 1] - 1;
        }
        std::vector<int64_t> cro(ncoup);
        std::vector<int32_t> clen(ncoup);
        for (int64_t e = 0; e < neb; e++)
            for (int64_t c = eliminated_ptr[e] - 1; c < eliminated_ptr[e + 1] - 1; c++) {
                co[c] = coffsets[c] - 1;
                cr[c] = (int32_t)(entry_retained[c] - 1);
                cro[c] = boff[cr[c]];
                clen[c] = (int32_t)(dim[nrb + e] * dim[cr[c]]);
            }
        for (int64_t r = 0; r < nrb; r++) {
            br[r] = boff[r];
            rdv[r] = (int32_t)dim[r];
        }
        if (h->dense) {
            for (int64_t q = 0; q < nrb; q++)
                for (int64_t p = 0; p <= q; p++) {
                    if (p == q) rdg[q] = (int64_t)rb.size();
                    rb.push_back(dR[q * (q + 1) / 2 + p].blk);
                    rp.push_back((int32_t)p);
                    rq.push_back((int32_t)q);
                }
        } else {
            for (int64_t q = 0; q < nrb; q++)
                for (int64_t k = r_colptr[q] - 1; k < r_colptr[q + 1] - 1; k++) {
                    if (r_rowval[k] - 1 == q) rdg[q] = (int64_t)rb.size();
                    rb.push_back(dR[rcol[q] + (k - (r_colptr[q] - 1))].blk);
                    rp.push_back((int32_t)(r_rowval[k] - 1));
                    rq.push_back((int32_t)q);
                }
        }
        int rc1;
        if ((rc1 = h->d_eoff.upload(eo)) || (rc1 = h->d_boff_e.upload(be)) || (rc1 = h->d_cptr.upload(cp)) ||
            (rc1 = h->d_coff.upload(co)) || (rc1 = h->d_boff_r.upload(br)) || (rc1 = h->d_de.upload(dd)) ||
            (rc1 = h->d_c_r.upload(cr)) || (rc1 = h->d_rd.upload(rdv)) || (rc1 = h->d_rp.upload(rp)) ||
            (rc1 = h->d_rq.upload(rq)) || (rc1 = h->d_rblk.upload(rb)) || (rc1 = h->d_rdg.upload(rdg)) ||
            (rc1 = h->d_cro.upload(cro)) || (rc1 = h->d_clen.upload(clen)) ||
            (rc1 = h->d_part.alloc((size_t)(neb + (int64_t)rb.size() + FSUM_CHUNKS))))
            return rc1;
        h->nrblk = (int64_t)rb.size();
    }
    std::vector<Dest> *lists[bls_asm::ND] = {&dE, &dC, &dR, &dbr, &dbe};
    for (int l = 0; l < bls_asm::ND; l++) {
        // light blocks first, then the rest; each keeps its contributions in row-block order
        std::vector<RGBlock> blocks;
        std::vector<int64_t> ptr(1, 0), contrib;
        int maxsz = 1;
        for (int pass = 0; pass < 2; pass++)
            for (const Dest &d : *lists[l]) {
                if (d.direct || (d.contrib.size() <= (size_t)LIGHT_MAX) != (pass == 0)) continue;
                blocks.push_back(d.blk);
                contrib.insert(contrib.end(), d.contrib.begin(), d.contrib.end());
                ptr.push_back((int64_t)contrib.size());
                if (pass == 0) maxsz = std::max(maxsz, d.blk.d1 * d.blk.d2);
            }
        int64_t nl = 0;
        for (const Dest &d : *lists[l]) nl += !d.direct && d.contrib.size() <= (size_t)LIGHT_MAX;
        h->nlight[l] = nl;
        h->slot[l] = std::min(maxsz, 64);
        if ((rc = h->d_blocks[l].upload(blocks)) || (rc = h->d_ptr[l].upload(ptr)) || (rc = h->d_contrib[l].upload(contrib)))
            return rc;
        h->nblk[l] = (int64_t)blocks.size();
    }
    *out = guard.release();
    return BLS_OK;
}

extern "C" void bls_asm_destroy(bls_asm *a) { delete a; }

extern "C" int bls_asm_lengths(const bls_asm *a, int64_t *lengths) {
    if (!a || !lengths) return fail(BLS_ERR_ARG, "bls_asm_lengths: bad arguments");
    lengths[0] = a->E_len;
    lengths[1] = a->C_len;
    lengths[2] = a->R_len;
    lengths[3] = a->nr;
    return BLS_OK;
}

extern "C" int bls_asm_hessian_f64(bls_asm *a, const double *J, const double *r, double *E_data, double *C_data,
                                   double *R, double *b_r, double *b_e, void *stream) {
    if (!a) return fail(BLS_ERR_ARG, "null handle");
    if ((a->glen && (!J || !r)) || (a->neb && (!E_data || !b_e)) || (a->ncoup && !C_data) || (a->nrb && (!R || !b_r)))
        return fail(BLS_ERR_ARG, "bls_asm_hessian_f64: a required array is NULL");
    hipStream_t s = (hipStream_t)stream;
    int rc;
    if (a->dense && a->R_len)
        k_rg_fill<double><<<(unsigned)std::min<int64_t>((a->R_len + 255) / 256, 4096), 256, 0, s>>>(R, a->R_len);
    if ((rc = rg_launch<double>(a->rg, J, r, a->d_G.ptr, a->d_gp.ptr, s))) return rc;
    if (a->ntiny) {
        const int64_t per_wg = 4 * (64 / a->tiny_slot);
        k_asm_tiny<double><<<(unsigned)((a->ntiny + per_wg - 1) / per_wg), 256, 0, s>>>(
            a->d_tiny.ptr, a->ntiny, a->tiny_slot, a->d_ecd.ptr, J, r, a->d_G.ptr, a->d_gp.ptr, C_data,
            a->d_tiny_cw.ptr);
    }
    double *dst[bls_asm::ND] = {E_data, C_data, R, b_r, b_e};
    for (int l = 0; l < bls_asm::ND; l++) {
        if (!a->nblk[l]) continue;
        const double *src = l < 3 ? a->d_G.ptr : a->d_gp.ptr;
        const int64_t nl = a->nlight[l], nh = a->nblk[l] - nl;
        if (nl) {
            const int64_t per_wg = 4 * (64 / a->slot[l]);
            k_red_gram_sum_light<double><<<(unsigned)((nl + per_wg - 1) / per_wg), 256, 0, s>>>(
                nl, a->slot[l], a->d_blocks[l].ptr, a->d_ptr[l].ptr, a->d_contrib[l].ptr, src, dst[l]);
        }
        if (nh)
            k_red_gram_sum<double><<<(unsigned)nh, 256, 0, s>>>(nh, a->d_blocks[l].ptr + nl, a->d_ptr[l].ptr + nl,
                                                                 a->d_contrib[l].ptr, src, dst[l]);
    }
    BLS_HIP(hipGetLastError());
    return BLS_OK;
}

namespace {
HDotArgs hdot_args(const bls_asm *a) {
    HDotArgs d{};
    d.nr = a->nr;
    d.nrb = a->nrb;
    d.neb = a->neb;
    d.nrblk = a->nrblk;
    d.eoff = a->d_eoff.ptr;
    d.boff_e = a->d_boff_e.ptr;
    d.cptr = a->d_cptr.ptr;
    d.coff = a->d_coff.ptr;
    d.boff_r = a->d_boff_r.ptr;
    d.de = a->d_de.ptr;
    d.c_r = a->d_c_r.ptr;
    d.rd = a->d_rd.ptr;
    d.rblk = a->d_rblk.ptr;
    d.rp = a->d_rp.ptr;
    d.rq = a->d_rq.ptr;
    d.dense = a->dense;
    d.rdg = a->d_rdg.ptr;
    d.cro = a->d_cro.ptr;
    d.clen = a->d_clen.ptr;
    return d;
}
}  // namespace

namespace {
template <class T>
int hessian_dot(bls_asm *a, const double *x, const double *y, const T *E_data, const T *C_data, const T *R,
                double *result, hipStream_t s) {
    if (!a) return fail(BLS_ERR_ARG, "null handle");
    if (!x || !y || !result || (a->neb && !E_data) || (a->ncoup && !C_data) || (a->nrblk && !R))
        return fail(BLS_ERR_ARG, "dot(x, H, y): a required array is NULL");
    const HDotArgs d = hdot_args(a);
    constexpr int per_wg = 256 / HD_LANES;
    if (a->neb && a->win_de > 0) {
        const unsigned grid = (unsigned)((a->neb + 7) / 8);
        auto go = [&](auto de_c) {
            constexpr int DE = decltype(de_c)::value;
            k_hdot_win<DE, T><<<grid, 256, 0, s>>>(a->neb, a->nr, a->win_dr, a->d_cptr.ptr, a->d_cro.ptr, x, y, E_data,
                                                   C_data, a->d_part.ptr);
        };
        switch (a->win_de) {
        case 1: go(std::integral_constant<int, 1>{}); break;
        case 2: go(std::integral_constant<int, 2>{}); break;
        case 3: go(std::integral_constant<int, 3>{}); break;
        default: go(std::integral_constant<int, 4>{}); break;
        }
    } else if (a->neb) {
        k_hdot_e<T><<<(unsigned)((a->neb + per_wg - 1) / per_wg), 256, 0, s>>>(d, x, y, E_data, C_data, a->d_part.ptr);
    }
    if (a->nrblk) k_hdot_r<T><<<(unsigned)((a->nrblk + 255) / 256), 256, 0, s>>>(d, x, y, R, a->d_part.ptr);
    fsum_launch(a->neb + a->nrblk, a->d_part.ptr, a->d_part.ptr + a->neb + a->nrblk, result, s);   // FSUM_CHUNKS slots after the partials
    BLS_HIP(hipGetLastError());
    return BLS_OK;
}

template <class T> int hessian_diag(bls_asm *a, const T *E_data, const T *R, double *d, hipStream_t s) {
    if (!a) return fail(BLS_ERR_ARG, "null handle");
    if (!d || (a->neb && !E_data) || (a->nrblk && !R)) return fail(BLS_ERR_ARG, "getdiag!: a required array is NULL");
    const int64_t nb = a->nrb + a->neb;
    if (nb) k_hdiag<T><<<grid_for(nb, 256), 256, 0, s>>>(hdot_args(a), E_data, R, d);
    BLS_HIP(hipGetLastError());
    return BLS_OK;
}
}  // namespace

extern "C" int bls_asm_dot_f64(bls_asm *a, const double *x, const double *y, const double *E_data, const double *C_data,
                               const double *R, double *result, void *stream) {
    return hessian_dot<double>(a, x, y, E_data, C_data, R, result, (hipStream_t)stream);
}
extern "C" int bls_asm_dot_f32(bls_asm *a, const double *x, const double *y, const float *E_data, const float *C_data,
                               const float *R, double *result, void *stream) {
    return hessian_dot<float>(a, x, y, E_data, C_data, R, result, (hipStream_t)stream);
}
extern "C" int bls_asm_diag_f64(bls_asm *a, const double *E_data, const double *R, double *d, void *stream) {
    return hessian_diag<double>(a, E_data, R, d, (hipStream_t)stream);
}
extern "C" int bls_asm_diag_f32(bls_asm *a, const float *E_data, const float *R, double *d, void *stream) {
    return hessian_diag<float>(a, E_data, R, d, (hipStream_t)stream);
}
