// The Gram pass: for many small column-major "tables" X_t (n_t rows; entries = contiguous column groups, ascending),
// every pair block X_t(:, cols(u1))' X_t(:, cols(u2)) (u1 <= u2) on the FP64 matrix instruction into a contribution
// buffer, and the per-entry partials X_t(:, cols(u))' r_t; then fixed-order sums of the contributions of each
// destination block. Shared by reduced_normal! (tables of the reduced Jacobian) and the assembly of a Schur Hessian
// from row blocks (addblock!). Included inside each unit's anonymous namespace.
#pragma once

constexpr int RG_MAXE = 64;    // entries of one table (one wavefront scans their dims)
constexpr int RG_MAXW = 1024;  // columns spanned by the entries of one table
constexpr int RG_STATIC_LDS = 6144;   // bytes of static LDS of the Gram pass (its per-column words)
static_assert(RG_MAXW * (sizeof(int16_t) + sizeof(int)) == RG_STATIC_LDS, "static LDS of the Gram pass");

typedef double rg_v4d __attribute__((ext_vector_type(4)));

// Launch classes by the size of a table's staged block (rows rounded up to 4, columns to 16, one pad column): up to
// 4096 doubles (four workgroups per CU), up to 9472 (two: 2 x (9472 x 8 + 6144) bytes fit 160 KB), up to the
// device's LDS per workgroup (one), and the rest, which read their operands from global memory.
constexpr int RG_CLASS_CAP0 = 4096, RG_CLASS_CAP1 = 9472, RG_CLASS_CAP2 = 19456;
static_assert(16 * 256 >= RG_CLASS_CAP0 && 19 * 512 >= RG_CLASS_CAP1 && 19 * 1024 >= RG_CLASS_CAP2, "prefetch covers the classes");

// One table of the Gram pass, in launch order (built on the host): where its block and residual lie, where its pair
// blocks and per-entry partials go, its reduced rows n and the W columns its entries span from column c0.
struct RGTab {
    int64_t raw_off, res_off, goff, eoff;
    int32_t M, top, n, u0, ne, c0, W, pad;
};

__device__ inline RGTab rg_load(const RGTab *p) {   // field by field, so the record stays in registers
    RGTab g;
    g.raw_off = p->raw_off;
    g.res_off = p->res_off;
    g.goff = p->goff;
    g.eoff = p->eoff;
    g.M = p->M;
    g.top = p->top;
    g.n = p->n;
    g.u0 = p->u0;
    g.ne = p->ne;
    g.c0 = p->c0;
    g.W = p->W;
    g.pad = 0;
    return g;
}

// X = the reduced rows of table t over the columns of its entries (n x W, column c of the table at c0 + c). Each
// wavefront forms 16 x 16 tiles (I <= J) of X'X with the FP64 matrix instruction (16x16x4: lane l holds A(l % 16,
// l / 16), B(l / 16, l % 16), and D(l / 16 + 4 v, l % 16) in accumulator v), from X staged row-major in LDS when it
// fits, and writes each element (a, b), a <= b, to its pair block (e1, e2): the pairs of the table's entries i1 <= i2
// in the order i2 (i2 + 1) / 2 + i1, each d1 x d2 row-major, at offset Q[e2] + d2 P[e1] with P the prefix sums of
// the entry dims and Q[e] = sum_{j < e} d_j P[j + 1]. Lower entries of a diagonal block (row > col) are not written.
// Persistent: workgroup b takes tables tabs[b], tabs[b + grid], ...; while it works on one table, the next one's
// staged block (X row-major with row stride SW = Wp + 1, r~ in column Wp, zeros elsewhere; PF values per thread,
// PF NT >= stage_cap) and entry words (ecd[u] = (entry column - c0) << 16 | entry dim) are loading into registers, and
// the record of the one after it is loading too.
// at least 4 waves per SIMD: 16 wavefronts per CU in every class (4, 2 and 1 workgroups of 4, 8 and 16)
template <class T, int NT, int PF>
__global__ __launch_bounds__(NT, 4) void k_red_gram(const RGTab *__restrict__ tabs, int64_t cnt,
                                                 const int32_t *__restrict__ ecd, int stage_cap, const T *__restrict__ F,
                                                 const double *__restrict__ Fr, double *__restrict__ G,
                                                 double *__restrict__ gp) {
    extern __shared__ double XL[];
    // per staged column c of the current table (-1 where no entry holds it): s_colA[c] = P[e] + j, its column in
    // the entries' packed order (entry e, column j of it), and s_colB[c] = (Q[e] + j) << 5 | d_e, so that element
    // (a, b) of X'X goes to pair-block offset (s_colB[b] >> 5) + (s_colB[b] & 31) s_colA[a]
    __shared__ int16_t s_colA[RG_MAXW];
    __shared__ int s_colB[RG_MAXW];
    // the next table's staged values as loaded: Float32 tables keep theirs in Float32 registers, widened as they go
    // into LDS, and a lane's one residual value (the residual column holds np4 <= 64 < NT elements) in pres at slot
    // jres, so that no register a pending load writes is reused and no wait for the prefetch sits in its issue
    std::conditional_t<sizeof(T) == 8, double, float> pf[PF];
    double pres = 0.0;
    int jres = -1;
    int my_ecd = 0;
    auto prefetch = [&](const RGTab g) {
        my_ecd = (int)threadIdx.x < g.ne ? ecd[g.u0 + threadIdx.x] : 0;
        jres = -1;
        const int Wp = (g.W + 15) & ~15, np4 = (g.n + 3) & ~3, SW = Wp + 1;
        if ((int64_t)np4 * SW > stage_cap) return;
        const T *Xg = F + g.raw_off + g.top;
        const double *rg = Fr + g.res_off + g.top;
        // element q = threadIdx.x + j NT is (k, c) = (q % np4, q / np4), walked without divisions
        const int dc = NT / np4, dk = NT - dc * np4;
        int c = threadIdx.x / np4, k = threadIdx.x - c * np4;
#pragma unroll
        for (int j = 0; j < PF; j++) {
            T v = (T)0;
            if (c < SW && k < g.n) {
                if (c < g.W) {
                    v = Xg[(int64_t)(g.c0 + c) * g.M + k];
                } else if (c == Wp) {
                    if constexpr (sizeof(T) == 8) {
                        v = rg[k];
                    } else {
                        pres = rg[k];
                        jres = j;
                    }
                }
            }
            pf[j] = v;
            c += dc;
            k += dk;
            if (k >= np4) {
                k -= np4;
                c++;
            }
        }
    };
    const int64_t grid = gridDim.x;
    RGTab cur = rg_load(tabs + blockIdx.x);   // the grid is at most cnt
    RGTab nxt = rg_load(tabs + (blockIdx.x + grid < cnt ? blockIdx.x + grid : blockIdx.x));
    prefetch(cur);
    for (int64_t it = blockIdx.x; it < cnt; it += grid) {
        __syncthreads();   // the previous table is done with XL and the maps
        const int M = cur.M, n = cur.n, ne = cur.ne, c0 = cur.c0, W = cur.W;
        const int Wp = (W + 15) & ~15, np4 = (n + 3) & ~3, SW = Wp + 1;
        const bool fits = (int64_t)np4 * SW <= stage_cap;
        const T *Xg = F + cur.raw_off + cur.top;        // X(k, c) = Xg[(c0 + c) M + k]
        const double *rg = Fr + cur.res_off + cur.top;  // r~(k)
        double *Gt = G + cur.goff;
        double *gpt = gp + cur.eoff;
        if (fits) {
            const int dc = NT / np4, dk = NT - dc * np4;
            int c = threadIdx.x / np4, k = threadIdx.x - c * np4;
#pragma unroll
            for (int j = 0; j < PF; j++) {
                if (c < SW) {
                    double v = (double)pf[j];
                    if constexpr (sizeof(T) == 4)
                        if (j == jres) v = pres;
                    XL[k * SW + c] = v;
                }
                c += dc;
                k += dk;
                if (k >= np4) {
                    k -= np4;
                    c++;
                }
            }
        }
        for (int i = threadIdx.x; i < Wp; i += NT) {
            s_colA[i] = -1;
            s_colB[i] = -1;
        }
        const int e_cst = my_ecd >> 16, e_dim = my_ecd & 0xffff;   // entry threadIdx.x (< ne)
        int e_P = 0, e_Q = 0;
        if (threadIdx.x < 64) {   // prefix sums of the entry dims: P (columns) and Q (pair-block offsets)
            const int e = threadIdx.x;
            const int d = e < ne ? e_dim : 0;
            int p = d;
            for (int o = 1; o < 64; o <<= 1) {
                const int v = __shfl_up(p, o);
                if (e >= o) p += v;
            }
            const int q0 = d * p;   // d_e P[e + 1]
            int q = q0;
            for (int o = 1; o < 64; o <<= 1) {
                const int v = __shfl_up(q, o);
                if (e >= o) q += v;
            }
            e_P = p - d;
            e_Q = q - q0;
        }
        __syncthreads();
        if ((int)threadIdx.x < ne)
            for (int j = 0; j < e_dim; j++) {
                s_colA[e_cst + j] = (int16_t)(e_P + j);
                s_colB[e_cst + j] = ((e_Q + j) << 5) | e_dim;
            }
        __syncthreads();
        if (it + grid < cnt) {   // the next table's loads go out before this one's work
            cur = nxt;
            prefetch(cur);
            if (it + 2 * grid < cnt) nxt = rg_load(tabs + it + 2 * grid);
        }
        const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
        // writes tile (I, J) of X'X: element (a, b), a <= b, to its pair block
        auto store = [&](const rg_v4d &acc, int I, int J) {
            const int b = 16 * J + li;
            const int cb = b < W ? s_colB[b] : -1;
            if (cb < 0) return;
            const int baseB = cb >> 5, d2 = cb & 31;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int aa = 16 * I + lk + 4 * v;
                const int pa = aa <= b ? s_colA[aa] : -1;
                if (pa >= 0) Gt[baseB + d2 * pa] = acc[v];
            }
        };
        // work items, dealt to the wavefronts in turn: column strip J of the upper triangle in pairs of tiles
        // (I0, I0 + 1) that share B, then the per-entry partials of J~' r~ (as k_red_tmul_part writes them), 64
        // columns per item, one lane per column -- last, so that they go to the wavefronts with one tile item fewer
        const int ngp = (W + 63) >> 6;
        const int ntl = Wp >> 4;
        const int ntitems = (ntl / 2 + 1) * (ntl / 2) + ((ntl & 1) ? (ntl + 1) / 2 : 0);
        const int nitems = ntitems + ngp;
        for (int item = wave; item < nitems; item += NT / 64) {
            if (item >= ntitems) {
                const int c = 64 * (item - ntitems) + lane;
                const int pc = c < W ? s_colA[c] : -1;
                if (pc >= 0) {
                    double acc = 0.0;
                    if (fits) {
#pragma unroll 8
                        for (int k = 0; k < n; k++) acc += XL[k * SW + c] * XL[k * SW + Wp];
                    } else {
                        const T *x = Xg + (int64_t)(c0 + c) * M;
                        for (int k = 0; k < n; k++) acc += (double)x[k] * rg[k];
                    }
                    gpt[pc] = acc;
                }
                continue;
            }
            const int titem = item;
            int J = 0, base = 0;
            while (base + (J + 2) / 2 <= titem) {
                base += (J + 2) / 2;
                J++;
            }
            const int I0 = 2 * (titem - base);
            const bool two = I0 + 1 <= J;
            rg_v4d acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
            if (fits) {
                const double *pb = XL + lk * SW + 16 * J + li, *pa = XL + lk * SW + 16 * I0 + li;
                if (two) {
#pragma unroll 4
                    for (int k0 = 0; k0 < np4; k0 += 4) {
                        const double bv = pb[k0 * SW], a0 = pa[k0 * SW], a1 = pa[k0 * SW + 16];
                        acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, bv, acc0, 0, 0, 0);
                        acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, bv, acc1, 0, 0, 0);
                    }
                } else {
#pragma unroll 4
                    for (int k0 = 0; k0 < np4; k0 += 4)
                        acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(pa[k0 * SW], pb[k0 * SW], acc0, 0, 0, 0);
                }
            } else {
                const bool vb = 16 * J + li < W, va0 = 16 * I0 + li < W, va1 = two && 16 * (I0 + 1) + li < W;
                const T *xb = Xg + (int64_t)(c0 + 16 * J + li) * M, *xa = Xg + (int64_t)(c0 + 16 * I0 + li) * M;
                for (int k0 = 0; k0 < np4; k0 += 4) {
                    const int k = k0 + lk;
                    const double bv = (vb && k < n) ? (double)xb[k] : 0.0;
                    const double a0 = (va0 && k < n) ? (double)xa[k] : 0.0;
                    const double a1 = (va1 && k < n) ? (double)xa[(int64_t)16 * M + k] : 0.0;
                    acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, bv, acc0, 0, 0, 0);
                    if (two) acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, bv, acc1, 0, 0, 0);
                }
            }
            store(acc0, I0, J);
            if (two) store(acc1, I0 + 1, J);
        }
    }
}

struct RGBlock {
    int64_t dst;       // S index of entry (0, 0) of the block
    int64_t cstride;   // distance between its columns in S
    int32_t d1, d2, diag, flags;   // flags: RG_TRANSPOSED, RG_FULL
};
constexpr int RG_TRANSPOSED = 1;   // entry (row, col) of the contributions goes to S[dst + row cstride + col]
constexpr int RG_FULL = 2;         // a diagonal block is written whole: its lower triangle mirrors the upper

// One workgroup per block of S: entry w = row d2 + col of the block is the sum of its contributions in ascending
// table order. A contribution c >= 0 is a d1 x d2 row-major pair block of the Gram buffer at c; ~c (negative) is one
// stored the other way round, d2 x d1 row-major at c (a table that holds the two blocks in the opposite column
// order). The sum is written at S[dst + col cstride + row] (RG_TRANSPOSED:
// S[dst + row cstride + col]). Below the diagonal of a diagonal block there are no contributions: zero, or with
// RG_FULL the mirror of the entry above. Thread (w, run) = (tid % sz, tid / sz) adds the run-th of 256 / sz
// contiguous runs of the block's contribution list, and the runs are then added in order.
template <class T>
__global__ __launch_bounds__(256) void k_red_gram_sum(int64_t nblk, const RGBlock *__restrict__ blocks,
                                                      const int64_t *__restrict__ ptr, const int64_t *__restrict__ contrib,
                                                      const double *__restrict__ G, T *__restrict__ S) {
    __shared__ double part[256];
    const int64_t b = blockIdx.x;
    const RGBlock bk = blocks[b];
    const int sz = bk.d1 * bk.d2;   // <= 256
    const int nrun = 256 / sz;
    const int w = threadIdx.x % sz, run = threadIdx.x / sz;
    const int row = w / bk.d2, col = w - row * bk.d2;
    const bool lower = bk.diag && row > col;
    double acc = 0.0;
    if (run < nrun && !lower) {
        const int64_t cb = ptr[b], L = ptr[b + 1] - cb;
        const int64_t len = (L + nrun - 1) / nrun;
        const int64_t kb = cb + std::min<int64_t>(L, run * len), ke = cb + std::min<int64_t>(L, (run + 1) * len);
        int64_t c = kb;
        // eight contributions per round; the next round's indices load while this round's values are in flight
        const int wt = col * bk.d1 + row;   // the same entry in a contribution stored the other way round
        auto at = [&](int64_t ci) { return ci >= 0 ? ci + w : ~ci + wt; };
        int64_t idx[8];
        if (c + 8 <= ke) {
#pragma unroll
            for (int j = 0; j < 8; j++) idx[j] = contrib[c + j];
        }
        for (; c + 8 <= ke; c += 8) {
            double v[8];
#pragma unroll
            for (int j = 0; j < 8; j++) v[j] = G[at(idx[j])];
            if (c + 16 <= ke) {
#pragma unroll
                for (int j = 0; j < 8; j++) idx[j] = contrib[c + 8 + j];
            }
#pragma unroll
            for (int j = 0; j < 8; j++) acc += v[j];
        }
        for (; c < ke; c++) acc += G[at(contrib[c])];
    }
    part[threadIdx.x] = acc;
    __syncthreads();
    if ((int)threadIdx.x < sz) {
        double sum = 0.0;
        for (int q = 0; q < nrun; q++) sum += part[q * sz + threadIdx.x];
        const bool tr = bk.flags & RG_TRANSPOSED, full = bk.flags & RG_FULL;
        const int64_t at = tr ? bk.dst + (int64_t)row * bk.cstride + col : bk.dst + (int64_t)col * bk.cstride + row;
        if (!(lower && full)) S[at] = (T)sum;
        if (full && bk.diag && row < col) S[bk.dst + (int64_t)row * bk.cstride + col] = (T)sum;
    }
}

template <class T> __global__ void k_rg_fill(T *p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = (T)0;
}

// Where the tables of a Gram pass go: per-table records in launch order (classes 0..3 at cls_off[c] ..
// cls_off[c + 1]), the entry words, and the device's limits.
struct RGPlan {
    const RGTab *tabs = nullptr;
    const int32_t *ecd = nullptr;
    int64_t cls_off[5] = {0, 0, 0, 0, 0};
    int cap2 = 0, ncu = 1, lds_cu = 0;
};

// Orders records into the four launch classes (staged doubles: rows rounded up to 4 times columns rounded up to 16
// plus one) and reads the device's LDS limits; the caller uploads `ordered`.
inline int rg_classify(const std::vector<RGTab> &recs, std::vector<RGTab> &ordered, RGPlan &plan) {
    int dev = 0, lds = 0;
    BLS_HIP(hipGetDevice(&dev));
    BLS_HIP(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    BLS_HIP(hipDeviceGetAttribute(&plan.ncu, hipDeviceAttributeMultiprocessorCount, dev));
    BLS_HIP(hipDeviceGetAttribute(&plan.lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev));
    plan.cap2 = (int)std::min<int64_t>(RG_CLASS_CAP2, std::max<int64_t>(0, (lds - RG_STATIC_LDS) / 8));
    const int64_t caps[3] = {RG_CLASS_CAP0, RG_CLASS_CAP1, plan.cap2};
    std::vector<RGTab> cls[4];
    for (const RGTab &g : recs) {
        const int64_t st = ((int64_t)(g.n + 3) & ~3LL) * ((((int64_t)g.W + 15) & ~15LL) + 1);
        int c = 3;
        for (int k = 2; k >= 0; k--)
            if (st <= caps[k]) c = k;
        cls[c].push_back(g);
    }
    ordered.clear();
    for (int c = 0; c < 4; c++) {
        plan.cls_off[c] = (int64_t)ordered.size();
        ordered.insert(ordered.end(), cls[c].begin(), cls[c].end());
    }
    plan.cls_off[4] = (int64_t)ordered.size();
    return BLS_OK;
}

// One launch per non-empty class: persistent grids of as many workgroups as the compute units hold at once.
template <class T>
int rg_launch(const RGPlan &plan, const T *F, const double *Fr, double *G, double *gp, hipStream_t s) {
    const int caps[4] = {RG_CLASS_CAP0, RG_CLASS_CAP1, plan.cap2, 0};
    for (int c = 0; c < 4; c++) {
        const int64_t cnt = plan.cls_off[c + 1] - plan.cls_off[c];
        if (!cnt) continue;
        const RGTab *tabs = plan.tabs + plan.cls_off[c];
        const size_t lds = (size_t)caps[c] * sizeof(double);
        const int nt = c == 2 ? 1024 : c == 1 ? 512 : 256;
        const int64_t per_cu =
            std::max<int64_t>(1, std::min<int64_t>(plan.lds_cu / (int64_t)(lds + RG_STATIC_LDS), 32 / (nt / 64)));
        const unsigned grid = (unsigned)std::min<int64_t>(cnt, per_cu * std::max(1, plan.ncu));
        if (c == 0 || c == 3) {
            BLS_HIP(hipFuncSetAttribute((const void *)k_red_gram<T, 256, 16>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
            k_red_gram<T, 256, 16><<<grid, 256, lds, s>>>(tabs, cnt, plan.ecd, caps[c], F, Fr, G, gp);
        } else if (c == 1) {
            BLS_HIP(hipFuncSetAttribute((const void *)k_red_gram<T, 512, 19>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
            k_red_gram<T, 512, 19><<<grid, 512, lds, s>>>(tabs, cnt, plan.ecd, caps[c], F, Fr, G, gp);
        } else {
            BLS_HIP(hipFuncSetAttribute((const void *)k_red_gram<T, 1024, 19>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
            k_red_gram<T, 1024, 19><<<grid, 1024, lds, s>>>(tabs, cnt, plan.ecd, caps[c], F, Fr, G, gp);
        }
    }
    BLS_HIP(hipGetLastError());
    return BLS_OK;
}
