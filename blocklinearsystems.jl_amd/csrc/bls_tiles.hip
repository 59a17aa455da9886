// Phase 2 of the Schur reduction on the FP64 matrix cores, for structures whose landmarks couple contiguous windows
// of retained blocks (every block of uniform size d_r >= 6, every landmark block of uniform size d_e <= 4).
//
// S is cut into 16 x 16 scalar tiles. A workgroup owns a run of up to four tiles of one tile row, (ti, tj .. tj+3),
// tj >= ti, and sums into each tile, for every landmark e that couples a block of tile row ti and a block of that
// tile's column,
//     D[i][j] += sum_k W1[k][i] C2[k][j],   W1 = E_inv_e C_e(row i),  C2 = C_e(column j)
// which is the reference's S[r1, r2] -= W_{c1}' C_{c2} (src/schur_reduction.jl:315-380) for the block pairs the tile
// covers, as a v_mfma_f64_16x16x4f64 per landmark and tile (K = d_e components, padded to 4). The tile's 16 rows and
// columns are scalar rows and columns, so a tile spans parts of up to four blocks on each side; a landmark adds zero
// where it does not couple a block (buffer loads outside its window return 0).
//
// The landmarks of tile row ti are one list, sorted by the last tile column they reach (descending, then landmark
// order): the landmarks reaching tile column tj + t are a prefix of that list, and the run's list splits into
// segments of landmarks reaching 4, 3, 2 and 1 of its tiles. Per landmark a lane loads one value of the row side
// (component k of row i's coupling) once for all the tiles it reaches, and W1 = E_inv C1 is one 16 x 16 x 4 product
// (A' = E_inv in rows 0..d_e-1 and z in row 4, from LDS) whose output lands in the layout of the tile products' A
// operand and carries the g_S part C1' z in row 4; then per tile one column-side load and one product. E_inv and z
// come from one coalesced load per group of landmarks through LDS.
//
// The four waves of a workgroup take equal contiguous shares of each segment; their partial tiles are added in wave
// order through LDS, and wave t writes tile t. Each S entry is owned by one tile: no atomics, and the sums are taken
// in a fixed order (bitwise reproducible). g_S = b_r s_r - sum C' z is formed by the run starting at the diagonal
// tile (tj == ti), whose list is the whole row's.
#include "bls_internal.h"

#include <algorithm>
#include <cstring>
#include <type_traits>

namespace bls {

namespace {

constexpr int TW = 4;   // waves per workgroup
constexpr int NT = 4;   // tiles per run
typedef double d4 __attribute__((ext_vector_type(4)));

// landmarks per group: what one cooperative E_inv / z load covers (U (d_e^2 + d_e) <= 64 lanes), fewer when each
// landmark holds the values of several tiles (registers)
template <int DE, int SEG> constexpr int tile_group() {
    constexpr int cap = 64 / (DE * DE + DE);
    constexpr int want = SEG >= 3 ? 3 : (SEG == 2 ? 4 : 5);
    return cap < want ? cap : want;
}

template <int U, int SEG> struct TileGroup {
    double ez;   // lane l < U (DE^2 + DE): value l % (DE^2 + DE) of slot l / (DE^2 + DE): E_inv, then z
    double x[U], b[U][SEG];
};

template <int DE, class TC, class TEI, class TS>
__global__ __launch_bounds__(64 * TW) void k_tiles(const TileUnit *__restrict__ units, const TileEntry *__restrict__ entries,
                                                  const int32_t *__restrict__ tabs, const TEI *__restrict__ Einv,
                                                  const double *__restrict__ z, const TC *__restrict__ C,
                                                  const double *__restrict__ b_r, const double *__restrict__ s_r,
                                                  TS *S, double *g, const int64_t *__restrict__ S_colptr, int64_t nr,
                                                  int dr, const unsigned long long *flag, unsigned *fin_counter,
                                                  unsigned long long *fin_flag, int64_t *fin_out) {
    constexpr int EZN = DE * DE + DE, EZS = EZN + 1;
    constexpr int UMAX = tile_group<DE, 1>();
    __shared__ double red[TW][NT][256];
    __shared__ double gred[TW - 1][16];
    __shared__ double ezs[TW][2 * UMAX * EZS];
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63, k = lane >> 4, i = lane & 15;
    if (*flag == (unsigned long long)NO_FAILURE) {
        const TileUnit *up = units + blockIdx.x;
        const int ti = __builtin_amdgcn_readfirstlane(up->ti), tj = __builtin_amdgcn_readfirstlane(up->tj);
        const int nt = __builtin_amdgcn_readfirstlane(up->nt);
        const int rb0 = __builtin_amdgcn_readfirstlane(up->rb0), spanc = __builtin_amdgcn_readfirstlane(up->spanc);
        int cnt[NT + 1];
#pragma unroll
        for (int t = 0; t < NT; t++) cnt[t] = __builtin_amdgcn_readfirstlane(up->cnt[t]);
        cnt[NT] = 0;
        const int cst = DE * dr;
        const TileEntry *ep = entries + up->elo;
        // the lane's row (16 ti + i) and, per tile, column (16 (tj + t) + i): block relative to rb0 and byte offset of
        // its value (component k) inside a coupling block, d_e x d_r column-major
        const int64_t R = 16ll * ti + i;
        const bool rv = R < nr && k < DE;
        const int rb = (int)(R / dr);
        const int qr = rv ? rb - rb0 : -1;
        const int aoff = rv ? ((k + DE * (int)(R - (int64_t)rb * dr)) + qr * cst) * (int)sizeof(TC) : 0;
        int qc[NT], boff[NT];
#pragma unroll
        for (int t = 0; t < NT; t++) {
            const int64_t Cc = 16ll * (tj + t) + i;
            const bool cv = Cc < nr && k < DE && t < nt;
            const int cb = (int)(Cc / dr);
            qc[t] = cv ? cb - rb0 : -1;
            boff[t] = cv ? ((k + DE * (int)(Cc - (int64_t)cb * dr)) + qc[t] * cst) * (int)sizeof(TC) : 0;
        }
        const bool diag = tj == ti;
        const __amdgpu_buffer_rsrc_t crs = __builtin_amdgcn_make_buffer_rsrc(
            (void *)(C + up->cbase * cst), (short)0, spanc * cst * (int)sizeof(TC), 0x00020000);
        auto ldc = [&](int off) -> double {
            if constexpr (sizeof(TC) == 8) {
                auto v = __builtin_amdgcn_raw_buffer_load_b64(crs, off, 0, 0);
                return __hiloint2double((int)v[1], (int)v[0]);
            } else {
                return (double)__builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(crs, off, 0, 0));
            }
        };
        d4 D[NT];
#pragma unroll
        for (int t = 0; t < NT; t++) D[t] = d4{0.0, 0.0, 0.0, 0.0};
        double gacc = 0.0;
        // the group's E_inv and z in LDS, slot u at u * EZS: E_inv (column-major), z, then a zero
        const int ez_slot = lane / EZN, ez_q = lane - ez_slot * EZN;
        if (lane < 2 * UMAX) ezs[w][lane * EZS + EZN] = 0.0;
        // the lane's value of the first product's A operand, A'[i][k] (lane 16 k + i): E_inv[i][k] in rows i < d_e,
        // z[k] in row 4, else zero
        const int aq = (i < DE && k < DE) ? i + DE * k : (i == 4 && k < DE) ? DE * DE + k : EZN;
        // the landmarks [lo, hi) of the list, each reaching the run's first SEG tiles
        auto run = [&](auto diag_c, auto seg_c, int lo, int hi) {
            constexpr bool DIAG = decltype(diag_c)::value;
            constexpr int SEG = decltype(seg_c)::value;
            constexpr int U = tile_group<DE, SEG>();
            if (lo >= hi) return;
            const int ngroups = (hi - lo + U - 1) / U;
            double *buf0 = ezs[w], *buf1 = ezs[w] + U * EZS;
            auto load_ev = [&](int gi) {
                const int t = lo + gi * U + lane < hi ? lo + gi * U + lane : hi - 1;
                return *reinterpret_cast<const int4 *>(ep + t);
            };
            // the group's E_inv / z (one coalesced load; every lane's address valid, only lanes < U (d_e^2 + d_e)
            // keep theirs) and, per landmark, the row side and each tile's column side
            auto issue = [&](int gi, TileGroup<U, SEG> &G, const int4 ev) {
                const int e_l = __shfl(ev.x, ez_slot < U ? ez_slot : 0, 64);
                const TEI *pe = Einv + (int64_t)e_l * (DE * DE) + (ez_q < DE * DE ? ez_q : 0);
                const double *pz = z + (int64_t)e_l * DE + (ez_q < DE * DE ? 0 : ez_q - DE * DE);
                if constexpr (sizeof(TEI) == sizeof(double)) {
                    G.ez = *(ez_q < DE * DE ? reinterpret_cast<const double *>(pe) : pz);
                } else {
                    const double ve = (double)*pe, vz = *pz;
                    G.ez = ez_q < DE * DE ? ve : vz;
                }
                const int nvalid = hi - (lo + gi * U);
#pragma unroll
                for (int u = 0; u < U; u++) {
                    const int cw = __builtin_amdgcn_readlane(ev.y, u);
                    const int qq = __builtin_amdgcn_readlane(ev.z, u);
                    const int qf = qq & 0xffff;
                    const unsigned wl = (unsigned)((qq >> 16) - qf);   // window: blocks qf .. qf + wl
                    const int base = (cw - qf) * (cst * (int)sizeof(TC));
                    const bool pr = (unsigned)(qr - qf) <= wl && u < nvalid;
                    G.x[u] = ldc(pr ? base + aoff : (int)0x80000000);
#pragma unroll
                    for (int t = DIAG ? 1 : 0; t < SEG; t++) {
                        const bool pc = (unsigned)(qc[t] - qf) <= wl && u < nvalid;
                        G.b[u][t] = ldc(pc ? base + boff[t] : (int)0x80000000);
                    }
                }
            };
            // per landmark: W = A' x (rows k < d_e: W1[k][i] = sum_n E_inv[k][n] C1[n][i], the A operand of the tile
            // products in their layout; row 4: the g_S part sum_n z[n] C1[n][i]), then D[t] += W1' C2 per tile
            auto compute = [&](const TileGroup<U, SEG> &G, double *buf) {
                if (lane < U * EZN) buf[ez_slot * EZS + ez_q] = G.ez;
                __builtin_amdgcn_wave_barrier();
#pragma unroll
                for (int u = 0; u < U; u++) {
                    const d4 W = __builtin_amdgcn_mfma_f64_16x16x4f64(buf[u * EZS + aq], G.x[u],
                                                                      d4{0.0, 0.0, 0.0, 0.0}, 0, 0, 0);
                    if constexpr (DIAG) gacc += W[1];
#pragma unroll
                    for (int t = 0; t < SEG; t++)
                        D[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(W[0], (DIAG && t == 0) ? G.x[u] : G.b[u][t], D[t],
                                                                    0, 0, 0);
                }
                __builtin_amdgcn_wave_barrier();
            };
            // two groups in flight, their records loaded two groups ahead; the loads of a group past the last are
            // issued all the same (records clamped to the last one, every C load outside the buffer: zeros), so
            // that the loop body has no branch
            TileGroup<U, SEG> GA, GB;
            int4 evA = load_ev(0), evB = load_ev(1);
            issue(0, GA, evA);
            evA = load_ev(2);
            for (int gi = 0; gi < ngroups; gi += 2) {
                issue(gi + 1, GB, evB);
                evB = load_ev(gi + 3);
                compute(GA, buf0);
                if (gi + 1 >= ngroups) break;
                issue(gi + 2, GA, evA);
                evA = load_ev(gi + 4);
                compute(GB, buf1);
            }
        };
        // segment s: the landmarks reaching exactly the first s tiles, [cnt[s], cnt[s - 1]); the wave's share of it
        auto seg = [&](auto diag_c, auto seg_c) {
            constexpr int SEG = decltype(seg_c)::value;
            const int a = cnt[SEG], n = cnt[SEG - 1] - cnt[SEG];
            run(diag_c, seg_c, a + n * w / TW, a + n * (w + 1) / TW);
        };
        auto all_segs = [&](auto diag_c) {
            seg(diag_c, std::integral_constant<int, 4>{});
            seg(diag_c, std::integral_constant<int, 3>{});
            seg(diag_c, std::integral_constant<int, 2>{});
            seg(diag_c, std::integral_constant<int, 1>{});
        };
        if (diag) all_segs(std::true_type{});
        else all_segs(std::false_type{});
        // the waves' partial tiles added in wave order: wave t writes tile t (upper part of S), wave 0 also g
#pragma unroll
        for (int t = 0; t < NT; t++)
            if (t < nt)
#pragma unroll
                for (int v = 0; v < 4; v++) red[w][t][lane * 4 + v] = D[t][v];
        if (w > 0 && k == 0) gred[w - 1][i] = gacc;
        __syncthreads();
        if (w == 0 && diag) {
#pragma unroll
            for (int q = 0; q < TW - 1; q++) gacc += gred[q][i];
            if (k == 0 && R < nr) g[R] = b_r[R] * (s_r ? s_r[R] : 1.0) - gacc;
        }
        const int t = w;
        const int64_t Cc = 16ll * (tj + t) + i;
        if (t < nt && cnt[t] > 0 && Cc < nr) {
            const int cb = (int)(Cc / dr);
            const int cl = (int)(Cc - (int64_t)cb * dr);
            const int cbl = (int)((16ll * (tj + t)) / dr);
            const int64_t colbase = S_colptr[Cc] - 1;
            const int32_t *tab = tabs + ((int64_t)blockIdx.x * NT + t) * 16;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                double dv = 0.0;
#pragma unroll
                for (int q = 0; q < TW; q++) dv += red[q][t][lane * 4 + v];
                // D layout of v_mfma_f64_16x16x4f64: lane l holds rows l / 16 + 4 v of column l % 16
                const int64_t Rr = 16ll * ti + k + 4 * v;
                if (Rr >= nr) continue;
                const int rbb = (int)(Rr / dr);
                const int rll = (int)(Rr - (int64_t)rbb * dr);
                if (rbb > cb || (rbb == cb && cl < rll)) continue;   // below the diagonal block
                const int off = tab[(rbb - rb0) * 4 + (cb - cbl)];
                if (off < 0) continue;
                TS *d = S + colbase + off + rll;
                *d = (TS)((double)*d - dv);
            }
        }
    }
    // the reduction's verdict from the last workgroup to finish: the lowest failing block, the flag re-armed
    if (fin_counter) {
        __syncthreads();
        if (threadIdx.x == 0 && atomicAdd(fin_counter, 1u) == gridDim.x - 1) {
            const unsigned long long f = *fin_flag;
            *fin_out = f == (unsigned long long)NO_FAILURE ? 0 : (int64_t)f;
            *fin_flag = (unsigned long long)NO_FAILURE;
            *fin_counter = 0u;
        }
    }
}

template <int DE, class TC, class TEI, class TS>
void launch_de(const TilePlan &P, const void *E_inv, const double *z, const void *C, const double *b_r,
               const double *s_r, void *S, double *g, const int64_t *S_colptr, int64_t nr,
               const unsigned long long *flag, unsigned *fin_counter, unsigned long long *fin_flag, int64_t *fin_out,
               hipStream_t s) {
    k_tiles<DE, TC, TEI, TS><<<(unsigned)P.nunits, 64 * TW, 0, s>>>(
        P.units.ptr, P.entries.ptr, P.tabs.ptr, (const TEI *)E_inv, z, (const TC *)C, b_r, s_r, (TS *)S, g, S_colptr,
        nr, P.dr, flag, fin_counter, fin_flag, fin_out);
}

template <class TC, class TEI, class TS>
void launch_variant(const TilePlan &P, const void *E_inv, const double *z, const void *C, const double *b_r,
                    const double *s_r, void *S, double *g, const int64_t *S_colptr, int64_t nr,
                    const unsigned long long *flag, unsigned *fin_counter, unsigned long long *fin_flag,
                    int64_t *fin_out, hipStream_t s) {
    switch (P.de) {
    case 1: launch_de<1, TC, TEI, TS>(P, E_inv, z, C, b_r, s_r, S, g, S_colptr, nr, flag, fin_counter, fin_flag, fin_out, s); break;
    case 2: launch_de<2, TC, TEI, TS>(P, E_inv, z, C, b_r, s_r, S, g, S_colptr, nr, flag, fin_counter, fin_flag, fin_out, s); break;
    case 3: launch_de<3, TC, TEI, TS>(P, E_inv, z, C, b_r, s_r, S, g, S_colptr, nr, flag, fin_counter, fin_flag, fin_out, s); break;
    default: launch_de<4, TC, TEI, TS>(P, E_inv, z, C, b_r, s_r, S, g, S_colptr, nr, flag, fin_counter, fin_flag, fin_out, s); break;
    }
}

}  // namespace

int tiles_launch(const TilePlan &P, int variant, const void *E_inv, const double *z, const void *C, const double *b_r,
                 const double *s_r, void *S, double *g, const int64_t *S_colptr, int64_t nr,
                 const unsigned long long *flag, unsigned *fin_counter, unsigned long long *fin_flag, int64_t *fin_out,
                 hipStream_t s) {
    if (P.nunits <= 0) return fail(BLS_ERR_STATE, "no tile plan");
    switch (variant) {
    case 0: launch_variant<double, double, double>(P, E_inv, z, C, b_r, s_r, S, g, S_colptr, nr, flag, fin_counter, fin_flag, fin_out, s); break;
    case 1: launch_variant<float, float, float>(P, E_inv, z, C, b_r, s_r, S, g, S_colptr, nr, flag, fin_counter, fin_flag, fin_out, s); break;
    default: launch_variant<float, double, double>(P, E_inv, z, C, b_r, s_r, S, g, S_colptr, nr, flag, fin_counter, fin_flag, fin_out, s); break;
    }
    BLS_HIP(hipGetLastError());
    return BLS_OK;
}

// The tile plan of a sparse layout. Leaves P.nunits == 0 (and says why) when the structure does not fit: blocks not
// uniform in size (d_e 1..4, d_r 6..16), storage offsets that are not the uniform ones, a landmark whose couplings
// are not a contiguous window of blocks, or a tile row whose couplings span more than the 32-bit buffer offsets allow.
int tiles_build(TilePlan &P, const TileStructure &T, std::string *why) {
    P.units.reset();
    P.entries.reset();
    P.tabs.reset();
    P.nunits = P.nentries = P.nmfma = 0;
    auto no = [&](const std::string &m) {
        if (why) *why = m;
        return BLS_OK;
    };
    const int64_t nr = T.nr, nrb = T.nrb, neb = T.neb;
    if (neb == 0 || nrb == 0) return no("no landmarks");
    const int de = (int)T.e_dims[0], dr = (int)T.r_dims[0];
    if (de < 1 || de > 4) return no("landmark blocks above 4");
    if (dr < 6 || dr > 16) return no("retained blocks outside 6..16");
    for (int64_t e = 0; e < neb; e++)
        if (T.e_dims[e] != de || T.eoffsets[e] != e * de * de || T.boffsets_e[e] != e * de)
            return no("landmark blocks not uniform");
    for (int64_t r = 0; r < nrb; r++)
        if (T.r_dims[r] != dr || T.boffsets_r[r] != r * dr) return no("retained blocks not uniform");
    const int64_t cst = (int64_t)de * dr;
    for (int64_t c = 0; c < (int64_t)T.coffsets.size(); c++)
        if (T.coffsets[c] != c * cst) return no("coupling offsets not uniform");
    struct Raw {
        int64_t ti;
        int32_t e, c0, clast, qf, ql, last;   // last: last tile column reached, relative to ti
    };
    std::vector<Raw> raw;
    raw.reserve((size_t)(T.eptr[neb] * 2));
    for (int64_t e = 0; e < neb; e++) {
        const int64_t p0 = T.eptr[e], p1 = T.eptr[e + 1];
        if (p1 == p0) continue;
        for (int64_t c = p0 + 1; c < p1; c++)
            if (T.eret[c] != T.eret[c - 1] + 1) return no("a landmark's couplings are not a contiguous window");
        const int64_t first_row = T.eret[p0] * dr, last_col = T.eret[p1 - 1] * dr + dr - 1;
        for (int64_t ti = first_row / 16; ti <= last_col / 16; ti++) {
            const int64_t rb0 = (16 * ti) / dr;
            const int64_t rbh = std::min(16 * ti + 15, nr - 1) / dr;
            if (T.eret[p1 - 1] < rb0 || T.eret[p0] > rbh) continue;   // no block in the tile row's rows
            int64_t c0 = p0;
            while (T.eret[c0] < rb0) c0++;
            const int64_t qf = T.eret[c0] - rb0, ql = T.eret[p1 - 1] - rb0;
            if (ql >= 0x7fff) return no("window too long");
            raw.push_back(Raw{ti, (int32_t)e, (int32_t)c0, (int32_t)(p1 - 1), (int32_t)qf, (int32_t)ql,
                              (int32_t)(last_col / 16 - ti)});
        }
    }
    std::stable_sort(raw.begin(), raw.end(), [](const Raw &a, const Raw &b) {
        return a.ti != b.ti ? a.ti < b.ti : a.last > b.last;
    });
    struct Pending {
        int64_t weight, tab;
        TileUnit u;
    };
    std::vector<Pending> us;
    std::vector<TileEntry> ents;
    std::vector<int32_t> tabs;
    ents.reserve(raw.size());
    const int64_t max_bytes = 0x7fffff00ll;
    const int64_t ntr = (nr + 15) / 16;
    // the 4 x 4 offsets in S's scalar columns of the (row block, column block) pairs of tile (ti, tj); -1: not stored
    auto push_tab = [&](int64_t ti, int64_t tj, bool any) {
        const int64_t t0 = (int64_t)tabs.size();
        tabs.resize(tabs.size() + 16, -1);
        if (!any || 16 * tj >= nr) return;
        const int64_t rb0 = (16 * ti) / dr, cbl = (16 * tj) / dr;
        for (int q = 0; q < 4; q++)
            for (int p = 0; p < 4; p++) {
                const int64_t rb = rb0 + q, c = cbl + p;
                if (rb >= nrb || c >= nrb || rb > c) continue;
                for (int64_t kk = (*T.bcp)[c]; kk < (*T.bcp)[c + 1]; kk++)
                    if ((*T.brv)[kk] == rb) tabs[t0 + q * 4 + p] = (int32_t)(*T.bro)[kk];
            }
    };
    size_t r0 = 0;
    for (int64_t ti = 0; ti < ntr; ti++) {
        const int64_t rb0 = (16 * ti) / dr;
        if (r0 == raw.size() || raw[r0].ti != ti) {
            // no landmark couples the tile row's blocks: one run that only writes g = b_r s_r
            TileUnit u{};
            u.ti = u.tj = (int32_t)ti;
            u.nt = 1;
            u.rb0 = (int32_t)rb0;
            u.elo = (int64_t)ents.size();
            const int64_t t0 = (int64_t)tabs.size() / 16;
            for (int t = 0; t < NT; t++) push_tab(ti, ti + t, false);
            us.push_back(Pending{0, t0, u});
            continue;
        }
        size_t r1 = r0;
        int64_t cb = INT64_MAX, cm = -1;
        while (r1 < raw.size() && raw[r1].ti == ti) {
            cb = std::min<int64_t>(cb, raw[r1].c0);
            cm = std::max<int64_t>(cm, raw[r1].clast);
            r1++;
        }
        const int64_t spanc = cm + 1 - cb;
        if (spanc * cst * 8 >= max_bytes) return no("a tile row's couplings span more than 2 GB");
        if ((int64_t)(r1 - r0) > INT32_MAX) return no("too many landmarks in a tile row");
        const int64_t elo = (int64_t)ents.size();
        for (size_t q = r0; q < r1; q++)
            ents.push_back(TileEntry{raw[q].e, (int32_t)(raw[q].c0 - cb), raw[q].qf | (raw[q].ql << 16), raw[q].last});
        // prefix[s]: the landmarks reaching tile column ti + s (the list is sorted by reach, descending)
        const int maxlast = raw[r0].last;
        std::vector<int64_t> prefix(maxlast + 2, 0);
        for (size_t q = r0; q < r1; q++) prefix[raw[q].last]++;
        for (int s = maxlast - 1; s >= 0; s--) prefix[s] += prefix[s + 1];
        for (int s0 = 0; s0 <= maxlast && 16 * (ti + s0) < nr; s0 += NT) {
            TileUnit u{};
            u.ti = (int32_t)ti;
            u.tj = (int32_t)(ti + s0);
            u.rb0 = (int32_t)rb0;
            u.spanc = (int32_t)spanc;
            u.elo = elo;
            u.cbase = cb;
            int nt = 0;
            int64_t weight = prefix[s0];   // the W products
            for (int t = 0; t < NT; t++) {
                const int s = s0 + t;
                if (s > maxlast || 16 * (ti + s) >= nr) break;
                const int64_t cbh = std::min(16 * (ti + s) + 15, nr - 1) / dr;
                if (cbh - rb0 >= 64) return no("a tile column more than 64 blocks from its row");
                u.cnt[t] = (int32_t)prefix[s];
                weight += prefix[s];
                nt = t + 1;
            }
            u.nt = nt;
            const int64_t t0 = (int64_t)tabs.size() / 16;
            for (int t = 0; t < NT; t++) push_tab(ti, ti + s0 + t, t < nt);
            us.push_back(Pending{weight, t0, u});
            P.nmfma += weight;
        }
        r0 = r1;
    }
    // Dispatch order. The dispatcher places workgroup p on XCD p % 8, and the runs of one tile row (and of
    // neighbouring rows) read the same coupling values: each XCD gets a contiguous range of tile rows of about equal
    // weight, in its own L2, and runs its runs heaviest first. Positions of an XCD past its last run hold empty
    // units (no landmarks, not diagonal: they write nothing).
    std::vector<std::vector<Pending>> xcd(8);
    {
        int64_t wsum = 0;
        for (const Pending &x : us) wsum += x.weight + 64;   // 64: a run's fixed cost in landmark units
        int64_t acc = 0;
        for (size_t q = 0; q < us.size();) {   // us is in tile-row order
            size_t q1 = q;
            int64_t rw = 0;
            while (q1 < us.size() && us[q1].u.ti == us[q].u.ti) rw += us[q1++].weight + 64;
            const int x = (int)std::min<int64_t>(7, (acc + rw / 2) * 8 / std::max<int64_t>(wsum, 1));
            xcd[x].insert(xcd[x].end(), us.begin() + q, us.begin() + q1);
            acc += rw;
            q = q1;
        }
    }
    size_t per = 0;
    for (auto &v : xcd) {
        std::stable_sort(v.begin(), v.end(), [](const Pending &a, const Pending &b) { return a.weight > b.weight; });
        per = std::max(per, v.size());
    }
    std::vector<TileUnit> units;
    std::vector<int32_t> otabs;
    units.reserve(per * 8);
    otabs.reserve(per * 8 * NT * 16);
    for (size_t j = 0; j < per; j++)
        for (int x = 0; x < 8; x++) {
            if (j < xcd[x].size()) {
                units.push_back(xcd[x][j].u);
                const int64_t t0 = xcd[x][j].tab * 16;
                otabs.insert(otabs.end(), tabs.begin() + t0, tabs.begin() + t0 + NT * 16);
            } else {
                TileUnit e{};
                e.tj = 1;   // ti 0, tj 1: not diagonal; no tiles, no landmarks
                units.push_back(e);
                otabs.insert(otabs.end(), NT * 16, -1);
            }
        }
    if (units.size() > (size_t)INT32_MAX) return no("too many tiles");
    int rc;
    if ((rc = P.units.upload(units)) || (rc = P.entries.upload(ents)) || (rc = P.tabs.upload(otabs))) return rc;
    P.nunits = (int64_t)units.size();
    P.nentries = (int64_t)ents.size();
    P.de = de;
    P.dr = dr;
    if (why) why->clear();
    return BLS_OK;
}

}  // namespace bls
