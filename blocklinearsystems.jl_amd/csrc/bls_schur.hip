// Schur-complement elimination on gfx950: reset_schur_{sparse,dense}!, reduce_schur_{sparse,dense}!,
// backsubstitute! and the copyto! gather, behind the C ABI of include/bls_amd.h.
//
// Phase 1 (factor_eliminated!, src/schur_reduction.jl:121-162): one lane per eliminated block of a run of equal
// d_e; closed-form inverse for d_e <= 3 and a register Cholesky for 4..8 (inv_sym_block, src/kernels.jl:14-83),
// written without contraction so the acceptance test makes the reference's decisions; caches E_inv and z.
//
// Phase 2 (accumulate_*_columns!, src/schur_reduction.jl:262-380) is owner-computes, like the reference's
// col_chunks: a workgroup owns an "item" -- a run of consecutive block rows ("slots") of one block column r2 of S
// -- and each lane owns one (slot, row) of that column, holding the d_r2 entries of that row in registers. The
// item's landmarks (the couplings of r2 in ascending eliminated block, the reference's order) are streamed in
// batches: their E_inv, z and C_c2 blocks are staged in LDS with coalesced loads, then every lane applies, in
// order, the landmarks that touch its slot -- W_c1[:,row] = E_inv C_c1[:,row] formed in registers from its own
// column of C_c1 and S[row, col] -= W_c1[:,row] . C_c2[:,col]. Each S entry receives its updates from one lane
// (or from a few lanes whose partial sums are merged in a fixed order): no float atomics, and the result is
// bitwise reproducible run to run.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <type_traits>

#include "bls_internal.h"

using namespace bls;

namespace {

// Element types of a precision variant: E and C as the caller stores them, E_inv as cached, R and S of the
// reduced system. z, g_S, b and the damping are Float64 in all of them (src/schur_reduction.jl:16-92).
//   f64    SchurBlockHessian{Float64}
//   f32    SchurBlockHessian{Float32}: the landmark factorisation in Float32 (inv_sym_block in T, so that blocks are
//          accepted and rejected as the reference's Float32 path decides), E_inv and S stored in Float32
//   mixed  Float32 blocks E, C, R in; the landmark factorisation, E_inv, S and g_S in Float64 (no reference
//          counterpart: halves the bytes of the blocks and leaves the reduced system exact for the stored blocks)
// Products are formed and accumulated in Float64 registers in every variant: without packed FP32 math a Float32
// FMA issues at the Float64 rate on this hardware, so Float32 pays only through the bytes it saves.
// TSC: the type a conditioned coupling value scaled(C, s_e s_r) is held in -- the element type T for f64 and f32,
// as the reference narrows it (src/conditioning.jl:202); Float64 for mixed, whose reduced system is the Float64
// one of the stored blocks. With conditioning scales phase 2 reads that scaled copy, with the policy `Scaled`.
struct PrecF64 {
    using TE = double; using TEI = double; using TC = double; using TS = double; using TR = double;
    using TSC = double;
    static constexpr bool F32_FACTOR = false;
};
struct PrecF32 {
    using TE = float; using TEI = float; using TC = float; using TS = float; using TR = float;
    using TSC = float;
    static constexpr bool F32_FACTOR = true;
};
struct PrecMixed {
    using TE = float; using TEI = double; using TC = float; using TS = double; using TR = float;
    using TSC = double;
    static constexpr bool F32_FACTOR = false;
};
// the policy phase 2 runs with when it reads the scaled coupling copy
template <class P> struct ScaledPolicy { using type = P; };
template <> struct ScaledPolicy<PrecMixed> { using type = PrecF64; };
template <class P> constexpr bool all_f64() {
    return sizeof(typename P::TE) == 8 && sizeof(typename P::TEI) == 8 && sizeof(typename P::TC) == 8 &&
           sizeof(typename P::TS) == 8;
}

constexpr int P2_THREADS = 256;
constexpr int P2_MAXE = 64;        // entries (landmarks) per staged batch
constexpr int P2_BASE_LANES = 96;  // default: (slot, row) lanes of one item before its heavy slots get lane groups
constexpr int MAXDR = 16;          // largest retained block handled

struct P2Item {
    int64_t base;        // S index of (row 0, col 0) of the column, before the slot's row base
    int64_t colstride;   // distance between scalar columns of the block column
    int64_t entry_lo, entry_hi;
    int64_t pbase;       // split items: this part's rows in the partial buffer
    int32_t r2, dr2, ro2;
    int32_t slot_lo, nslots, ntasks, task_lo;
    int32_t has_diag, uniform_dr;
    int32_t part, nparts;
};

// A split item: its parts' row sums (nparts x rows x (d_r2 + 1), the last column g) are added in part order.
struct P2Merge {
    int64_t base, colstride, pbase;
    int32_t slot_lo, nslots, rows, dr2, ro2, has_diag, nparts;
};

struct P2Entry {
    int32_t e, c_first, c2, pad;
    uint64_t mask;
};

// A lane's share of an item: one row of one slot. A slot with many landmarks is split over `nchunk` lane groups
// (tasks slot_base + chunk * d_r1 + row); group `chunk` takes that share of the slot's landmarks in every batch,
// and the groups' partial sums are merged in chunk order at the end.
struct P2Task {
    int16_t slot, row, chunk, nchunk;
};

template <class P> struct P2Args {
    using TEI = typename P::TEI;
    using TC = typename P::TC;
    using TS = typename P::TS;
    const P2Item *items;
    int64_t nitems;
    const P2Entry *entries;
    const P2Task *tasks;
    const int64_t *slot_rowbase;
    const int32_t *slot_dr, *slot_roff;
    const int64_t *eoffsets, *boffsets_e, *coffsets;
    const int32_t *e_dims;
    const TEI *E_inv;
    const double *z;
    const TC *C;
    const double *b_r, *s_r;
    TS *S;
    double *g;
    double *partial;
    const unsigned long long *flag;
    int uniform_offsets;   // one run of d_e == DE and one d_r: E/z/C offsets are e*DE^2, e*DE, c*cstride
    int64_t cstride;
    // the reduction's verdict from the last workgroup to finish (no split items): the lowest failing block into
    // fin_out, the flag re-armed; fin_counter counts finished workgroups and is left at 0 (NULL: a separate kernel)
    unsigned int *fin_counter;
    unsigned long long *fin_flag;
    int64_t *fin_out;
};

// ---------------------------------------------------------------------------------------------------------
// inv_sym_block (src/kernels.jl:14-83). H: DE x DE column-major (scaled), damp: lambda*scale (already formed).
// ---------------------------------------------------------------------------------------------------------
#pragma clang fp contract(off)
template <int DE, class F>
__device__ inline bool inv_sym(const F *H, const F *damp, F *inv) {
    const F tol = (F)4504 * (sizeof(F) == 8 ? (F)2.220446049250313e-16 : (F)1.1920928955078125e-07);
    if constexpr (DE == 1) {
        const F h11 = H[0] + damp[0];
        if (!(h11 > 0) || !isfinite(h11)) return false;
        inv[0] = (F)1 / h11;
        return true;
    } else if constexpr (DE == 2) {
        const F H11 = H[0] + damp[0], H22 = H[3] + damp[1], H12 = H[2];
        const F det = H11 * H22 - H12 * H12;
        if (!(H11 > 0) || !(det > tol * (H11 * H22)) || !isfinite(det)) return false;
        const F id = (F)1 / det;
        inv[0] = H22 * id; inv[1] = -H12 * id; inv[2] = -H12 * id; inv[3] = H11 * id;
        return true;
    } else if constexpr (DE == 3) {
        const F H11 = H[0] + damp[0], H22 = H[4] + damp[1], H33 = H[8] + damp[2];
        const F H12 = H[3], H13 = H[6], H23 = H[7];
        const F c11 = H22 * H33 - H23 * H23;
        const F c12 = H13 * H23 - H12 * H33;
        const F c13 = H12 * H23 - H13 * H22;
        const F c22 = H11 * H33 - H13 * H13;
        const F c23 = H12 * H13 - H11 * H23;
        const F c33 = H11 * H22 - H12 * H12;
        const F det = H11 * c11 + H12 * c12 + H13 * c13;
        if (!(H11 > 0) || !(c33 > tol * (H11 * H22)) || !(det > tol * (c33 * H33)) || !isfinite(det)) return false;
        const F id = (F)1 / det;
        inv[0] = c11 * id; inv[1] = c12 * id; inv[2] = c13 * id;
        inv[3] = c12 * id; inv[4] = c22 * id; inv[5] = c23 * id;
        inv[6] = c13 * id; inv[7] = c23 * id; inv[8] = c33 * id;
        return true;
    } else {
        // Cholesky of Symmetric(H + diag(damp), :U): U'U, every pivot ratio |U_ii|^2 / Hd_ii > tol; inv = U^-1 U^-T
        F U[DE * DE];
#pragma unroll
        for (int q = 0; q < DE * DE; q++) U[q] = 0;
#pragma unroll
        for (int j = 0; j < DE; j++) {
#pragma unroll
            for (int i = 0; i <= j; i++) {
                F a = H[j * DE + i] + (i == j ? damp[i] : (F)0);
                for (int k = 0; k < i; k++) a -= U[i * DE + k] * U[j * DE + k];
                if (i == j) {
                    if (!(a > 0) || !isfinite(a)) return false;
                    U[j * DE + j] = sqrt(a);
                } else {
                    U[j * DE + i] = a / U[i * DE + i];
                }
            }
        }
        F min_ratio = (F)INFINITY;
        bool nan_ratio = false;
#pragma unroll
        for (int i = 0; i < DE; i++) {
            const F r = (U[i * DE + i] * U[i * DE + i]) / (H[i * DE + i] + damp[i]);
            if (isnan(r)) nan_ratio = true;
            min_ratio = r < min_ratio ? r : min_ratio;
        }
        if (nan_ratio || !(min_ratio > tol && isfinite(min_ratio))) return false;
        F Ui[DE * DE];
#pragma unroll
        for (int q = 0; q < DE * DE; q++) Ui[q] = 0;
#pragma unroll
        for (int j = 0; j < DE; j++) {
            Ui[j * DE + j] = (F)1 / U[j * DE + j];
            for (int i = j - 1; i >= 0; i--) {
                F s = 0;
                for (int k = i + 1; k <= j; k++) s += U[k * DE + i] * Ui[j * DE + k];
                Ui[j * DE + i] = -s / U[i * DE + i];
            }
        }
#pragma unroll
        for (int j = 0; j < DE; j++)
#pragma unroll
            for (int i = 0; i < DE; i++) {
                F s = 0;
                for (int k = (i > j ? i : j); k < DE; k++) s += Ui[k * DE + i] * Ui[k * DE + j];
                inv[j * DE + i] = s;
            }
        return true;
    }
}

// Phase 1 over one run of equal d_e: blocks first..first+count-1 (0-based), E/E_inv at eoff0 + i*DE^2,
// b_e/z/scaling at boff0 + i*DE.
// The damped, scaled block is formed as the reference forms it (scaled(E, s_i s_j) and T((lambda sc) s^2), in
// Float64 and narrowed once) and inverted in Float32 for the f32 variant, Float64 otherwise; z in Float64.
template <int DE, class P>
__global__ __launch_bounds__(256) void k_phase1(int64_t first, int64_t count, int64_t eoff0, int64_t boff0,
                                                const typename P::TE *__restrict__ E, const double *__restrict__ b_e,
                                                const double *__restrict__ scaling_e, double lambda,
                                                const double *__restrict__ s_e, typename P::TEI *__restrict__ E_inv,
                                                double *__restrict__ z, unsigned long long *flag) {
    using F = typename std::conditional<P::F32_FACTOR, float, double>::type;
    using TEI = typename P::TEI;
    // blocks of up to 4 x 4: the workgroup's E blocks are read, and its E_inv blocks written, as one contiguous
    // run through LDS (coalesced), each lane then taking its own block from LDS
    constexpr bool STG = DE <= 4;
    constexpr int NQ = DE * DE;
    __shared__ double sbuf[STG ? 256 * NQ : 1];
    __shared__ unsigned char sok[STG ? 256 : 1];
    const int t = threadIdx.x;
    const int64_t w0 = (int64_t)blockIdx.x * blockDim.x;
    const int64_t i = w0 + t;
    const bool valid = i < count;
    const int64_t nw = count - w0 < (int64_t)blockDim.x ? count - w0 : (int64_t)blockDim.x;
    if constexpr (STG) {
        for (int q = t; q < nw * NQ; q += 256) sbuf[q] = (double)E[eoff0 + w0 * NQ + q];
        __syncthreads();
    } else {
        if (!valid) return;
    }
    const int64_t eo = eoff0 + i * NQ, bo = boff0 + i * DE;
    double se[DE];
    F damp[DE], H[NQ], inv[NQ];
    bool ok = false;
    if (valid) {
#pragma unroll
        for (int k = 0; k < DE; k++) {
            se[k] = s_e ? s_e[bo + k] : 1.0;
            damp[k] = (F)((lambda * scaling_e[bo + k]) * (se[k] * se[k]));
        }
#pragma unroll
        for (int k = 0; k < NQ; k++) {
            const typename P::TE ev = STG ? (typename P::TE)sbuf[t * NQ + k] : E[eo + k];
            H[k] = s_e ? (F)((double)ev * (se[k % DE] * se[k / DE])) : (F)ev;
        }
        ok = inv_sym<DE, F>(H, damp, inv);
        if (!ok) atomicMin(flag, (unsigned long long)(first + i + 1));
    }
    if constexpr (STG) {
        __syncthreads();   // every lane has read its E block
        if (valid && ok) {
#pragma unroll
            for (int k = 0; k < NQ; k++) sbuf[t * NQ + k] = (double)(TEI)inv[k];
        }
        sok[t] = valid && ok;
        __syncthreads();
        for (int q = t; q < nw * NQ; q += 256)
            if (sok[q / NQ]) E_inv[eoff0 + w0 * NQ + q] = (TEI)sbuf[q];   // a failing block's E_inv is not written
    } else {
        if (!ok) return;
#pragma unroll
        for (int k = 0; k < NQ; k++) E_inv[eo + k] = (TEI)inv[k];
    }
    if (!valid || !ok) return;
#pragma unroll
    for (int r = 0; r < DE; r++) {
        double acc = 0.0;
#pragma unroll
        for (int j = 0; j < DE; j++) acc += (double)inv[j * DE + r] * (b_e[bo + j] * se[j]);
        z[bo + r] = acc;
    }
}
#pragma clang fp contract(fast)

__global__ void k_fill_u64(unsigned long long *p, unsigned long long v) { *p = v; }

// The verdict of a reduction (0, or the lowest failing block); the flag is left at NO_FAILURE for the next one.
__global__ void k_finalize_flag(unsigned long long *flag, int64_t *out) {
    const unsigned long long f = *flag;
    *out = f == (unsigned long long)NO_FAILURE ? 0 : (int64_t)f;
    *flag = (unsigned long long)NO_FAILURE;
}

// ---------------------------------------------------------------------------------------------------------
// Phase 2
//
// A workgroup streams its item's landmark entries in batches of up to P2_MAXE. Per batch, each landmark's shared
// data [E_inv | z | C_c2] (d_e^2 + d_e + d_e d_r2 doubles) is staged in LDS with coalesced loads. Every lane then
// applies the batch's landmarks that touch its slot, in order, loading its own column C_c1[:, row] (d_e
// contiguous doubles) from global memory, P2_G landmarks at a time so that their loads are in flight together:
//   w = E_inv C_c1[:, row]   (a column of the reference's W_data, src/schur_reduction.jl:149-160)
//   S[row, c] -= w . C_c2[:, c]   and, on the diagonal slot, g[row] -= C_c2[:, row] . z
// Software pipeline: while batch b is applied, the staged data of batch b+1 is in flight into registers and
// wave 0 describes batch b+2 (descriptions triple-buffered, staging double-buffered: one barrier per batch).
// A slot touched by many landmarks is split over `nchunk` lane groups; in every batch, group k takes the k-th of
// nchunk equal contiguous shares of the slot's landmarks, so the groups stay balanced batch by batch. An item
// with many landmarks is split into `nparts` workgroups over consecutive landmark ranges, whose row sums go to a
// partial buffer and are added in part order by k_phase2_merge. Every sum is taken in a fixed order:
// deterministic, no atomics.
//
// With one landmark dimension DE (the common case) a staged record has a fixed size SZ, record j sits at j*SZ
// and four lanes stage each record. Mixed landmark dimensions (DE == 0) pack records of varying size; a lane
// finds the record holding a staged position from bit masks of the records' starts.
// ---------------------------------------------------------------------------------------------------------
constexpr int P2_STAGE = 2048;                 // doubles per staging buffer (16 KB)
constexpr int P2_Q = P2_STAGE / P2_THREADS;    // staged doubles per lane per batch (mixed dimensions)
constexpr int P2_G = 4;                        // landmarks whose C_c1 columns a lane loads together
constexpr int P2_G4 = 2;                       // the same at four workgroups per CU (d_e 3, d_r <= 6)
constexpr int P2_WORDS = P2_STAGE / 64;
constexpr int P2_PART = 384;                   // default: landmarks per workgroup before an item is split

// Loaded doubles per lane per batch. One landmark dimension: four lanes per landmark, lane 0 loading E_inv and z
// (DE^2 + DE), lanes 1-3 every third column of C_c2 (DE each, up to DRM columns); mixed dimensions: P2_Q.
template <int DE, int DRM> struct P2Width {
    static constexpr int c2cols = (DRM + 2) / 3;
    static constexpr int value = DE > 0 ? (DE * DE + DE > DE * c2cols ? DE * DE + DE : DE * c2cols) : P2_Q;
};

// Doubles of a landmark's staged record: with one landmark dimension [W2 = E_inv C_c2 (DE x d_r2) | gz = C_c2' z
// (d_r2)], formed by the staging lanes; with mixed dimensions the raw [E_inv | z | C_c2].
template <int DE> __device__ __host__ inline int p2_record(int de, int dr2) {
    return DE > 0 ? (DE + 1) * dr2 : de * de + de + de * dr2;
}

struct P2Meta {
    int nb;
    int off[P2_MAXE + 1];       // staged offset of each entry's [E_inv | z | C_c2]
    int de[P2_MAXE];
    unsigned long long mask[P2_MAXE];
    int64_t src_e[P2_MAXE], src_z[P2_MAXE], src_c1[P2_MAXE], src_c2[P2_MAXE];
    unsigned long long slotbits[64];   // per slot of the item: which entries of the batch touch it
    // mixed dimensions: staged positions where an entry starts, one bit per position, with the number of starts
    // before each 64-position word: the entry holding position p is found with one popcount
    unsigned long long startbits[P2_WORDS];
    int startcnt[P2_WORDS];
};

// Load up to 64 entries starting at `start` (lane i takes entry start + i). The load is unconditional, from a
// clamped index (the entry array ends with an unused record, so entry_lo is always a valid index), so that no
// branch joins on its result and waiting for it never waits for loads issued after it.
template <class P>
__device__ inline bool p2_load_entries(const P2Args<P> &a, int64_t start, int64_t lo, int64_t end, int lane,
                                       P2Entry &en) {
    const bool valid = start + lane < end;
    en = a.entries[valid ? start + lane : lo];
    return valid;
}

__device__ inline int wave_incl_scan(int v, int lane) {
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(v, d, 64);
        if (lane >= d) v += y;
    }
    return v;
}

// Wave 0: form the batch from the loaded entries and describe it in `m`; returns its entry count. With one
// landmark dimension the batch holds up to `nbmax` records of SZ doubles; otherwise it is the longest prefix
// whose records fit P2_STAGE.
template <int DE, class P>
__device__ inline int p2_process(const P2Args<P> &a, const P2Item &it, P2Meta &m, bool valid, const P2Entry &en,
                                 int lane, int nbmax) {
    int de = DE;
    if (DE == 0) de = valid ? a.e_dims[en.e] : 1;
    int64_t se = 0, sz = 0, sc1 = 0, sc2 = 0;
    if (valid) {
        if (a.uniform_offsets) {
            se = (int64_t)en.e * (DE * DE);
            sz = (int64_t)en.e * DE;
            sc1 = (int64_t)en.c_first * a.cstride;
            sc2 = (int64_t)en.c2 * a.cstride;
        } else {
            se = a.eoffsets[en.e];
            sz = a.boffsets_e[en.e];
            sc1 = a.coffsets[en.c_first];
            sc2 = a.coffsets[en.c2];
        }
    }
    const int size = p2_record<DE>(de, it.dr2);
    int nb, excl;
    if constexpr (DE > 0) {
        nb = __popcll(__ballot(valid && lane < nbmax));
        excl = lane * size;
    } else {
        const int sz_l = valid ? size : 0;
        const int incl = wave_incl_scan(sz_l, lane);
        excl = incl - sz_l;
        nb = __popcll(__ballot(valid && incl <= P2_STAGE));
    }
    if (lane < nb) {
        m.off[lane] = excl;
        m.de[lane] = de;
        m.mask[lane] = en.mask;
        m.src_e[lane] = se;
        m.src_z[lane] = sz;
        m.src_c1[lane] = sc1;
        m.src_c2[lane] = sc2;
        if (lane == nb - 1) m.off[nb] = excl + size;
    }
    if (lane == 0) {
        m.nb = nb;
        if (nb == 0) m.off[0] = 0;
    }
    const unsigned long long mk = lane < nb ? en.mask : 0ull;
    for (int sl = 0; sl < it.nslots; sl++) {
        const unsigned long long b = __ballot((mk >> sl) & 1ull);
        if (lane == 0) m.slotbits[sl] = b;
    }
    if constexpr (DE == 0) {
        if (lane < P2_WORDS) m.startbits[lane] = 0ull;
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        if (lane < nb) atomicOr(&m.startbits[excl >> 6], 1ull << (excl & 63));
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        const int ps = lane < P2_WORDS ? __popcll(m.startbits[lane]) : 0;
        const int ips = wave_incl_scan(ps, lane);
        if (lane < P2_WORDS) m.startcnt[lane] = ips - ps;
    }
    return nb;
}

// Mixed dimensions: index of the entry holding staged position p (0 <= p < P2_STAGE): starts at or before p,
// less one. Past the last entry it returns the last one.
__device__ inline int seg_lookup(const P2Meta &m, int p) {
    const int w = p >> 6;
    const unsigned long long upto = (2ull << (p & 63)) - 1ull;   // bits 0..p%64 (all 64 when p%64 == 63)
    return m.startcnt[w] + __popcll(m.startbits[w] & upto) - 1;
}

// All lanes: issue the loads of a batch's staged data into registers (addresses first, then the loads).
template <int DE, int DRM, class P>
__device__ inline void p2_issue(const P2Args<P> &a, const P2Meta &m, int t, int dr2, double *v) {
    constexpr int QN = P2Width<DE, DRM>::value;
    if constexpr (!all_f64<P>()) {
        // typed loads: E_inv (TEI), z (Float64) and C (TC), each promoted to Float64
        if constexpr (DE > 0) {
            const int j = t >> 2, part = t & 3;
            const bool has = j < m.nb;
            const int jj = has ? j : 0;
            const int64_t se = m.src_e[jj], sz = m.src_z[jj], sc = m.src_c2[jj];
#pragma unroll
            for (int k = 0; k < QN; k++) {
                const int c = (part - 1) + 3 * (k / DE);
                const bool ok0 = has && part == 0 && k < DE * DE + DE;
                const bool ok1 = has && part != 0 && c < dr2;
                double x = 0.0;
                if (ok0) x = k < DE * DE ? (double)a.E_inv[se + k] : a.z[sz + (k - DE * DE)];
                else if (ok1) x = (double)a.C[sc + c * DE + (k % DE)];
                v[k] = x;
            }
        } else {
            const int L = m.off[m.nb];
#pragma unroll
            for (int q = 0; q < QN; q++) {
                const int p = t + q * P2_THREADS;
                double x = 0.0;
                if (p < L) {
                    const int j = seg_lookup(m, p);
                    const int de = m.de[j];
                    const int r = p - m.off[j];
                    const int d2 = de * de, d3 = d2 + de;
                    x = r < d2 ? (double)a.E_inv[m.src_e[j] + r]
                               : (r < d3 ? a.z[m.src_z[j] + (r - d2)] : (double)a.C[m.src_c2[j] + (r - d3)]);
                }
                v[q] = x;
            }
        }
        return;
    } else {
    const double *src[QN];
    if constexpr (DE > 0) {
        // lanes 4j..4j+3 load landmark j: lane 4j its E_inv then z, lane 4j+p (p = 1..3) the C_c2 columns
        // p-1, p+2, ... (DE doubles each)
        const int j = t >> 2, part = t & 3;
        const bool has = j < m.nb;
        const int jj = has ? j : 0;
        const int64_t se = m.src_e[jj], sz = m.src_z[jj], sc = m.src_c2[jj];
#pragma unroll
        for (int k = 0; k < QN; k++) {
            const int c = (part - 1) + 3 * (k / DE);   // C_c2 column of this double (lanes 1-3)
            const double *s0 = k < DE * DE ? a.E_inv + se + k : a.z + sz + (k - DE * DE);
            const double *s1 = a.C + sc + c * DE + (k % DE);
            const bool ok0 = part == 0 && k < DE * DE + DE;
            const bool ok1 = part != 0 && c < dr2;
            src[k] = has && (ok0 || ok1) ? (part == 0 ? s0 : s1) : a.C;
        }
    } else {
        const int L = m.off[m.nb];
#pragma unroll
        for (int q = 0; q < QN; q++) {
            const int p = t + q * P2_THREADS;
            const int j = seg_lookup(m, p);
            const int de = m.de[j];
            const int r = p - m.off[j];
            const int d2 = de * de, d3 = d2 + de;
            const double *s1 = a.E_inv + m.src_e[j] + r;
            const double *s2 = a.z + m.src_z[j] + (r - d2);
            const double *s3 = a.C + m.src_c2[j] + (r - d3);
            const double *sp = r < d2 ? s1 : (r < d3 ? s2 : s3);
            src[q] = p < L ? sp : a.C;
        }
    }
#pragma unroll
    for (int q = 0; q < QN; q++) v[q] = *src[q];
    }
}

template <int DE, int DRM>
__device__ inline void p2_store(const P2Meta &m, int t, int dr2, const double *v, double *stage) {
    constexpr int QN = P2Width<DE, DRM>::value;
    if constexpr (DE > 0) {
        // lanes 4j+1..4j+3 take E_inv and z from lane 4j and write W2 = E_inv C_c2 and gz = C_c2' z for their
        // columns of C_c2 into record j (the reference's W_data for the coupling of the block column,
        // src/schur_reduction.jl:149-160, and its g term)
        constexpr int CC = P2Width<DE, DRM>::c2cols;
        const int j = t >> 2, part = t & 3;
        const int src_lane = (t & 63) & ~3;
        double ez[DE * DE + DE];
#pragma unroll
        for (int q = 0; q < DE * DE + DE; q++) ez[q] = __shfl(v[q], src_lane, 64);
        if (j < m.nb && part != 0) {
            const int size = (DE + 1) * dr2;
            double *rec = stage + j * size;
#pragma unroll
            for (int cc = 0; cc < CC; cc++) {
                const int c = (part - 1) + 3 * cc;
                if (c < dr2) {
                    const double *col = v + cc * DE;
                    double gz = 0.0;
#pragma unroll
                    for (int i = 0; i < DE; i++) gz += col[i] * ez[DE * DE + i];
#pragma unroll
                    for (int k = 0; k < DE; k++) {
                        double w = 0.0;
#pragma unroll
                        for (int i = 0; i < DE; i++) w += ez[k + i * DE] * col[i];
                        rec[c * DE + k] = w;
                    }
                    rec[DE * dr2 + c] = gz;
                }
            }
        }
    } else {
        const int L = m.off[m.nb];
#pragma unroll
        for (int q = 0; q < QN; q++) {
            const int p = t + q * P2_THREADS;
            if (p < L) stage[p] = v[q];
        }
    }
}

// Position of the r-th (0-based) set bit of x; 64 when x has at most r set bits.
__device__ inline int nth_set_bit(unsigned long long x, int r) {
    if (r >= __popcll(x)) return 64;
    int pos = 0;
#pragma unroll
    for (int w = 32; w >= 1; w >>= 1) {
        const int c = __popcll((x >> pos) & ((1ull << w) - 1ull));
        if (r >= c) {
            r -= c;
            pos += w;
        }
    }
    return pos;
}

__device__ inline unsigned long long bits_below(int p) { return p >= 64 ? ~0ull : ((1ull << p) - 1ull); }

// Staging capacity: one landmark dimension needs at most P2_MAXE records of (DE + 1) DRM doubles, so d_e 3 with
// d_r up to 6 stages a full batch in 1536 doubles and four workgroups fit a CU's LDS.
template <int DE, int DRM>
constexpr int p2_stage_cap() {
    return DE > 0 && P2_MAXE * (DE + 1) * DRM < P2_STAGE ? P2_MAXE * (DE + 1) * DRM : P2_STAGE;
}

// 4 waves per SIMD (4 workgroups per CU) for d_e 3, d_r up to 6; at least 3 elsewhere where the row registers allow
template <int DE, int DRM, class P>
__global__ __launch_bounds__(P2_THREADS, (DE > 0 && DRM <= 6) ? 4 : (DRM <= 9 ? 3 : 2)) void k_phase2(P2Args<P> a) {
    using TC = typename P::TC;
    using TS = typename P::TS;
    constexpr int MERGE = P2_THREADS * (DRM + 1);
    // the update reads DE*DRM doubles of C_c2 (and a column at a row up to MAXDR) whatever d_r2 is: pad each
    // staging buffer by the largest such read
    constexpr int SCAP = p2_stage_cap<DE, DRM>();
    constexpr int STAGE = SCAP + 8 * (DRM > MAXDR ? DRM : MAXDR);
    constexpr int LDSN = 2 * STAGE > MERGE ? 2 * STAGE : MERGE;
    constexpr int DEC = DE ? DE : 8;   // register width of a C_c1 column
    constexpr int QN = P2Width<DE, DRM>::value;
    constexpr int GL = (DE > 0 && DRM <= 6) ? P2_G4 : P2_G;   // landmarks per C_c1 load group
    __shared__ double lds[LDSN];
    __shared__ P2Meta meta[3];
    __shared__ int s_slot_dr[64];

    // phase 1 failed: S and g_S stay untouched. The dispatcher deals workgroups to the 8 XCDs round robin, so
    // with items longest first every XCD gets a longest-first share of the work
    const int64_t idx = blockIdx.x;
    if (*a.flag == (unsigned long long)NO_FAILURE && idx < a.nitems) {
    const P2Item it = a.items[idx];
    const int t = threadIdx.x;
    const int lane = t & 63;
    const bool wave0 = t < 64;
    if (t < it.nslots) s_slot_dr[t] = a.slot_dr[it.slot_lo + t];
    int my_slot = -1, my_row = 0, chunk = 0, nchunk = 1;
    if (t < it.ntasks) {
        const P2Task tk = a.tasks[it.task_lo + t];
        my_slot = tk.slot;
        my_row = tk.row;
        chunk = tk.chunk;
        nchunk = tk.nchunk;
    }
    const bool diag = it.has_diag && my_slot == it.nslots - 1;
    const int dr2 = it.dr2;
    const int size = p2_record<DE>(DE, dr2);   // staged record (one landmark dimension)
    const int nbmax = DE > 0 ? (SCAP / size < P2_MAXE ? SCAP / size : P2_MAXE) : P2_MAXE;
    double acc[DRM];
    double gacc = 0.0;
    int64_t rowaddr = 0;
#pragma unroll
    for (int c = 0; c < DRM; c++) acc[c] = 0.0;
    if (my_slot >= 0) {
        rowaddr = it.base + a.slot_rowbase[it.slot_lo + my_slot] + my_row;
        if (chunk == 0 && it.part == 0) {
#pragma unroll
            for (int c = 0; c < DRM; c++)
                if (c < dr2 && (!diag || c >= my_row)) acc[c] = (double)a.S[rowaddr + c * it.colstride];
            if (diag) gacc = a.b_r[it.ro2 + my_row] * (a.s_r ? a.s_r[it.ro2 + my_row] : 1.0);
        }
    }
    const unsigned long long lowmask = my_slot > 0 ? ((1ull << my_slot) - 1ull) : 0ull;
    const int udr = it.uniform_dr;

    // prologue: describe the first two batches, stage the first
    int64_t start = it.entry_lo;
    P2Entry en;
    bool ev = p2_load_entries(a, start, it.entry_lo, it.entry_hi, lane, en);
    if (wave0) start += p2_process<DE, P>(a, it, meta[0], ev, en, lane, nbmax);
    ev = p2_load_entries(a, start, it.entry_lo, it.entry_hi, lane, en);
    if (wave0) start += p2_process<DE, P>(a, it, meta[1], ev, en, lane, nbmax);
    __syncthreads();
    double v[QN];
    if (meta[0].nb > 0) {
        p2_issue<DE, DRM, P>(a, meta[0], t, dr2, v);
        p2_store<DE, DRM>(meta[0], t, dr2, v, lds);
    }
    __syncthreads();
    int mc = 0, sb = 0;
    while (meta[mc].nb > 0) {
        const int mn = mc == 2 ? 0 : mc + 1;
        const int mnn = mn == 2 ? 0 : mn + 1;
        const bool have_next = meta[mn].nb > 0;
        ev = p2_load_entries(a, start, it.entry_lo, it.entry_hi, lane, en);   // wave 0 describes the batch after next
        if (have_next) p2_issue<DE, DRM, P>(a, meta[mn], t, dr2, v);
        const P2Meta &m = meta[mc];
        const double *stg = lds + sb * STAGE;
        unsigned long long bits = 0;
        if (my_slot >= 0) {
            const unsigned long long sbits = m.slotbits[my_slot];
            if (nchunk == 1) {
                bits = sbits;
            } else {
                const int h = __popcll(sbits);
                const int lo = chunk * h / nchunk, hi = (chunk + 1) * h / nchunk;
                bits = sbits & ~bits_below(nth_set_bit(sbits, lo)) & bits_below(nth_set_bit(sbits, hi));
            }
        }
        while (bits) {
            int jj[GL];
            const TC *cp[GL];
#pragma unroll
            for (int gi = 0; gi < GL; gi++) {
                const bool hv = bits != 0;
                const int j = hv ? __builtin_ctzll(bits) : 0;
                bits &= bits - 1ull;
                const int de = DE ? DE : m.de[j];
                int pre;
                if (udr) {
                    pre = __popcll(m.mask[j] & lowmask) * de * udr;
                } else {
                    pre = 0;
                    unsigned long long mm = m.mask[j] & lowmask;
                    while (mm) {
                        const int sl = __builtin_ctzll(mm);
                        mm &= mm - 1;
                        pre += de * s_slot_dr[sl];
                    }
                }
                jj[gi] = hv ? j : -1;
                cp[gi] = hv ? a.C + m.src_c1[j] + pre + my_row * de : a.C;
            }
            double cv[GL][DEC];
#pragma unroll
            for (int gi = 0; gi < GL; gi++) {
                if constexpr (DE > 0) {
#pragma unroll
                    for (int k = 0; k < DE; k++) cv[gi][k] = (double)cp[gi][k];
                } else {
                    const int de = jj[gi] >= 0 ? m.de[jj[gi]] : 1;
#pragma unroll
                    for (int k = 0; k < 8; k++) cv[gi][k] = (double)cp[gi][k < de ? k : 0];
                }
            }
#pragma unroll
            for (int gi = 0; gi < GL; gi++) {
                const int j = jj[gi];
                if (j < 0) continue;
                if constexpr (DE > 0) {
                    // S[row, c] -= C_c1[:, row] . W2[:, c] for every column (a column past d_r2, or below the
                    // diagonal of the diagonal block, collects values that are never written); g[row] -= gz[row]
                    // on the diagonal slot
                    const double *W2 = stg + j * size;
#pragma unroll
                    for (int c = 0; c < DRM; c++) {
                        double s = 0.0;
#pragma unroll
                        for (int k = 0; k < DE; k++) s += cv[gi][k] * W2[c * DE + k];
                        acc[c] -= s;
                    }
                    if (diag) gacc -= W2[DE * dr2 + my_row];
                } else {
                    const double *Ei = stg + m.off[j];
                    const int de = m.de[j];
                    const double *zz = Ei + de * de;
                    const double *C2 = zz + de;
                    double w[8];
#pragma unroll
                    for (int i = 0; i < 8; i++) {
                        double s = 0.0;
                        for (int k = 0; k < de; k++) s += Ei[i + k * de] * cv[gi][k];
                        w[i] = s;
                    }
#pragma unroll
                    for (int c = 0; c < DRM; c++) {
                        if (c < dr2 && (!diag || c >= my_row)) {
                            double s = 0.0;
                            for (int i = 0; i < de; i++) s += w[i] * C2[c * de + i];
                            acc[c] -= s;
                        }
                    }
                    if (diag) {
                        double s = 0.0;
                        for (int i = 0; i < de; i++) s += C2[my_row * de + i] * zz[i];
                        gacc -= s;
                    }
                }
            }
        }
        if (have_next) p2_store<DE, DRM>(meta[mn], t, dr2, v, lds + (sb ^ 1) * STAGE);
        if (wave0) start += p2_process<DE, P>(a, it, meta[mnn], ev, en, lane, nbmax);
        __syncthreads();
        mc = mn;
        sb ^= 1;
    }
    // merge the chunks of each (slot, row) in chunk order, then write S (upper triangle) and g, or this part's
    // row sums to the partial buffer
    if (my_slot >= 0) {
#pragma unroll
        for (int c = 0; c < DRM; c++) lds[t * (DRM + 1) + c] = acc[c];
        lds[t * (DRM + 1) + DRM] = gacc;
    }
    __syncthreads();
    if (my_slot >= 0 && chunk == 0) {
        const int dr1 = s_slot_dr[my_slot];
        for (int k = 1; k < nchunk; k++) {
            const int sib = t + k * dr1;
#pragma unroll
            for (int c = 0; c < DRM; c++) acc[c] += lds[sib * (DRM + 1) + c];
            gacc += lds[sib * (DRM + 1) + DRM];
        }
        if (it.nparts == 1) {
#pragma unroll
            for (int c = 0; c < DRM; c++)
                if (c < dr2 && (!diag || c >= my_row)) a.S[rowaddr + c * it.colstride] = (TS)acc[c];
            if (diag) a.g[it.ro2 + my_row] = gacc;
        } else {
            double *prow = a.partial + it.pbase + (int64_t)(a.slot_roff[it.slot_lo + my_slot] + my_row) * (dr2 + 1);
#pragma unroll
            for (int c = 0; c < DRM; c++)
                if (c < dr2) prow[c] = acc[c];
            prow[dr2] = gacc;
        }
    }
    }
    if (a.fin_counter) {
        __syncthreads();
        if (threadIdx.x == 0 && atomicAdd(a.fin_counter, 1u) == gridDim.x - 1) {   // every workgroup is done
            const unsigned long long f = *a.fin_flag;
            *a.fin_out = f == (unsigned long long)NO_FAILURE ? 0 : (int64_t)f;
            *a.fin_flag = (unsigned long long)NO_FAILURE;
            *a.fin_counter = 0u;
        }
    }
}

// Sum the parts of split items in part order and write S (upper triangle) and g: one workgroup per split item,
// one lane per row of its slots.
template <class TS>
__global__ void k_phase2_merge(const P2Merge *merges, const int64_t *slot_rowbase, const int32_t *slot_dr,
                               const double *partial, TS *S, double *g, const unsigned long long *flag) {
    if (*flag != (unsigned long long)NO_FAILURE) return;
    const P2Merge mg = merges[blockIdx.x];
    const int t = threadIdx.x;
    if (t >= mg.rows) return;
    int sl = 0, r0 = 0;
    while (r0 + slot_dr[mg.slot_lo + sl] <= t) r0 += slot_dr[mg.slot_lo + sl++];
    const int row = t - r0;
    const bool diag = mg.has_diag && sl == mg.nslots - 1;
    const int W = mg.dr2 + 1;
    const int64_t rowaddr = mg.base + slot_rowbase[mg.slot_lo + sl] + row;
    const int64_t pstride = (int64_t)mg.rows * W;
    for (int c = diag ? row : 0; c < mg.dr2; c++) {
        const double *P = partial + mg.pbase + (int64_t)t * W + c;
        double s = P[0];
        for (int p = 1; p < mg.nparts; p++) s += P[p * pstride];
        S[rowaddr + c * mg.colstride] = (TS)s;
    }
    if (diag) {
        const double *P = partial + mg.pbase + (int64_t)t * W + mg.dr2;
        double s = P[0];
        for (int p = 1; p < mg.nparts; p++) s += P[p * pstride];
        g[mg.ro2 + row] = s;
    }
}

// ---------------------------------------------------------------------------------------------------------
// reset_schur_* and copyto!
// ---------------------------------------------------------------------------------------------------------
template <class T>
__global__ void k_fill(T *p, int64_t n, T v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) p[i] = v;
}

template <class TS, class TR>
__global__ void k_scatter_r(TS *nzval, const TR *R, const int64_t *r_to_s, int64_t nR) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nR; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t idx = r_to_s[i];
        if (idx > 0) nzval[idx - 1] = (TS)R[i];
    }
}

template <class TS>
__global__ void k_add_damping(TS *nzval, const int64_t *diag_inds, int64_t nr, const double *scaling_r,
                              double lambda, const double *s_r) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nr) return;
    {
#pragma clang fp contract(off)
        const double si = s_r ? s_r[i] : 1.0;
        const int64_t k = diag_inds[i] - 1;
        nzval[k] = (TS)((double)nzval[k] + lambda * scaling_r[i] * (si * si));
    }
}

template <class TS>
__global__ void k_scale_sparse(TS *nzval, const int64_t *S_colptr, const int64_t *S_rowval, int64_t nr,
                               const double *s_r) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nr) return;
    for (int64_t k = S_colptr[j] - 1; k < S_colptr[j + 1] - 1; k++) nzval[k] = (TS)((double)nzval[k] * (s_r[S_rowval[k] - 1] * s_r[j]));
}

// reset_schur_sparse! in one pass over the stored entries (src/schur_reduction.jl:416-436): entry k takes R[q]
// (q + 1 = code[k] >> 1: the last R entry r_to_s_map sends there; 0 where none does), scaled(., s_r[row] s_r[col])
// when conditioned ((row, col) from rc, read only then), and the damping when it is a diagonal entry (code bit 0;
// indices nnz + i of the loop: diagonal entry i), with the same operations per entry as fill!, the scatter,
// scale_sparse! and the damping loop one after another.
template <class TS, class TR>
__global__ void k_reset_sparse1(TS *nzval, const TR *R, const int32_t *code, const int2 *rc, const int64_t *diag_inds,
                                int64_t nnz, int64_t nr, const double *scaling_r, double lambda, const double *s_r) {
    // blocks of one XCD take a contiguous range: the R entries an S column reads (R is stored the other way round)
    // are shared by the neighbouring S columns, in that XCD's L2
    const int64_t lb = xcd_remap(blockIdx.x, gridDim.x);
    for (int64_t x = lb * blockDim.x + threadIdx.x; x < nnz + nr; x += (int64_t)gridDim.x * blockDim.x) {
        if (x < nnz) {
            const int32_t c = code[x];
            if (c & 1) continue;   // a diagonal entry: below
            const int32_t q = (c >> 1) - 1;
            TS v = q >= 0 ? (TS)R[q] : (TS)0;
            if (s_r) {
                const int2 p = rc[x];
                v = (TS)((double)v * (s_r[p.x] * s_r[p.y]));
            }
            nzval[x] = v;
        } else {
            const int64_t i = x - nnz, k = diag_inds[i] - 1;
            const int32_t q = (code[k] >> 1) - 1;
            TS v = q >= 0 ? (TS)R[q] : (TS)0;
            if (s_r) v = (TS)((double)v * (s_r[i] * s_r[i]));
            {
#pragma clang fp contract(off)
                const double si = s_r ? s_r[i] : 1.0;
                nzval[k] = (TS)((double)v + lambda * scaling_r[i] * (si * si));
            }
        }
    }
}

template <class TS, class TR>
__global__ void k_reset_dense(TS *S, const TR *R, int64_t nr, const double *scaling_r, double lambda,
                              const double *s_r) {
    const int64_t n2 = nr * nr;
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < n2; k += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = k % nr, j = k / nr;
        TS v = s_r ? (TS)((double)R[k] * (s_r[i] * s_r[j])) : (TS)R[k];
        if (i == j) {
#pragma clang fp contract(off)
            const double si = s_r ? s_r[i] : 1.0;
            v = (TS)((double)v + lambda * scaling_r[i] * (si * si));
        }
        S[k] = v;
    }
}

template <class T>
__global__ void k_gather(T *dest, const T *src, const int64_t *map, int64_t n) {
    // XCD-contiguous ranges: neighbouring CSC columns gather from the same stored blocks (the lower triangle reads S
    // the other way round), in one XCD's L2
    const int64_t lb = xcd_remap(blockIdx.x, gridDim.x);
    for (int64_t i = lb * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dest[i] = src[map[i] - 1];
}

// backsubstitute! (src/schur_reduction.jl:541-576): one lane per eliminated block
constexpr int BS_LANES = 16;   // lanes per eliminated block in k_backsub
template <class TC, class TEI, class TSC>
__global__ void k_backsub(int64_t neb, const int32_t *e_dims, const int64_t *eoffsets, const int64_t *boffsets_e,
                          const int64_t *eliminated_ptr, const int32_t *entry_retained, const int64_t *coffsets,
                          const int32_t *r_dims, const int64_t *boffsets_r, const TC *C, const TEI *E_inv,
                          const double *z, const double *x_r, const double *s_r, const double *s_e, double *x_e) {
    // BS_LANES lanes per eliminated block: lane q adds the couplings q, q + BS_LANES, ... of the block; the lanes'
    // partials are added in a fixed tree, and the lanes i < d_e apply E_inv
    const int64_t e = (int64_t)blockIdx.x * (256 / BS_LANES) + threadIdx.x / BS_LANES;
    const int lane = threadIdx.x % BS_LANES;
    const bool valid = e < neb;
    const int de = valid ? e_dims[e] : 0;
    const int64_t eo = valid ? eoffsets[e] : 0, bo = valid ? boffsets_e[e] : 0;
    double t[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const int64_t cb0 = valid ? eliminated_ptr[e] : 0, cb1 = valid ? eliminated_ptr[e + 1] : 0;
    for (int64_t c = cb0 + lane; c < cb1; c += BS_LANES) {
        const int r = entry_retained[c];
        const int dr = r_dims[r];
        const int64_t ro = boffsets_r[r], co = coffsets[c];
        for (int ir = 0; ir < dr; ir++) {
            const double xc = x_r[ro + ir];
            const double f = s_r ? s_r[ro + ir] : 1.0;
            for (int i = 0; i < de; i++) {
                // scaled(T, C, s_e s_r) is held in TSC before it is promoted (src/conditioning.jl:222-235)
                const double cv = (s_r || s_e) ? (double)(TSC)((double)C[co + ir * de + i] * ((s_e ? s_e[bo + i] : 1.0) * f))
                                               : (double)C[co + ir * de + i];
                t[i] += cv * xc;
            }
        }
    }
#pragma unroll
    for (int i = 0; i < 8; i++)
#pragma unroll
        for (int o = BS_LANES / 2; o >= 1; o >>= 1) t[i] += __shfl_xor(t[i], o, 64);
    if (!valid || lane >= de) return;
    double acc = 0.0;   // lane i: row i of E_inv t
#pragma unroll
    for (int j = 0; j < 8; j++)
        if (j < de) acc += (double)E_inv[eo + j * de + lane] * t[j];
    x_e[bo + lane] = z[bo + lane] - acc;
}

// backsubstitute! (src/schur_reduction.jl:541-576) for layouts of uniform blocks (d_e = DE) with the default offsets
// whose landmarks couple contiguous runs of retained blocks: a landmark's coupling values then form one run in which
// value (i, q) — component i, row q of its window's scalar retained rows — sits at DE q + i, against x_r at the
// window's first row + q. 32 lanes per landmark, a lane per window row (coalesced over C and x_r), the lanes'
// partials added in a fixed tree, the lanes i < d_e applying E_inv.
template <int DE, class TC, class TEI, class TSC>
__global__ __launch_bounds__(256) void k_backsub_win(int64_t neb, const int64_t *__restrict__ eliminated_ptr,
                                                     const int32_t *__restrict__ entry_retained, int dr,
                                                     const TC *__restrict__ C, const TEI *__restrict__ E_inv,
                                                     const double *__restrict__ z, const double *__restrict__ x_r,
                                                     const double *__restrict__ s_r, const double *__restrict__ s_e,
                                                     double *__restrict__ x_e) {
    const int64_t e = (int64_t)blockIdx.x * 8 + threadIdx.x / 32;
    const int lane = threadIdx.x % 32;
    double t[DE];
#pragma unroll
    for (int i = 0; i < DE; i++) t[i] = 0.0;
    if (e < neb) {
        const int64_t c0 = eliminated_ptr[e], nq = (eliminated_ptr[e + 1] - c0) * dr;
        if (nq > 0) {
            const int64_t xb = (int64_t)entry_retained[c0] * dr;
            const TC *cp = C + c0 * DE * dr;
            double se[DE];
#pragma unroll
            for (int i = 0; i < DE; i++) se[i] = s_e ? s_e[e * DE + i] : 1.0;
            for (int64_t q = lane; q < nq; q += 32) {
                const double x = x_r[xb + q];
                if (s_r || s_e) {
                    // scaled(T, C, s_e s_r), held in TSC before it is promoted (src/conditioning.jl:222-235)
                    const double f = s_r ? s_r[xb + q] : 1.0;
#pragma unroll
                    for (int i = 0; i < DE; i++) t[i] += (double)(TSC)((double)cp[DE * q + i] * (se[i] * f)) * x;
                } else {
#pragma unroll
                    for (int i = 0; i < DE; i++) t[i] += (double)cp[DE * q + i] * x;
                }
            }
        }
    }
#pragma unroll
    for (int i = 0; i < DE; i++)
#pragma unroll
        for (int o = 16; o >= 1; o >>= 1) t[i] += __shfl_xor(t[i], o, 32);
    if (e >= neb || lane >= DE) return;
    double acc = 0.0;   // lane i: row i of E_inv t
#pragma unroll
    for (int j = 0; j < DE; j++) acc += (double)E_inv[e * DE * DE + j * DE + lane] * t[j];
    x_e[e * DE + lane] = z[e * DE + lane] - acc;
}

// Conditioning scales (src/conditioning.jl:99-141): Cs = the value the reference reads for every coupling entry,
// scaled(C_er[i, ir], s_e[i] s_r[ir]) = C * (s_e * s_r) in Float64, for phase 2 to use in place of C (phase 1 scales
// E and b_e itself). One lane per eliminated block.
template <class TC, class TSC>
__global__ void k_scale_c(int64_t neb, const int32_t *e_dims, const int64_t *boffsets_e, const int64_t *eliminated_ptr,
                          const int32_t *entry_retained, const int64_t *coffsets, const int32_t *r_dims,
                          const int64_t *boffsets_r, const TC *C, const double *s_r, const double *s_e, TSC *Cs) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= neb) return;
    const int de = e_dims[e];
    const int64_t bo = boffsets_e[e];
    for (int64_t c = eliminated_ptr[e]; c < eliminated_ptr[e + 1]; c++) {
        const int r = entry_retained[c];
        const int dr = r_dims[r];
        const int64_t ro = boffsets_r[r], co = coffsets[c];
        for (int ir = 0; ir < dr; ir++) {
            const double f = s_r ? s_r[ro + ir] : 1.0;
            for (int i = 0; i < de; i++) Cs[co + ir * de + i] = (TSC)((double)C[co + ir * de + i] * ((s_e ? s_e[bo + i] : 1.0) * f));
        }
    }
}

// ---------------------------------------------------------------------------------------------------------
// Blocks above the register path: eliminated blocks of d_e > 8 or retained blocks of d_r > MAXDR. The reference
// takes its run-time-size route for them (src/schur_reduction.jl:164-225 phase 1, :262-354 phase 2, :578-614
// backsubstitute). A handle holding any such block reduces its whole instance with these kernels ("wide" mode):
// phase 1 of every large block by one workgroup (Cholesky, pivot-ratio test and inverse through a global
// workspace), W = E_inv C per coupling as the reference forms W_data, every stored entry of S by one thread
// walking its block pair's (c1, c2) list in the reference's order, and g_S by one thread per retained row. Every
// sum is taken in a fixed order; no atomics on values.
// ---------------------------------------------------------------------------------------------------------
constexpr int WIDE_MAXDE = 4096;   // largest eliminated block (backsubstitute keeps t in LDS)
constexpr int WIDE_THREADS = 256;
constexpr int64_t WIDE_P1_GRID = 1024;   // workgroups per run of large blocks (each owns 2 d_e^2 of workspace)

struct WidePair {      // one stored block (r1, r2), r1 <= r2, of S with its coupling pairs cc[lo, hi)
    int32_t r1, r2;
    int64_t base;      // S index of the block's (row 0, col 0)
    int64_t stride;    // S index step from one scalar column of the block to the next
    int64_t lo, hi;
};

#pragma clang fp contract(off)
// Phase 1 of `count` blocks of dimension de > 8 (the run at first, E at eoff0 + b de^2, b_e at boff0 + b de):
// LAPACK potrf!('U') / pivot ratio / potri!('U') of src/schur_reduction.jl:164-225, in T for the Float32 variant
// (Float64 otherwise). U is formed right-looking: entry (i, j) receives its updates in ascending k, as the
// column-by-column form subtracts them, and is divided by U_ii last.
template <class P>
__global__ __launch_bounds__(WIDE_THREADS) void k_phase1_wide(
    int64_t first, int64_t count, int de, int64_t eoff0, int64_t boff0, const typename P::TE *__restrict__ E,
    const double *__restrict__ b_e, const double *__restrict__ scaling_e, double lambda,
    const double *__restrict__ s_e, typename P::TEI *__restrict__ E_inv, double *__restrict__ z,
    unsigned long long *flag, typename std::conditional<P::F32_FACTOR, float, double>::type *ws) {
    using F = typename std::conditional<P::F32_FACTOR, float, double>::type;
    using TEI = typename P::TEI;
    const double tol = 4504.0 * (P::F32_FACTOR ? 1.1920928955078125e-07 : 2.220446049250313e-16);
    const int t = threadIdx.x;
    const int64_t dd = (int64_t)de * de;
    F *U = ws + (int64_t)blockIdx.x * 2 * dd, *Ui = U + dd;
    __shared__ int bad;
    for (int64_t b = blockIdx.x; b < count; b += gridDim.x) {
        const int64_t eo = eoff0 + b * dd, bo = boff0 + b * de;
        for (int64_t q = t; q < dd; q += WIDE_THREADS) {
            const int i = (int)(q % de), j = (int)(q / de);
            const double si = s_e ? s_e[bo + i] : 1.0, sj = s_e ? s_e[bo + j] : 1.0;
            F v = s_e ? (F)((double)E[eo + q] * (si * sj)) : (F)E[eo + q];
            if (i == j) v = (F)((double)v + lambda * scaling_e[bo + i] * (si * si));
            U[q] = v;
        }
        if (t == 0) bad = 0;
        __syncthreads();
        for (int k = 0; k < de; k++) {
            if (t == 0) {
                const F a = U[(int64_t)k * de + k];
                if (!(a > 0)) bad = 1;   // potrf's info != 0
                else U[(int64_t)k * de + k] = sqrt(a);
            }
            __syncthreads();
            if (bad) break;   // read by every thread after the barrier: uniform
            const F ukk = U[(int64_t)k * de + k];
            for (int j = k + 1 + t; j < de; j += WIDE_THREADS) U[(int64_t)j * de + k] = U[(int64_t)j * de + k] / ukk;
            __syncthreads();
            const int64_t nt = de - k - 1;
            for (int64_t q = t; q < nt * nt; q += WIDE_THREADS) {
                const int64_t i = k + 1 + q % nt, j = k + 1 + q / nt;
                if (i <= j) U[j * de + i] = U[j * de + i] - U[i * de + k] * U[j * de + k];
            }
            __syncthreads();
        }
        if (!bad) {
            // every pivot safely positive against its damped diagonal entry, recomputed in Float64
            for (int i = t; i < de; i += WIDE_THREADS) {
                const double si = s_e ? s_e[bo + i] : 1.0;
                const double hii = (double)E[eo + (int64_t)i * (de + 1)] * (si * si) + lambda * scaling_e[bo + i] * (si * si);
                const double u = (double)U[(int64_t)i * de + i];
                const double ratio = (u * u) / hii;
                if (!(isfinite(ratio) && ratio > tol)) bad = 1;   // every writer stores 1
            }
        }
        __syncthreads();
        if (bad) {
            if (t == 0) atomicMin(flag, (unsigned long long)(first + b + 1));
        } else {
            // potri: Ui = U^-1 a column per thread, then the upper triangle of Ui Ui', mirrored
            for (int j = t; j < de; j += WIDE_THREADS) {
                Ui[(int64_t)j * de + j] = (F)1 / U[(int64_t)j * de + j];
                for (int i = j - 1; i >= 0; i--) {
                    F s = 0;
                    for (int k = i + 1; k <= j; k++) s += U[(int64_t)k * de + i] * Ui[(int64_t)j * de + k];
                    Ui[(int64_t)j * de + i] = -s / U[(int64_t)i * de + i];
                }
            }
            __syncthreads();
            for (int64_t q = t; q < dd; q += WIDE_THREADS) {
                const int i = (int)(q % de), j = (int)(q / de);
                if (i > j) continue;
                F s = 0;
                for (int k = j; k < de; k++) s += Ui[(int64_t)k * de + i] * Ui[(int64_t)k * de + j];
                E_inv[eo + (int64_t)j * de + i] = (TEI)s;
                E_inv[eo + (int64_t)i * de + j] = (TEI)s;
            }
            __syncthreads();
            // z_e = E_inv (b_e .* s_e), Float64 accumulation
            for (int i = t; i < de; i += WIDE_THREADS) {
                double acc = 0.0;
                for (int j = 0; j < de; j++)
                    acc += (double)E_inv[eo + (int64_t)j * de + i] * (b_e[bo + j] * (s_e ? s_e[bo + j] : 1.0));
                z[bo + i] = acc;
            }
        }
        __syncthreads();   // every thread has read `bad` before the next block resets it
    }
}

// W = E_inv C of every coupling (the reference's W_data, src/schur_reduction.jl:210-223), in Float64: one
// wavefront per coupling, lane q the entries q, q + 64, ... of its d_e x d_r block. C is the (scaled) coupling data.
template <class TEI, class TCS>
__global__ __launch_bounds__(WIDE_THREADS) void k_wide_w(int64_t ncoup, const int32_t *coup_e, const int32_t *e_dims,
                                                        const int64_t *eoffsets, const int64_t *coffsets,
                                                        const int32_t *entry_retained, const int32_t *r_dims,
                                                        const TEI *E_inv, const TCS *C, double *W) {
    const int64_t c = (int64_t)blockIdx.x * (WIDE_THREADS / 64) + threadIdx.x / 64;
    if (c >= ncoup) return;
    const int lane = threadIdx.x & 63;
    const int e = coup_e[c];
    const int de = e_dims[e], dr = r_dims[entry_retained[c]];
    const int64_t eo = eoffsets[e], co = coffsets[c];
    for (int64_t q = lane; q < (int64_t)de * dr; q += 64) {
        const int i = (int)(q % de);
        const int64_t col = co + (q - i);
        double acc = 0.0;
        for (int j = 0; j < de; j++) acc += (double)E_inv[eo + (int64_t)j * de + i] * (double)C[col + j];
        W[co + q] = acc;
    }
}

// S[r1, r2] -= W_c1' C_c2 over the block pair's coupling pairs, in the order the reference's column sweep visits
// them (src/schur_reduction.jl:262-354): one workgroup per stored block, one thread per entry (upper triangle of
// a diagonal block). Nothing is written when phase 1 failed.
template <class TS, class TCS>
__global__ __launch_bounds__(WIDE_THREADS) void k_phase2_wide(const WidePair *pairs, const int2 *cc,
                                                             const int32_t *coup_e, const int32_t *e_dims,
                                                             const int64_t *coffsets, const int32_t *r_dims,
                                                             const double *W, const TCS *C, TS *S,
                                                             const unsigned long long *flag) {
    if (*flag != (unsigned long long)NO_FAILURE) return;
    const WidePair p = pairs[blockIdx.x];
    const int dr1 = r_dims[p.r1], dr2 = r_dims[p.r2];
    for (int64_t q = threadIdx.x; q < (int64_t)dr1 * dr2; q += WIDE_THREADS) {
        const int row = (int)(q % dr1), col = (int)(q / dr1);
        if (p.r1 == p.r2 && row > col) continue;
        const int64_t idx = p.base + (int64_t)col * p.stride + row;
        double v = (double)S[idx];
        for (int64_t k = p.lo; k < p.hi; k++) {
            const int2 pc = cc[k];
            const int de = e_dims[coup_e[pc.x]];
            const double *w = W + coffsets[pc.x] + (int64_t)row * de;
            const TCS *cs = C + coffsets[pc.y] + (int64_t)col * de;
            double acc = 0.0;
            for (int i = 0; i < de; i++) acc += w[i] * (double)cs[i];
            v -= acc;
        }
        S[idx] = (TS)v;
    }
}

// g_S = b_r .* s_r - C' z over each retained block's couplings in the reference's order: one workgroup per
// retained block, one thread per row.
template <class TCS>
__global__ __launch_bounds__(WIDE_THREADS) void k_g_wide(const int64_t *retained_ptr, const int2 *rm_ce,
                                                        const int32_t *e_dims, const int64_t *boffsets_e,
                                                        const int64_t *coffsets, const int32_t *r_dims,
                                                        const int64_t *boffsets_r, const TCS *C, const double *z,
                                                        const double *b_r, const double *s_r, double *g,
                                                        const unsigned long long *flag) {
    if (*flag != (unsigned long long)NO_FAILURE) return;
    const int64_t r = blockIdx.x;
    const int dr = r_dims[r];
    const int64_t ro = boffsets_r[r];
    for (int ir = threadIdx.x; ir < dr; ir += WIDE_THREADS) {
        double v = b_r[ro + ir] * (s_r ? s_r[ro + ir] : 1.0);
        for (int64_t k = retained_ptr[r]; k < retained_ptr[r + 1]; k++) {
            const int2 ce = rm_ce[k];
            const int de = e_dims[ce.y];
            const int64_t bo = boffsets_e[ce.y];
            const TCS *cs = C + coffsets[ce.x] + (int64_t)ir * de;
            double acc = 0.0;
            for (int i = 0; i < de; i++) acc += (double)cs[i] * z[bo + i];
            v -= acc;
        }
        g[ro + ir] = v;
    }
}

// backsubstitute! for any block size: one workgroup per eliminated block, t = sum_c C_c x_r in LDS (Float64), then
// x_e = z - E_inv t; blocks above 8 subtract term by term as the reference's run-time-size loop does
// (src/schur_reduction.jl:578-614), smaller ones subtract the product once (its static path).
template <class TC, class TEI, class TSC>
__global__ __launch_bounds__(WIDE_THREADS) void k_backsub_wide(int64_t neb, const int32_t *e_dims,
                                                              const int64_t *eoffsets, const int64_t *boffsets_e,
                                                              const int64_t *eliminated_ptr,
                                                              const int32_t *entry_retained, const int64_t *coffsets,
                                                              const int32_t *r_dims, const int64_t *boffsets_r,
                                                              const TC *C, const TEI *E_inv, const double *z,
                                                              const double *x_r, const double *s_r,
                                                              const double *s_e, double *x_e) {
    extern __shared__ double tsh[];
    for (int64_t e = blockIdx.x; e < neb; e += gridDim.x) {
        const int de = e_dims[e];
        const int64_t eo = eoffsets[e], bo = boffsets_e[e];
        for (int i = threadIdx.x; i < de; i += WIDE_THREADS) {
            const double se = s_e ? s_e[bo + i] : 1.0;
            double t = 0.0;
            for (int64_t c = eliminated_ptr[e]; c < eliminated_ptr[e + 1]; c++) {
                const int r = entry_retained[c];
                const int dr = r_dims[r];
                const int64_t ro = boffsets_r[r], co = coffsets[c];
                for (int ir = 0; ir < dr; ir++) {
                    const double f = s_r ? s_r[ro + ir] : 1.0;
                    const double cv = (s_r || s_e) ? (double)(TSC)((double)C[co + (int64_t)ir * de + i] * (se * f))
                                                   : (double)C[co + (int64_t)ir * de + i];
                    t += cv * x_r[ro + ir];
                }
            }
            tsh[i] = t;
        }
        __syncthreads();
        for (int i = threadIdx.x; i < de; i += WIDE_THREADS) {
            double acc;
            if (de > 8) {
                acc = z[bo + i];
                for (int j = 0; j < de; j++) acc -= (double)E_inv[eo + (int64_t)j * de + i] * tsh[j];
            } else {
                double m = 0.0;
                for (int j = 0; j < de; j++) m += (double)E_inv[eo + (int64_t)j * de + i] * tsh[j];
                acc = z[bo + i] - m;
            }
            x_e[bo + i] = acc;
        }
        __syncthreads();
    }
}
#pragma clang fp contract(fast)

// The retained rows' part of the full system's residual: out[r] = (b_r .* s_r)[r] - sum over the couplings c of
// retained block r (ascending eliminated block) of scaled(C_c)' x_e[e(c)] -- with x_e = backsubstitute!(x_r)
// this is g_S - (S - R) x_r of the reduction; with x_e = z it is g_S itself. One workgroup per retained block;
// thread t sums a contiguous share of the block's couplings (given as (coupling, eliminated block) pairs in the
// retained-major order, one load per coupling before its data), and the shares are added in a fixed tree (no
// atomics; bitwise reproducible). DE, DR > 0: every block has those dimensions and the offsets are c DE DR, e DE.
constexpr int CTX_THREADS = 256;
template <class TC, int DE, int DR>
__global__ __launch_bounds__(CTX_THREADS) void k_ct_xe(const int64_t *retained_ptr, const int2 *rm_ce,
                                                       const int32_t *e_dims, const int64_t *boffsets_e,
                                                       const int64_t *coffsets, const int32_t *r_dims,
                                                       const int64_t *boffsets_r, const TC *C, const double *x_e,
                                                       const double *b_r, const double *s_r, const double *s_e,
                                                       double *out) {
    constexpr int NR = DR > 0 ? DR : MAXDR;
    __shared__ double red[CTX_THREADS / 64][NR];
    const int r = blockIdx.x, t = threadIdx.x;
    const int dr = DR > 0 ? DR : r_dims[r];
    const int64_t ro = boffsets_r[r];
    const int64_t k0 = retained_ptr[r], n = retained_ptr[r + 1] - k0;
    const int64_t per = (n + CTX_THREADS - 1) / CTX_THREADS;
    const int64_t lo = k0 + (int64_t)t * per, hi = lo + per < k0 + n ? lo + per : k0 + n;
    double acc[NR];
#pragma unroll
    for (int i = 0; i < NR; i++) acc[i] = 0.0;
    for (int64_t k = lo; k < hi; k++) {
        const int2 ce = rm_ce[k];
        if constexpr (DE > 0 && DR > 0) {
            const TC *Cb = C + (int64_t)ce.x * (DE * DR);
            const double *xp = x_e + (int64_t)ce.y * DE;
            double cv[DE * DR], xe[DE];
#pragma unroll
            for (int q = 0; q < DE * DR; q++) cv[q] = (double)Cb[q];
#pragma unroll
            for (int i = 0; i < DE; i++) xe[i] = xp[i] * (s_e ? s_e[(int64_t)ce.y * DE + i] : 1.0);
#pragma unroll
            for (int ir = 0; ir < DR; ir++) {
                double sum = 0.0;
#pragma unroll
                for (int i = 0; i < DE; i++) sum += cv[ir * DE + i] * xe[i];
                acc[ir] += sum;
            }
        } else {
            const int de = e_dims[ce.y];
            const int64_t bo = boffsets_e[ce.y], co = coffsets[ce.x];
            double xe[8];
            for (int i = 0; i < 8; i++) xe[i] = i < de ? x_e[bo + i] * (s_e ? s_e[bo + i] : 1.0) : 0.0;
#pragma unroll
            for (int ir = 0; ir < NR; ir++) {
                if (ir < dr) {
                    double sum = 0.0;
                    for (int i = 0; i < de; i++) sum += (double)C[co + (int64_t)ir * de + i] * xe[i];
                    acc[ir] += sum;
                }
            }
        }
    }
    // fixed-order sums: a butterfly within each wavefront, then the wavefronts in order
#pragma unroll
    for (int ir = 0; ir < NR; ir++) {
        double v = acc[ir];
        for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
        if ((t & 63) == 0) red[t >> 6][ir] = v;
    }
    __syncthreads();
    if (t < dr) {
        double v = red[0][t];
        for (int w = 1; w < CTX_THREADS / 64; w++) v += red[w][t];
        const double sr = s_r ? s_r[ro + t] : 1.0;
        out[ro + t] = b_r[ro + t] * sr - sr * v;
    }
}


// ---------------------------------------------------------------------------------------------------------
// reduce_schur_dense! as a SYRK on the matrix cores. With E_inv_e = M M' (M lower, Cholesky of the landmark's
// inverse), C_e' E_inv_e C_e = Y_e' Y_e for Y_e = M' C_e (d_e rows, the landmark's retained columns), so the
// phase-2 update of the dense S is S -= Y' Y over the ne x nr matrix Y whose rows are the landmarks' Y_e. Y is
// built (zeros outside each landmark's retained blocks), the upper 128 x 128 tiles of Y' Y are formed on MFMA
// in K slices (rows of Y) whose partial tiles accumulate in T registers, and the slices are added in Float64, in
// slice order, into S (no atomics; bitwise reproducible).
// ---------------------------------------------------------------------------------------------------------
typedef float sy_v4f __attribute__((ext_vector_type(4)));
typedef double sy_v4d __attribute__((ext_vector_type(4)));
constexpr int SY_T = 128;    // output tile edge
constexpr int SY_KT = 16;    // rows of Y per LDS stage

template <class T> struct SyMfma;
template <> struct SyMfma<float> {
    using acc_t = sy_v4f;
    static __device__ inline acc_t step(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    // v_mfma_f32_16x16x4f32: lane l holds D[4 (l / 16) + v][l % 16]
    static __device__ inline int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};
template <> struct SyMfma<double> {
    using acc_t = sy_v4d;
    static __device__ inline acc_t step(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // v_mfma_f64_16x16x4f64: lane l holds D[l / 16 + 4 v][l % 16]
    static __device__ inline int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};

// One thread per landmark: M = chol(E_inv_e) (Float64), then Y[bo_e + i][col] = (M' C_e)[i][col] for the columns
// of each coupled retained block, rounded to TY.
template <class TC, class TY>
__global__ void k_ybuild(int64_t neb, const int32_t *e_dims, const int64_t *eoffsets, const int64_t *boffsets_e,
                         const int64_t *eliminated_ptr, const int32_t *entry_retained, const int64_t *coffsets,
                         const int32_t *r_dims, const int64_t *boffsets_r, const double *E_inv, const TC *C,
                         int64_t nr, TY *Y) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= neb) return;
    const int de = e_dims[e];
    const double *Ei = E_inv + eoffsets[e];
    double M[8][8];
    for (int j = 0; j < de; j++) {
        for (int i = j; i < de; i++) {
            double v = Ei[i + j * de];
            for (int k = 0; k < j; k++) v -= M[i][k] * M[j][k];
            if (i == j) M[j][j] = v > 0 ? sqrt(v) : 0.0;
            else M[i][j] = M[j][j] > 0 ? v / M[j][j] : 0.0;
        }
    }
    const int64_t row0 = boffsets_e[e];
    for (int64_t c = eliminated_ptr[e]; c < eliminated_ptr[e + 1]; c++) {
        const int r = entry_retained[c];
        const int dr = r_dims[r];
        const int64_t ro = boffsets_r[r], co = coffsets[c];
        for (int ir = 0; ir < dr; ir++) {
            const TC *col = C + co + (int64_t)ir * de;
            for (int i = 0; i < de; i++) {
                double v = 0.0;
                for (int j = i; j < de; j++) v += M[j][i] * (double)col[j];
                Y[(row0 + i) * nr + ro + ir] = (TY)v;
            }
        }
    }
}

// Partial tile (p, t) = sum over rows k of slice p of Y[k, I]' Y[k, J] for the upper tile t = (I, J) of S; four
// wavefronts, each a 64 x 64 quarter in 4 x 4 MFMA 16 x 16 x 4 tiles (a diagonal tile's lower-left quarter is
// skipped); Y's rows staged through LDS 16 at a time, the next stage's loads in registers meanwhile.
template <class T>
__global__ __launch_bounds__(256) void k_syrk(const T *Y, int64_t nr, int64_t K, int64_t kc, const int2 *tiles,
                                              int ntiles, T *partial) {
    using M = SyMfma<T>;
    __shared__ T As[SY_KT][SY_T + 4], Bs[SY_KT][SY_T + 4];
    const int t = blockIdx.x, p = blockIdx.y;
    const int2 tl = tiles[t];
    const int64_t I0 = (int64_t)tl.x * SY_T, J0 = (int64_t)tl.y * SY_T;
    const bool diag = tl.x == tl.y;
    const int64_t k_lo = (int64_t)p * kc, k_hi = k_lo + kc < K ? k_lo + kc : K;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, li = lane & 15, lk = lane >> 4;
    const int qi = wave & 1, qj = wave >> 1;
    const bool active = !(diag && qi > qj);
    typename M::acc_t acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] = typename M::acc_t{0, 0, 0, 0};
    const int lr = tid >> 4, lc = (tid & 15) * 8;
    T pa[8], pb[8];
    auto load = [&](int64_t k0) {
        const int64_t k = k0 + lr;
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const int64_t ca = I0 + lc + q, cb = J0 + lc + q;
            pa[q] = (k < k_hi && ca < nr) ? Y[k * nr + ca] : (T)0;
            pb[q] = (k < k_hi && cb < nr) ? Y[k * nr + cb] : (T)0;
        }
    };
    if (k_lo < k_hi) load(k_lo);
    for (int64_t k0 = k_lo; k0 < k_hi; k0 += SY_KT) {
#pragma unroll
        for (int q = 0; q < 8; q++) {
            As[lr][lc + q] = pa[q];
            Bs[lr][lc + q] = pb[q];
        }
        __syncthreads();
        if (k0 + SY_KT < k_hi) load(k0 + SY_KT);
        if (active) {
#pragma unroll
            for (int kk = 0; kk < SY_KT / 4; kk++) {
                T a[4], b[4];
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    a[i] = As[4 * kk + lk][64 * qi + 16 * i + li];
                    b[i] = Bs[4 * kk + lk][64 * qj + 16 * i + li];
                }
#pragma unroll
                for (int i = 0; i < 4; i++)
#pragma unroll
                    for (int j = 0; j < 4; j++) acc[i][j] = M::step(a[i], b[j], acc[i][j]);
            }
        }
        __syncthreads();
    }
    T *P = partial + ((int64_t)p * ntiles + t) * (SY_T * SY_T);
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int row = 64 * qi + 16 * i + M::row(lane, v), col = 64 * qj + 16 * j + li;
                P[row * SY_T + col] = acc[i][j][v];
            }
}

// S[a, b] -= sum over slices p (in order, Float64) of partial[p][t][a - I0][b - J0], for a <= b.
template <class T>
__global__ void k_syrk_sum(const T *partial, int nslices, int ntiles, const int2 *tiles, int64_t nr, double *S) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (int64_t)ntiles * SY_T * SY_T) return;
    const int t = (int)(idx / (SY_T * SY_T));
    const int el = (int)(idx % (SY_T * SY_T)), row = el / SY_T, col = el % SY_T;
    const int64_t a = (int64_t)tiles[t].x * SY_T + row, b = (int64_t)tiles[t].y * SY_T + col;
    if (a >= nr || b >= nr || a > b) return;
    double s = 0.0;
    for (int p = 0; p < nslices; p++) s += (double)partial[((int64_t)p * ntiles + t) * (SY_T * SY_T) + el];
    S[a + b * nr] -= s;
}

struct Run {
    int64_t first, count, de, eoff0, boff0;
    int64_t wsoff = -1;   // d_e > 8: offset (doubles) of the run's phase-1 workspace in d_wide_ws
    int64_t wsgrid = 0;   // ... and its workgroups
};

}  // namespace

// ---------------------------------------------------------------------------------------------------------
// The handle
// ---------------------------------------------------------------------------------------------------------
struct bls_schur {
    int64_t nr = 0, nrb = 0, neb = 0, ncoup = 0;
    std::vector<int64_t> r_dims, e_dims, boffsets_r, boffsets_e, eoffsets, eliminated_ptr, entry_retained, coffsets,
        retained_ptr, retained_entry;   // host, 0-based values
    std::vector<Run> runs;
    int uniform_de = 0;
    int mode = 0;   // 1 sparse, 2 dense
    int64_t nnz = 0, nR = 0;
    // device
    DevArray<int64_t> d_eoffsets, d_boffsets_e, d_coffsets, d_eliminated_ptr, d_boffsets_r;
    DevArray<int32_t> d_e_dims, d_r_dims, d_entry_retained;
    DevArray<int64_t> d_r_to_s, d_diag_inds, d_S_colptr, d_S_rowval;
    // the one-pass reset: per stored entry its R entry (-1 none) and its (row, column); empty when the registered
    // diag_inds are not the entries with row == column (the reset then runs fill, scatter, scale, damping)
    DevArray<int32_t> d_rst_code;   // (R entry + 1) << 1 | diagonal
    DevArray<int2> d_rst_rc;
    DevArray<P2Item> d_items;
    DevArray<P2Entry> d_entries;
    DevArray<P2Task> d_tasks;
    DevArray<int32_t> d_slot_dr, d_slot_roff;
    DevArray<P2Merge> d_merges;
    DevArray<double> d_partial;   // row sums of the parts of split items
    DevArray<double> d_Cs;        // scaled coupling blocks (in the variant's C type), when conditioning scales are given
    int64_t clen = 0;             // length of C_data
    int64_t nmerges = 0;
    int uniform_dr_all = 0;
    // d_e and d_r when every block has one size (d_e <= 4), the storage offsets are the default ones and every
    // landmark couples a contiguous run of retained blocks (k_backsub_win), else 0
    int win_de = 0, win_dr = 0;
    int max_dr = 0;
    DevArray<int64_t> d_slot_rowbase;
    DevArray<int64_t> d_retained_ptr;      // retained-major index of the couplings (0-based), for k_ct_xe
    DevArray<int2> d_rm_ce;                // ... its (coupling, eliminated block) pairs
    // the dense reduction by MFMA SYRK: Y (ne x nr, Float32 or Float64 as bytes), the upper tiles of S, the
    // split-K partial tiles
    DevArray<double> d_sy_Y, d_sy_partial;
    DevArray<int2> d_sy_tiles;
    int64_t sy_ntiles = 0;
    DevArray<unsigned long long> d_flag;
    DevArray<unsigned int> d_fin_counter;   // phase 2's finished workgroups (the folded verdict), 0 between calls
    bool flag_dirty = true;   // d_flag may not hold NO_FAILURE (before the first reduction, or after one that errored)
    DevArray<int64_t> d_failed_out;
    int64_t nitems = 0, nentries = 0;
    Timing timing;
    // the shape of the phase-2 work list (bls_schur_set_tuning / bls_schur_autotune_f64) and the registered sparse
    // layout it is rebuilt from (0-based block pattern, 1-based S_colptr, col_heights)
    int64_t p2_part = P2_PART, p2_lanes = P2_BASE_LANES;
    std::vector<int64_t> lay_bcp, lay_brv, lay_bro, lay_Scp, lay_ch;
    // blocks above the register path (d_e > 8 or d_r > MAXDR): the whole instance takes the run-time-size kernels
    bool wide = false;
    int wide_maxde = 0;
    DevArray<double> d_wide_ws;      // phase-1 workspace of the large blocks (U and U^-1 per workgroup)
    DevArray<double> d_W;            // W = E_inv C per coupling
    DevArray<int32_t> d_coup_e;      // eliminated block of each coupling
    DevArray<WidePair> d_wpairs;     // stored blocks of S with coupling pairs
    DevArray<int2> d_wcc;            // their (c1, c2) lists
    int64_t nwpairs = 0;
    // phase 2 on 16 x 16 tiles of S (matrix cores), for sparse layouts whose structure fits (bls_tiles.hip); p2_mode:
    // 0 the tiles when they fit, else the owner-computes work list; 1 always the work list; 2 the tiles (an error
    // when they do not fit)
    TilePlan tiles;
    std::string tiles_why;
    int p2_mode = 0;
    bool use_tiles() const { return mode == 1 && !wide && tiles.nunits > 0 && p2_mode != 1; }
};

namespace {

int check_structure(bls_schur *h) {
    for (int64_t e = 0; e < h->neb; e++) {
        if (h->e_dims[e] < 1 || h->e_dims[e] > WIDE_MAXDE)
            return fail(BLS_ERR_UNSUPPORTED, "eliminated block " + std::to_string(e + 1) + " has dimension " +
                                                 std::to_string(h->e_dims[e]) + "; the device path supports 1 to " +
                                                 std::to_string(WIDE_MAXDE));
        if (h->e_dims[e] > 8) h->wide = true;
        h->wide_maxde = std::max(h->wide_maxde, (int)h->e_dims[e]);
    }
    for (int64_t r = 0; r < h->nrb; r++) {
        if (h->r_dims[r] < 1 || h->r_dims[r] > INT32_MAX / 2)
            return fail(BLS_ERR_ARG, "retained block " + std::to_string(r + 1) + " has dimension " +
                                         std::to_string(h->r_dims[r]));
        if (h->r_dims[r] > MAXDR) h->wide = true;
    }
    return BLS_OK;
}

// Wide mode: the stored blocks of S that coupling pairs reach, each with its (c1, c2) list in the order of the
// reference's column sweep (src/schur_reduction.jl:300-312): block column r2 ascending, its couplings c2 in
// retained_entry order, and for each the couplings c1 <= c2 of the same eliminated block. `locate(r1, r2, base,
// stride)` gives a block's addressing in S, false when the pattern does not store it.
template <typename Locate>
int build_wide_pairs(bls_schur *h, Locate locate) {
    struct Key {
        int64_t key, seq;
        int32_t c1, c2;
    };
    std::vector<int32_t> coup_e((size_t)h->ncoup);
    for (int64_t e = 0; e < h->neb; e++)
        for (int64_t c = h->eliminated_ptr[e]; c < h->eliminated_ptr[e + 1]; c++) coup_e[c] = (int32_t)e;
    std::vector<Key> v;
    int64_t seq = 0;
    for (int64_t r2 = 0; r2 < h->nrb; r2++)
        for (int64_t k = h->retained_ptr[r2]; k < h->retained_ptr[r2 + 1]; k++) {
            const int64_t c2 = h->retained_entry[k], e = coup_e[c2];
            for (int64_t c1 = h->eliminated_ptr[e]; c1 <= c2; c1++) {
                const int64_t r1 = h->entry_retained[c1];
                if (r1 > r2) return fail(BLS_ERR_ARG, "couplings of an eliminated block are not in ascending retained order");
                v.push_back(Key{r2 * h->nrb + r1, seq++, (int32_t)c1, (int32_t)c2});
            }
        }
    std::sort(v.begin(), v.end(), [](const Key &a, const Key &b) { return a.key != b.key ? a.key < b.key : a.seq < b.seq; });
    std::vector<WidePair> pairs;
    std::vector<int2> cc(v.size());
    for (size_t i = 0; i < v.size();) {
        size_t j = i;
        while (j < v.size() && v[j].key == v[i].key) j++;
        WidePair p{};
        p.r2 = (int32_t)(v[i].key / h->nrb);
        p.r1 = (int32_t)(v[i].key % h->nrb);
        if (!locate(p.r1, p.r2, p.base, p.stride))
            return fail(BLS_ERR_ARG, "the Schur pattern has no block (" + std::to_string(p.r1 + 1) + ", " +
                                         std::to_string(p.r2 + 1) + ")");
        p.lo = (int64_t)i;
        p.hi = (int64_t)j;
        pairs.push_back(p);
        for (size_t q = i; q < j; q++) cc[q] = make_int2(v[q].c1, v[q].c2);
        i = j;
    }
    int rc;
    if ((rc = h->d_wpairs.upload(pairs)) || (rc = h->d_wcc.upload(cc)) || (rc = h->d_coup_e.upload(coup_e)) ||
        (rc = h->d_W.alloc((size_t)h->clen)))
        return rc;
    h->nwpairs = (int64_t)pairs.size();
    h->nitems = 0;
    h->nmerges = 0;
    return BLS_OK;
}

// One thread's share of the work list: block columns [c0, c1), with offsets local to it (fixed when the shares are
// joined in column order).
struct P2Share {
    std::vector<P2Item> items;
    std::vector<P2Entry> entries;
    std::vector<P2Task> tasks;
    std::vector<int64_t> slot_rowbase;
    std::vector<int32_t> slot_dr, slot_roff;
    std::vector<P2Merge> merges;
    int64_t npartial = 0;
    std::string error;
};

template <typename Slots, typename RowBase, typename Base>
void build_phase2_columns(const bls_schur *h, const std::vector<int32_t> &owner, int64_t c0, int64_t c1,
                          Slots slots_of, RowBase rowbase_of, Base base_of, P2Share &out) {
    std::vector<P2Item> &items = out.items;
    std::vector<P2Entry> &entries = out.entries;
    std::vector<P2Task> &tasks = out.tasks;
    std::vector<int64_t> &slot_rowbase = out.slot_rowbase;
    std::vector<int32_t> &slot_dr = out.slot_dr, &slot_roff = out.slot_roff;
    std::vector<P2Merge> &merges = out.merges;
    int64_t &npartial = out.npartial;
    const int64_t *er = h->entry_retained.data();
    std::vector<int64_t> col;
    std::vector<int32_t> pos((size_t)h->nrb, -1);   // slot of each block row in the current column, -1 none
    std::vector<int32_t> item_of;                   // item (within the column) of each slot
    std::vector<std::vector<P2Entry>> item_entries;
    std::vector<int64_t> hits(64);                  // landmarks touching each slot of the current item
    for (int64_t r2 = c0; r2 < c1; r2++) {
        col.clear();
        slots_of(r2, col);
        if (col.empty() || col.back() != r2) {
            out.error = "the reduced-system pattern has no diagonal block " + std::to_string(r2 + 1);
            return;
        }
        const int dr2 = (int)h->r_dims[r2];
        int64_t base, colstride;
        base_of(r2, base, colstride);
        // the column's items: runs of up to 64 slots and about p2_lanes rows
        std::vector<std::pair<size_t, size_t>> runs;
        item_of.assign(col.size(), 0);
        for (size_t k0 = 0; k0 < col.size();) {
            size_t k1 = k0;
            int rows = 0;
            while (k1 < col.size() && k1 - k0 < 64 && (k1 == k0 || rows + h->r_dims[col[k1]] <= h->p2_lanes))
                rows += (int)h->r_dims[col[k1++]];
            for (size_t k = k0; k < k1; k++) item_of[k] = (int32_t)runs.size();
            runs.emplace_back(k0, k1);
            k0 = k1;
        }
        for (size_t k = 0; k < col.size(); k++) pos[col[k]] = (int32_t)k;
        item_entries.resize(runs.size());
        for (auto &v : item_entries) v.clear();
        // each landmark of the column once: its couplings with block rows <= r2 (ascending), one entry per item they
        // fall in, in the column's landmark order
        for (int64_t kk = h->retained_ptr[r2]; kk < h->retained_ptr[r2 + 1]; kk++) {
            const int64_t c2 = h->retained_entry[kk];
            const int64_t e = owner[c2];
            for (int64_t c = h->eliminated_ptr[e]; c <= c2;) {
                const int32_t k = pos[er[c]];
                if (k < 0) {
                    out.error = "The sparse pattern has no block (" + std::to_string(er[c] + 1) + ", " +
                                std::to_string(r2 + 1) + ")";
                    return;
                }
                const int32_t it = item_of[k];
                const size_t k0 = runs[it].first, k1 = runs[it].second;
                P2Entry en{};
                en.e = (int32_t)e;
                en.c_first = (int32_t)c;
                en.c2 = (int32_t)c2;
                unsigned long long mask = 0;
                for (; c <= c2; c++) {
                    const int32_t kc = pos[er[c]];
                    if (kc < 0) {
                        out.error = "The sparse pattern has no block (" + std::to_string(er[c] + 1) + ", " +
                                    std::to_string(r2 + 1) + ")";
                        return;
                    }
                    if ((size_t)kc < k0 || (size_t)kc >= k1) break;
                    mask |= 1ull << (kc - (int32_t)k0);
                }
                en.mask = mask;
                item_entries[it].push_back(en);
            }
        }
        for (size_t k = 0; k < col.size(); k++) pos[col[k]] = -1;
        for (size_t it_i = 0; it_i < runs.size(); it_i++) {
            const size_t k0 = runs[it_i].first, k1 = runs[it_i].second;
            int rows = 0;
            for (size_t k = k0; k < k1; k++) rows += (int)h->r_dims[col[k]];
            P2Item it{};
            it.base = base;
            it.colstride = colstride;
            it.r2 = (int32_t)r2;
            it.dr2 = dr2;
            it.ro2 = (int32_t)h->boffsets_r[r2];
            it.slot_lo = (int32_t)slot_dr.size();
            it.nslots = (int32_t)(k1 - k0);
            it.has_diag = k1 == col.size();
            int udr = (int)h->r_dims[col[k0]];
            int roff = 0;
            for (size_t k = k0; k < k1; k++) {
                slot_dr.push_back((int32_t)h->r_dims[col[k]]);
                slot_roff.push_back(roff);
                slot_rowbase.push_back(rowbase_of(r2, k, col[k]));
                roff += (int)h->r_dims[col[k]];
                if (h->r_dims[col[k]] != udr) udr = 0;
            }
            it.uniform_dr = udr;
            const int64_t e_begin = (int64_t)entries.size();
            entries.insert(entries.end(), item_entries[it_i].begin(), item_entries[it_i].end());
            const int64_t e_end = (int64_t)entries.size();
            const int64_t ne = e_end - e_begin;
            const int np = ne > h->p2_part ? (int)((ne + h->p2_part - 1) / h->p2_part) : 1;
            it.nparts = np;
            if (np > 1) {
                P2Merge mg{};
                mg.base = base;
                mg.colstride = colstride;
                mg.pbase = npartial;
                mg.slot_lo = it.slot_lo;
                mg.nslots = it.nslots;
                mg.rows = rows;
                mg.dr2 = dr2;
                mg.ro2 = it.ro2;
                mg.has_diag = it.has_diag;
                mg.nparts = np;
                merges.push_back(mg);
            }
            for (int part = 0; part < np; part++) {
                P2Item pt = it;
                pt.part = part;
                pt.entry_lo = e_begin + part * ne / np;
                pt.entry_hi = e_begin + (part + 1) * ne / np;
                pt.pbase = np > 1 ? npartial + (int64_t)part * rows * (dr2 + 1) : 0;
                std::fill(hits.begin(), hits.end(), 0);
                for (int64_t x = pt.entry_lo; x < pt.entry_hi; x++)
                    for (unsigned long long mm = entries[x].mask; mm; mm &= mm - 1) hits[__builtin_ctzll(mm)]++;
                // chunks per slot: greedily give another lane group to the slot with the most hits per chunk
                const int ns = pt.nslots;
                std::vector<int> nch(ns, 1);
                int lanes = rows;
                while (true) {
                    int best = -1;
                    double bw = 0.0;
                    for (int sl = 0; sl < ns; sl++) {
                        const double w = (double)hits[sl] / nch[sl];
                        if (w > bw) { bw = w; best = sl; }
                    }
                    if (best < 0 || bw < 2.0) break;
                    if (lanes + slot_dr[pt.slot_lo + best] > P2_THREADS || nch[best] >= 32767) break;
                    nch[best]++;
                    lanes += slot_dr[pt.slot_lo + best];
                }
                pt.task_lo = (int32_t)tasks.size();
                for (int sl = 0; sl < ns; sl++) {
                    const int dr1 = slot_dr[pt.slot_lo + sl];
                    for (int ch = 0; ch < nch[sl]; ch++) {
                        for (int row = 0; row < dr1; row++) {
                            P2Task tk{};
                            tk.slot = (int16_t)sl;
                            tk.row = (int16_t)row;
                            tk.chunk = (int16_t)ch;
                            tk.nchunk = (int16_t)nch[sl];
                            tasks.push_back(tk);
                        }
                    }
                }
                pt.ntasks = (int32_t)(tasks.size() - pt.task_lo);
                items.push_back(pt);
            }
            if (np > 1) npartial += (int64_t)np * rows * (dr2 + 1);
        }
    }
}

// Build the phase-2 work: items (slot runs of block columns), their landmark entries, and the lanes' tasks.
// `slots_of(r2)` gives the block rows of column r2 (ascending, last == r2), `rowbase_of(r2, k, r1)` the row base of
// its k-th block, and base_of the column's addressing. An item with more than P2_PART landmarks is split into
// parts over consecutive landmark ranges (one workgroup each, summed in part order by k_phase2_merge). Within a
// part, a slot touched by many landmarks gets several lane groups (chunks), which share each batch's landmarks of
// the slot in contiguous runs and are merged in chunk order: every sum is taken in a fixed order. The block columns
// are shared out to host threads by their coupling counts and the shares joined in column order, so the list is
// the same whatever the thread count.
template <typename Slots, typename RowBase, typename Base>
int build_phase2(bls_schur *h, Slots slots_of, RowBase rowbase_of, Base base_of) {
    std::vector<int32_t> owner((size_t)h->ncoup);
    for (int64_t e = 0; e < h->neb; e++)
        for (int64_t c = h->eliminated_ptr[e]; c < h->eliminated_ptr[e + 1]; c++) owner[c] = (int32_t)e;
    // the work of column r2 grows with its landmarks times their couplings below it; share columns by coupling count
    const int64_t nrb = h->nrb;
    const int64_t total = h->ncoup;
    unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (total < (1 << 16)) nth = 1;
    std::vector<int64_t> cut{0};
    for (unsigned t = 1; t < nth; t++) {
        const int64_t target = total * t / nth;
        int64_t r = cut.back();
        while (r < nrb && h->retained_ptr[r] < target) r++;
        cut.push_back(r);
    }
    cut.push_back(nrb);
    std::vector<P2Share> shares(cut.size() - 1);
    {
        std::vector<std::thread> pool;
        for (size_t t = 0; t + 1 < cut.size(); t++)
            pool.emplace_back([&, t] {
                build_phase2_columns(h, owner, cut[t], cut[t + 1], slots_of, rowbase_of, base_of, shares[t]);
            });
        for (auto &th : pool) th.join();
    }
    for (auto &sh : shares)
        if (!sh.error.empty()) return fail(BLS_ERR_ARG, sh.error);
    std::vector<P2Item> items;
    std::vector<P2Entry> entries;
    std::vector<P2Task> tasks;
    std::vector<int64_t> slot_rowbase;
    std::vector<int32_t> slot_dr, slot_roff;
    std::vector<P2Merge> merges;
    int64_t npartial = 0;
    size_t n_it = 0, n_en = 0, n_tk = 0, n_sl = 0, n_mg = 0;
    for (auto &sh : shares) {
        n_it += sh.items.size(); n_en += sh.entries.size(); n_tk += sh.tasks.size(); n_sl += sh.slot_dr.size();
        n_mg += sh.merges.size();
    }
    items.reserve(n_it); entries.reserve(n_en + 1); tasks.reserve(n_tk); slot_rowbase.reserve(n_sl);
    slot_dr.reserve(n_sl); slot_roff.reserve(n_sl); merges.reserve(n_mg);
    for (auto &sh : shares) {
        const int64_t eo = (int64_t)entries.size(), to = (int64_t)tasks.size(), so = (int64_t)slot_dr.size();
        const int64_t po = npartial;
        for (P2Item it : sh.items) {
            it.entry_lo += eo;
            it.entry_hi += eo;
            it.task_lo += (int32_t)to;
            it.slot_lo += (int32_t)so;
            if (it.nparts > 1) it.pbase += po;
            items.push_back(it);
        }
        for (P2Merge mg : sh.merges) {
            mg.pbase += po;
            mg.slot_lo += (int32_t)so;
            merges.push_back(mg);
        }
        entries.insert(entries.end(), sh.entries.begin(), sh.entries.end());
        tasks.insert(tasks.end(), sh.tasks.begin(), sh.tasks.end());
        slot_rowbase.insert(slot_rowbase.end(), sh.slot_rowbase.begin(), sh.slot_rowbase.end());
        slot_dr.insert(slot_dr.end(), sh.slot_dr.begin(), sh.slot_dr.end());
        slot_roff.insert(slot_roff.end(), sh.slot_roff.begin(), sh.slot_roff.end());
        npartial += sh.npartial;
        sh = P2Share{};
    }
    if (entries.size() > (size_t)INT32_MAX || tasks.size() > (size_t)INT32_MAX || slot_dr.size() > (size_t)INT32_MAX)
        return fail(BLS_ERR_UNSUPPORTED, "too many phase-2 entries");
    // longest items first (their landmark count sets the workgroup's run time): the dispatcher then fills the
    // tail with short items. Items are independent, so the order does not change any result.
    std::stable_sort(items.begin(), items.end(), [](const P2Item &x, const P2Item &y) {
        return x.entry_hi - x.entry_lo > y.entry_hi - y.entry_lo;
    });
    const int64_t nentries = (int64_t)entries.size();
    // one unused record past the end: the kernel reads entries[entry_lo] unconditionally, also for an item without
    // entries (a block column no landmark touches), whose entry_lo may equal the entry count
    entries.push_back(P2Entry{});
    int rc;
    if ((rc = h->d_items.upload(items)) || (rc = h->d_entries.upload(entries)) || (rc = h->d_tasks.upload(tasks)) ||
        (rc = h->d_slot_rowbase.upload(slot_rowbase)) || (rc = h->d_slot_dr.upload(slot_dr)) ||
        (rc = h->d_slot_roff.upload(slot_roff)) || (rc = h->d_merges.upload(merges)) ||
        (rc = h->d_partial.alloc(npartial)))
        return rc;
    h->nitems = (int64_t)items.size();
    h->nmerges = (int64_t)merges.size();
    h->nentries = nentries;
    return BLS_OK;
}

template <int DE, class P>
void launch_phase1(const Run &r, const typename P::TE *E, const double *b_e, const double *sc, double lambda,
                   const double *s_e, typename P::TEI *E_inv, double *z, unsigned long long *flag, hipStream_t s) {
    k_phase1<DE, P><<<grid_for(r.count, 256), 256, 0, s>>>(r.first, r.count, r.eoff0, r.boff0, E, b_e, sc, lambda,
                                                            s_e, E_inv, z, flag);
}

// Device scratch of one reduction: the failure flag (phase 1's CAS-min, read by phase 2) and the row sums of
// the parts of split items. The handle owns one; bls_reduce_schur_sparse_f64_noalloc uses the library's static
// array instead.
struct P2Scratch {
    unsigned long long *flag;
    double *partial;
};

template <class P>
int phase1(bls_schur *h, const typename P::TE *E, const double *b_e, const double *sc, double lambda,
           const double *s_e, typename P::TEI *E_inv, double *z, unsigned long long *flag, hipStream_t s) {
    for (const Run &r : h->runs) {
        if (r.count == 0) continue;
        switch (r.de) {
        case 1: launch_phase1<1, P>(r, E, b_e, sc, lambda, s_e, E_inv, z, flag, s); break;
        case 2: launch_phase1<2, P>(r, E, b_e, sc, lambda, s_e, E_inv, z, flag, s); break;
        case 3: launch_phase1<3, P>(r, E, b_e, sc, lambda, s_e, E_inv, z, flag, s); break;
        case 4: launch_phase1<4, P>(r, E, b_e, sc, lambda, s_e, E_inv, z, flag, s); break;
        case 5: launch_phase1<5, P>(r, E, b_e, sc, lambda, s_e, E_inv, z, flag, s); break;
        case 6: launch_phase1<6, P>(r, E, b_e, sc, lambda, s_e, E_inv, z, flag, s); break;
        case 7: launch_phase1<7, P>(r, E, b_e, sc, lambda, s_e, E_inv, z, flag, s); break;
        case 8: launch_phase1<8, P>(r, E, b_e, sc, lambda, s_e, E_inv, z, flag, s); break;
        default: {
            using F = typename std::conditional<P::F32_FACTOR, float, double>::type;
            if (r.wsoff < 0) return fail(BLS_ERR_STATE, "no phase-1 workspace for blocks above 8");
            k_phase1_wide<P><<<(unsigned)r.wsgrid, WIDE_THREADS, 0, s>>>(
                r.first, r.count, (int)r.de, r.eoff0, r.boff0, E, b_e, sc, lambda, s_e, E_inv, z, flag,
                reinterpret_cast<F *>(h->d_wide_ws.ptr + r.wsoff));
            break;
        }
        }
    }
    BLS_HIP(hipGetLastError());
    return BLS_OK;
}

template <int DRM, class P> void launch_phase2_dr(int de, unsigned grid, hipStream_t s, const P2Args<P> &a) {
    switch (de) {
    case 1: k_phase2<1, DRM, P><<<grid, P2_THREADS, 0, s>>>(a); break;
    case 2: k_phase2<2, DRM, P><<<grid, P2_THREADS, 0, s>>>(a); break;
    case 3: k_phase2<3, DRM, P><<<grid, P2_THREADS, 0, s>>>(a); break;
    default: k_phase2<0, DRM, P><<<grid, P2_THREADS, 0, s>>>(a); break;
    }
}

template <class P> void launch_phase2(int de, int max_dr, unsigned grid, hipStream_t s, const P2Args<P> &a) {
    if (max_dr <= 6) launch_phase2_dr<6, P>(de, grid, s, a);
    else if (max_dr <= 9) launch_phase2_dr<9, P>(de, grid, s, a);
    else launch_phase2_dr<16, P>(de, grid, s, a);
}

template <class Q>
bool run_phase2(bls_schur *h, const typename Q::TEI *E_inv, const double *z, const typename Q::TC *C,
                const double *b_r, const double *s_r, typename Q::TS *S, double *g, const P2Scratch &ws,
                int64_t *d_failed, hipStream_t s) {
    P2Args<Q> a{};
    a.items = h->d_items.ptr;
    a.nitems = h->nitems;
    a.entries = h->d_entries.ptr;
    a.tasks = h->d_tasks.ptr;
    a.slot_rowbase = h->d_slot_rowbase.ptr;
    a.slot_dr = h->d_slot_dr.ptr;
    a.slot_roff = h->d_slot_roff.ptr;
    a.partial = ws.partial;
    a.eoffsets = h->d_eoffsets.ptr;
    a.boffsets_e = h->d_boffsets_e.ptr;
    a.coffsets = h->d_coffsets.ptr;
    a.e_dims = h->d_e_dims.ptr;
    a.E_inv = E_inv;
    a.z = z;
    a.C = C;
    a.b_r = b_r;
    a.s_r = s_r;
    a.S = S;
    a.g = g;
    a.flag = ws.flag;
    a.uniform_offsets = h->uniform_de > 0 && h->uniform_dr_all > 0 && h->runs.size() == 1;
    a.cstride = a.uniform_offsets ? (int64_t)h->uniform_de * h->uniform_dr_all : 0;
    // the verdict folded into the kernel's last workgroup when nothing runs after it (no split items) and the
    // handle's own flag is used
    if (h->use_tiles()) {
        // the tile plan: no split items, so the verdict is always folded into its last workgroup when the handle's
        // own flag is used
        const bool tfold = ws.flag == h->d_flag.ptr && h->d_fin_counter.ptr;
        const int variant = all_f64<Q>() ? 0 : (sizeof(typename Q::TS) == 4 ? 1 : 2);
        if (tiles_launch(h->tiles, variant, E_inv, z, C, b_r, s_r, S, g, h->d_S_colptr.ptr, h->nr, ws.flag,
                         tfold ? h->d_fin_counter.ptr : nullptr, tfold ? ws.flag : nullptr, tfold ? d_failed : nullptr,
                         s))
            return false;
        h->timing.mark(3, s);
        h->timing.mark(4, s);
        return tfold;
    }
    const bool fold = h->nitems > 0 && h->nmerges == 0 && ws.flag == h->d_flag.ptr && h->d_fin_counter.ptr;
    if (fold) {
        a.fin_counter = h->d_fin_counter.ptr;
        a.fin_flag = ws.flag;
        a.fin_out = d_failed;
    }
    if (h->nitems > 0) {
        launch_phase2<Q>(h->uniform_de, h->max_dr, (unsigned)h->nitems, s, a);
        h->timing.mark(3, s);
        if (h->nmerges > 0)
            k_phase2_merge<typename Q::TS><<<(unsigned)h->nmerges, P2_THREADS, 0, s>>>(
                h->d_merges.ptr, h->d_slot_rowbase.ptr, h->d_slot_dr.ptr, ws.partial, S, g, ws.flag);
    }
    if (h->nitems == 0) h->timing.mark(3, s);
    h->timing.mark(4, s);
    return fold;
}

// Wide mode's W, S and g_S (every block size): see k_phase2_wide.
template <class Q>
void run_phase2_wide(bls_schur *h, const typename Q::TEI *E_inv, const double *z, const typename Q::TC *C,
                     const double *b_r, const double *s_r, typename Q::TS *S, double *g, const P2Scratch &ws,
                     hipStream_t s) {
    using TC = typename Q::TC;
    if (h->ncoup)
        k_wide_w<typename Q::TEI, TC><<<grid_for(h->ncoup, WIDE_THREADS / 64), WIDE_THREADS, 0, s>>>(
            h->ncoup, h->d_coup_e.ptr, h->d_e_dims.ptr, h->d_eoffsets.ptr, h->d_coffsets.ptr, h->d_entry_retained.ptr,
            h->d_r_dims.ptr, E_inv, C, h->d_W.ptr);
    if (h->nwpairs)
        k_phase2_wide<typename Q::TS, TC><<<(unsigned)h->nwpairs, WIDE_THREADS, 0, s>>>(
            h->d_wpairs.ptr, h->d_wcc.ptr, h->d_coup_e.ptr, h->d_e_dims.ptr, h->d_coffsets.ptr, h->d_r_dims.ptr,
            h->d_W.ptr, C, S, ws.flag);
    h->timing.mark(3, s);
    if (h->nrb)
        k_g_wide<TC><<<(unsigned)h->nrb, WIDE_THREADS, 0, s>>>(h->d_retained_ptr.ptr, h->d_rm_ce.ptr, h->d_e_dims.ptr,
                                                              h->d_boffsets_e.ptr, h->d_coffsets.ptr, h->d_r_dims.ptr,
                                                              h->d_boffsets_r.ptr, C, z, b_r, s_r, g, ws.flag);
    h->timing.mark(4, s);
}

template <class P>
int reduce_async(bls_schur *h, int want_mode, typename P::TS *S, double *g, const double *b_r, const double *b_e,
                 const double *scaling_e, double lambda, const typename P::TE *E, const typename P::TC *C,
                 typename P::TEI *E_inv, double *z, const double *s_r, const double *s_e, int64_t *d_failed,
                 hipStream_t s, const P2Scratch *scratch = nullptr) {
    using TC = typename P::TC;
    if (!h) return fail(BLS_ERR_ARG, "null handle");
    if (h->mode != want_mode)
        return fail(BLS_ERR_STATE, want_mode == 1 ? "handle has no sparse layout (bls_schur_set_sparse)"
                                                  : "handle is not dense (bls_schur_set_dense)");
    if (!(std::isfinite(lambda) && lambda >= 0))
        return fail(BLS_ERR_ARG, "lambda must be finite and non-negative, got " + std::to_string(lambda));
    if (!S || !g || !b_r || !scaling_e || !E || !C || !E_inv || !z || !d_failed || (h->neb && !b_e))
        return fail(BLS_ERR_ARG, "reduce_schur: a required array is NULL");
    if ((s_r || s_e) && h->clen && !h->d_Cs.ptr) {
        int rc = h->d_Cs.alloc(h->clen);
        if (rc) return rc;
    }
    const P2Scratch ws = scratch ? *scratch : P2Scratch{h->d_flag.ptr, h->d_partial.ptr};
    h->timing.mark(1, s);
    const bool own_flag = ws.flag == h->d_flag.ptr;
    if (!own_flag || h->flag_dirty) k_fill_u64<<<1, 1, 0, s>>>(ws.flag, (unsigned long long)NO_FAILURE);
    if (own_flag) h->flag_dirty = true;   // until the finalize below is queued
    int rc = phase1<P>(h, E, b_e, scaling_e, lambda, s_e, E_inv, z, ws.flag, s);
    if (rc) return rc;
    bool folded = false;
    if ((s_r || s_e) && h->neb) {
        using Q = typename ScaledPolicy<P>::type;
        using TSC = typename P::TSC;
        TSC *Cs = reinterpret_cast<TSC *>(h->d_Cs.ptr);
        k_scale_c<TC, TSC><<<grid_for(h->neb, 256), 256, 0, s>>>(h->neb, h->d_e_dims.ptr, h->d_boffsets_e.ptr,
                                                                  h->d_eliminated_ptr.ptr, h->d_entry_retained.ptr,
                                                                  h->d_coffsets.ptr, h->d_r_dims.ptr,
                                                                  h->d_boffsets_r.ptr, C, s_r, s_e, Cs);
        h->timing.mark(2, s);
        if (h->wide) run_phase2_wide<Q>(h, E_inv, z, Cs, b_r, s_r, S, g, ws, s);
        else folded = run_phase2<Q>(h, E_inv, z, Cs, b_r, s_r, S, g, ws, d_failed, s);
    } else {
        h->timing.mark(2, s);
        if (h->wide) run_phase2_wide<P>(h, E_inv, z, C, b_r, s_r, S, g, ws, s);
        else folded = run_phase2<P>(h, E_inv, z, C, b_r, s_r, S, g, ws, d_failed, s);
    }
    if (!folded) k_finalize_flag<<<1, 1, 0, s>>>(ws.flag, d_failed);
    BLS_HIP(hipGetLastError());
    if (own_flag) h->flag_dirty = false;
    return BLS_OK;
}

template <class P>
int reduce_sync(bls_schur *h, int mode, typename P::TS *S, double *g, const double *b_r, const double *b_e,
                const double *sc, double lambda, const typename P::TE *E, const typename P::TC *C,
                typename P::TEI *E_inv, double *z, const double *s_r, const double *s_e, int64_t *failed_block,
                hipStream_t s) {
    int rc = reduce_async<P>(h, mode, S, g, b_r, b_e, sc, lambda, E, C, E_inv, z, s_r, s_e, h ? h->d_failed_out.ptr : nullptr, s);
    if (rc) return rc;
    int64_t fb = 0;
    BLS_HIP(hipMemcpyAsync(&fb, h->d_failed_out.ptr, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    BLS_HIP(hipStreamSynchronize(s));
    if (h->timing.enabled)
        for (int i = 1; i < 4; i++) (void)hipEventElapsedTime(&h->timing.ms[i], h->timing.ev[i], h->timing.ev[i + 1]);
    if (failed_block) *failed_block = fb;
    return fb ? BLS_FAILED : BLS_OK;
}

std::vector<int64_t> to0(const int64_t *p, int64_t n) {
    std::vector<int64_t> v(p, p + n);
    for (auto &x : v) x -= 1;
    return v;
}

std::vector<int32_t> narrow32(const std::vector<int64_t> &v) { return std::vector<int32_t>(v.begin(), v.end()); }

}  // namespace

namespace {
// The phase-2 work list of the registered layout, with the handle's shape parameters.
int build_sparse_phase2(bls_schur *h) {
    const std::vector<int64_t> &bcp = h->lay_bcp, &brv = h->lay_brv, &bro = h->lay_bro, &Scp = h->lay_Scp,
                               &ch = h->lay_ch;
    if (h->wide)
        return build_wide_pairs(h, [&](int64_t r1, int64_t r2, int64_t &base, int64_t &stride) {
            for (int64_t k = bcp[r2]; k < bcp[r2 + 1]; k++)
                if (brv[k] == r1) {
                    base = Scp[h->boffsets_r[r2]] - 1 + bro[k];
                    stride = ch[r2];
                    return true;
                }
            return false;
        });
    return build_phase2(
        h, [&](int64_t r2, std::vector<int64_t> &col) { col.assign(brv.begin() + bcp[r2], brv.begin() + bcp[r2 + 1]); },
        [&](int64_t r2, size_t k, int64_t) { return bro[bcp[r2] + (int64_t)k]; },
        [&](int64_t r2, int64_t &base, int64_t &cs) {
            base = Scp[h->boffsets_r[r2]] - 1;
            cs = ch[r2];
        });
}

int build_dense_phase2(bls_schur *h) {
    if (h->wide)
        return build_wide_pairs(h, [&](int64_t r1, int64_t r2, int64_t &base, int64_t &stride) {
            base = h->boffsets_r[r2] * h->nr + h->boffsets_r[r1];
            stride = h->nr;
            return true;
        });
    return build_phase2(
        h, [&](int64_t r2, std::vector<int64_t> &col) { for (int64_t r = 0; r <= r2; r++) col.push_back(r); },
        [&](int64_t, size_t, int64_t r1) { return h->boffsets_r[r1]; },
        [&](int64_t r2, int64_t &base, int64_t &cs) {
            base = h->boffsets_r[r2] * h->nr;
            cs = h->nr;
        });
}
}  // namespace

extern "C" int bls_schur_create(int64_t nr, int64_t nrblocks, int64_t neblocks, int64_t ncoup, const int64_t *r_dims,
                                const int64_t *e_dims, const int64_t *boffsets_r, const int64_t *boffsets_e,
                                const int64_t *eoffsets, const int64_t *eliminated_ptr, const int64_t *entry_retained,
                                const int64_t *coffsets, const int64_t *retained_ptr, const int64_t *retained_entry,
                                bls_schur **out) {
    if (!out || nr < 0 || nrblocks < 0 || neblocks < 0 || ncoup < 0) return fail(BLS_ERR_ARG, "bls_schur_create: bad arguments");
    *out = nullptr;
    if (ncoup > INT32_MAX || neblocks > INT32_MAX) return fail(BLS_ERR_UNSUPPORTED, "more than 2^31 couplings");
    auto *h = new bls_schur();
    h->nr = nr;
    h->nrb = nrblocks;
    h->neb = neblocks;
    h->ncoup = ncoup;
    h->r_dims.assign(r_dims, r_dims + nrblocks);
    h->e_dims.assign(e_dims, e_dims + neblocks);
    h->boffsets_r = to0(boffsets_r, nrblocks);
    h->boffsets_e = to0(boffsets_e, neblocks);
    h->eoffsets = to0(eoffsets, neblocks);
    h->eliminated_ptr = to0(eliminated_ptr, neblocks + 1);
    h->entry_retained = to0(entry_retained, ncoup);
    h->coffsets = to0(coffsets, ncoup);
    h->retained_ptr = to0(retained_ptr, nrblocks + 1);
    h->retained_entry = to0(retained_entry, ncoup);
    int rc = check_structure(h);
    if (rc) { delete h; return rc; }
    for (int64_t c = 0; c < ncoup; c++)
        if (h->entry_retained[c] < 0 || h->entry_retained[c] >= nrblocks) { delete h; return fail(BLS_ERR_ARG, "entry_retained out of range"); }
    for (int64_t e = 0; e < neblocks; e++)
        for (int64_t c = h->eliminated_ptr[e]; c < h->eliminated_ptr[e + 1]; c++)
            h->clen = std::max(h->clen, h->coffsets[c] + h->e_dims[e] * h->r_dims[h->entry_retained[c]]);
    // runs of equal d_e, with contiguous E / b offsets inside each
    h->uniform_de = neblocks ? (int)h->e_dims[0] : 3;
    for (int64_t e = 0; e < neblocks;) {
        int64_t f = e;
        while (e < neblocks && h->e_dims[e] == h->e_dims[f]) e++;
        h->runs.push_back(Run{f, e - f, h->e_dims[f], h->eoffsets[f], h->boffsets_e[f]});
        if (h->e_dims[f] != h->uniform_de) h->uniform_de = 0;
    }
    if (h->uniform_de > 3) h->uniform_de = 0;
    {
        int64_t wsn = 0;
        for (Run &r : h->runs)
            if (r.de > 8) {
                r.wsgrid = std::min<int64_t>(r.count, WIDE_P1_GRID);
                r.wsoff = wsn;
                wsn += 2 * r.de * r.de * r.wsgrid;
            }
        if (wsn && (rc = h->d_wide_ws.alloc((size_t)wsn))) {
            delete h;
            return rc;
        }
    }
    h->uniform_dr_all = nrblocks ? (int)h->r_dims[0] : 0;
    for (int64_t r = 0; r < nrblocks; r++) h->max_dr = std::max(h->max_dr, (int)h->r_dims[r]);
    for (int64_t r = 0; r < nrblocks; r++)
        if (h->r_dims[r] != h->uniform_dr_all) h->uniform_dr_all = 0;
    {
        const int64_t de = neblocks ? h->e_dims[0] : 0, dr = h->uniform_dr_all;
        bool ok = de >= 1 && de <= 4 && dr >= 1;
        for (int64_t e = 0; ok && e < neblocks; e++) {
            ok = h->e_dims[e] == de && h->eoffsets[e] == e * de * de && h->boffsets_e[e] == e * de;
            for (int64_t c = h->eliminated_ptr[e] + 1; ok && c < h->eliminated_ptr[e + 1]; c++)
                ok = h->entry_retained[c] == h->entry_retained[c - 1] + 1;
        }
        for (int64_t r = 0; ok && r < nrblocks; r++) ok = h->boffsets_r[r] == r * dr;
        for (int64_t c = 0; ok && c < ncoup; c++) ok = h->coffsets[c] == c * de * dr;
        if (ok) {
            h->win_de = (int)de;
            h->win_dr = (int)dr;
        }
    }
    if ((rc = h->d_eoffsets.upload(h->eoffsets)) || (rc = h->d_boffsets_e.upload(h->boffsets_e)) ||
        (rc = h->d_coffsets.upload(h->coffsets)) || (rc = h->d_eliminated_ptr.upload(h->eliminated_ptr)) ||
        (rc = h->d_boffsets_r.upload(h->boffsets_r)) || (rc = h->d_e_dims.upload(narrow32(h->e_dims))) ||
        (rc = h->d_r_dims.upload(narrow32(h->r_dims))) || (rc = h->d_entry_retained.upload(narrow32(h->entry_retained))) ||
        (rc = h->d_flag.alloc(1)) || (rc = h->d_failed_out.alloc(1)) || (rc = h->d_fin_counter.alloc(1)) ||
        (rc = hipMemset(h->d_fin_counter.ptr, 0, sizeof(unsigned int)) == hipSuccess
                  ? BLS_OK
                  : fail(BLS_ERR_HIP, "hipMemset of the phase-2 counter failed"))) {
        delete h;
        return rc;
    }
    {
        std::vector<int32_t> coup_e((size_t)h->ncoup);
        for (int64_t e = 0; e < h->neb; e++)
            for (int64_t c = h->eliminated_ptr[e]; c < h->eliminated_ptr[e + 1]; c++) coup_e[c] = (int32_t)e;
        std::vector<int2> ce((size_t)h->ncoup);
        for (int64_t k = 0; k < h->ncoup; k++) {
            const int64_t c = h->retained_entry[k];
            ce[k] = make_int2((int)c, coup_e[c]);
        }
        if ((rc = h->d_retained_ptr.upload(h->retained_ptr)) || (rc = h->d_rm_ce.upload(ce))) {
            delete h;
            return rc;
        }
    }
    *out = h;
    return BLS_OK;
}

extern "C" void bls_schur_destroy(bls_schur *h) { delete h; }

extern "C" int bls_schur_set_sparse(bls_schur *h, int64_t nnz, const int64_t *S_colptr, int64_t nnz_blocks,
                                    const int64_t *block_colptr, const int64_t *block_rowval,
                                    const int64_t *block_row_offsets, const int64_t *col_heights, int64_t nR,
                                    const int64_t *r_to_s_map, const int64_t *diag_inds) {
    if (!h || nnz < 0 || nnz_blocks < 0 || nR < 0) return fail(BLS_ERR_ARG, "bls_schur_set_sparse: bad arguments");
    std::vector<int64_t> bcp = to0(block_colptr, h->nrb + 1), brv = to0(block_rowval, nnz_blocks);
    std::vector<int64_t> Scp(S_colptr, S_colptr + h->nr + 1);
    for (int64_t r = 0; r < h->nrb; r++)
        for (int64_t k = bcp[r]; k < bcp[r + 1]; k++)
            if (brv[k] < 0 || brv[k] > r) return fail(BLS_ERR_ARG, "block pattern is not upper triangular");
    h->lay_bcp = bcp;
    h->lay_brv = brv;
    h->lay_bro.assign(block_row_offsets, block_row_offsets + nnz_blocks);
    h->lay_Scp = Scp;
    h->lay_ch.assign(col_heights, col_heights + h->nrb);
    int rc = build_sparse_phase2(h);
    if (rc) return rc;
    if (!h->wide) {
        const TileStructure T{h->nr, h->nrb, h->neb, h->e_dims, h->r_dims, h->eoffsets, h->boffsets_e, h->boffsets_r,
                              h->coffsets, h->eliminated_ptr, h->entry_retained, &h->lay_bcp, &h->lay_brv, &h->lay_bro};
        if ((rc = tiles_build(h->tiles, T, &h->tiles_why))) return rc;
    } else {
        h->tiles_why = "blocks above the register path";
    }
    std::vector<int64_t> rts(r_to_s_map, r_to_s_map + nR), di(diag_inds, diag_inds + h->nr);
    // scalar row of every stored entry (1-based), for the conditioning scale of reset_schur_sparse!: a scalar column
    // of block column r2 holds, in order, the rows of the blocks (r1, r2) of the pattern
    std::vector<int64_t> Srv((size_t)nnz);
    for (int64_t r2 = 0; r2 < h->nrb; r2++) {
        std::vector<int64_t> rows;
        for (int64_t k = bcp[r2]; k < bcp[r2 + 1]; k++)
            for (int64_t i = 0; i < h->r_dims[brv[k]]; i++) rows.push_back(h->boffsets_r[brv[k]] + i + 1);
        if ((int64_t)rows.size() != col_heights[r2]) return fail(BLS_ERR_ARG, "col_heights do not match the block pattern");
        for (int64_t j = h->boffsets_r[r2]; j < h->boffsets_r[r2] + h->r_dims[r2]; j++) {
            if (Scp[j + 1] - Scp[j] != col_heights[r2] || Scp[j + 1] - 1 > nnz)
                return fail(BLS_ERR_ARG, "S_colptr does not match the block pattern");
            std::copy(rows.begin(), rows.end(), Srv.begin() + (Scp[j] - 1));
        }
    }
    if ((rc = h->d_r_to_s.upload(rts)) || (rc = h->d_diag_inds.upload(di)) || (rc = h->d_S_colptr.upload(Scp)) ||
        (rc = h->d_S_rowval.upload(Srv)))
        return rc;
    {
        std::vector<int32_t> src((size_t)nnz, -1);
        std::vector<int2> rcv((size_t)nnz, make_int2(-1, -2));
        bool ok = nnz <= INT32_MAX && nR < (1 << 30);
        for (int64_t i = 0; ok && i < nR; i++)   // the last R entry mapped to an S entry wins, as the scatter loop's
            if (rts[i] > 0) {
                if (rts[i] > nnz) return fail(BLS_ERR_ARG, "r_to_s_map points past the stored entries");
                src[rts[i] - 1] = (int32_t)i;
            }
        for (int64_t j = 0; ok && j < h->nr; j++)
            for (int64_t k = Scp[j] - 1; k < Scp[j + 1] - 1; k++) rcv[k] = make_int2((int)(Srv[k] - 1), (int)j);
        for (int64_t i = 0; ok && i < h->nr; i++) {
            const int64_t k = di[i] - 1;
            if (k < 0 || k >= nnz || rcv[k].x != i || rcv[k].y != i) ok = false;
        }
        if (ok) {
            int64_t ndiag = 0;
            for (int64_t k = 0; k < nnz; k++) {
                if (rcv[k].x < 0) ok = false;   // an entry no column holds
                ndiag += rcv[k].x == rcv[k].y;
            }
            ok = ok && ndiag == h->nr;
        }
        if (ok) {
            std::vector<int32_t> code((size_t)nnz);
            for (int64_t k = 0; k < nnz; k++) code[k] = ((src[k] + 1) << 1) | (rcv[k].x == rcv[k].y ? 1 : 0);
            if ((rc = h->d_rst_code.upload(code)) || (rc = h->d_rst_rc.upload(rcv))) return rc;
        } else {
            h->d_rst_code.reset();
            h->d_rst_rc.reset();
        }
    }
    h->nnz = nnz;
    h->nR = nR;
    h->mode = 1;
    return BLS_OK;
}

extern "C" int bls_schur_set_dense(bls_schur *h) {
    if (!h) return fail(BLS_ERR_ARG, "null handle");
    int rc = build_dense_phase2(h);
    if (rc) return rc;
    h->mode = 2;
    return BLS_OK;
}

namespace {

template <class TS, class TR>
int reset_sparse(bls_schur *h, TS *nzval, const TR *R_nzval, const double *scaling_r, double lambda,
                 const double *s_r, hipStream_t s) {
    if (!h) return fail(BLS_ERR_ARG, "null handle");
    if (h->mode != 1) return fail(BLS_ERR_STATE, "handle has no sparse layout (bls_schur_set_sparse)");
    if (!(std::isfinite(lambda) && lambda >= 0))
        return fail(BLS_ERR_ARG, "lambda must be finite and non-negative, got " + std::to_string(lambda));
    if (!nzval || !scaling_r || (h->nR && !R_nzval)) return fail(BLS_ERR_ARG, "reset_schur_sparse: a required array is NULL");
    h->timing.mark(0, s);
    if (h->nnz && h->d_rst_code.ptr) {
        k_reset_sparse1<TS, TR><<<std::min<int64_t>(grid_for(h->nnz + h->nr, 256), 8192), 256, 0, s>>>(
            nzval, R_nzval, h->d_rst_code.ptr, h->d_rst_rc.ptr, h->d_diag_inds.ptr, h->nnz, h->nr, scaling_r, lambda,
            s_r);
        h->timing.mark(1, s);
        BLS_HIP(hipGetLastError());
        return BLS_OK;
    }
    k_fill<TS><<<std::min<int64_t>(grid_for(h->nnz, 256), 4096), 256, 0, s>>>(nzval, h->nnz, (TS)0);
    if (h->nR)
        k_scatter_r<TS, TR><<<std::min<int64_t>(grid_for(h->nR, 256), 4096), 256, 0, s>>>(nzval, R_nzval, h->d_r_to_s.ptr,
                                                                                          h->nR);
    if (s_r && h->nr)   // scale_sparse! (src/conditioning.jl:247-253)
        k_scale_sparse<TS><<<grid_for(h->nr, 256), 256, 0, s>>>(nzval, h->d_S_colptr.ptr, h->d_S_rowval.ptr, h->nr, s_r);
    if (h->nr)
        k_add_damping<TS><<<grid_for(h->nr, 256), 256, 0, s>>>(nzval, h->d_diag_inds.ptr, h->nr, scaling_r, lambda, s_r);
    h->timing.mark(1, s);
    BLS_HIP(hipGetLastError());
    return BLS_OK;
}

template <class TS, class TR>
int reset_dense(bls_schur *h, TS *S, const TR *R, const double *scaling_r, double lambda, const double *s_r,
                hipStream_t s) {
    if (!h) return fail(BLS_ERR_ARG, "null handle");
    if (!(std::isfinite(lambda) && lambda >= 0))
        return fail(BLS_ERR_ARG, "lambda must be finite and non-negative, got " + std::to_string(lambda));
    if (h->nr && (!S || !R || !scaling_r)) return fail(BLS_ERR_ARG, "reset_schur_dense: a required array is NULL");
    h->timing.mark(0, s);
    if (h->nr)
        k_reset_dense<TS, TR><<<std::min<int64_t>(grid_for(h->nr * h->nr, 256), 4096), 256, 0, s>>>(S, R, h->nr, scaling_r,
                                                                                                   lambda, s_r);
    h->timing.mark(1, s);
    BLS_HIP(hipGetLastError());
    return BLS_OK;
}

template <class TC, class TEI, class TSC>
int backsub(bls_schur *h, double *x_e, const double *x_r, const TC *C_data, const TEI *E_inv, const double *z,
            const double *s_r, const double *s_e, hipStream_t s) {
    if (!h) return fail(BLS_ERR_ARG, "null handle");
    if (h->neb && (!x_e || !x_r || !C_data || !E_inv || !z)) return fail(BLS_ERR_ARG, "backsubstitute: a required array is NULL");
    if (h->neb && h->wide)
        k_backsub_wide<TC, TEI, TSC><<<(unsigned)std::min<int64_t>(h->neb, 8192), WIDE_THREADS,
                                       (size_t)h->wide_maxde * sizeof(double), s>>>(
            h->neb, h->d_e_dims.ptr, h->d_eoffsets.ptr, h->d_boffsets_e.ptr, h->d_eliminated_ptr.ptr,
            h->d_entry_retained.ptr, h->d_coffsets.ptr, h->d_r_dims.ptr, h->d_boffsets_r.ptr, C_data, E_inv, z, x_r, s_r,
            s_e, x_e);
    else if (h->neb && h->win_de > 0) {
        const unsigned grid = grid_for(h->neb, 8);
        auto go = [&](auto de_c) {
            constexpr int DE = decltype(de_c)::value;
            k_backsub_win<DE, TC, TEI, TSC><<<grid, 256, 0, s>>>(h->neb, h->d_eliminated_ptr.ptr,
                                                                 h->d_entry_retained.ptr, h->win_dr, C_data, E_inv, z,
                                                                 x_r, s_r, s_e, x_e);
        };
        switch (h->win_de) {
        case 1: go(std::integral_constant<int, 1>{}); break;
        case 2: go(std::integral_constant<int, 2>{}); break;
        case 3: go(std::integral_constant<int, 3>{}); break;
        default: go(std::integral_constant<int, 4>{}); break;
        }
    } else if (h->neb)
        k_backsub<TC, TEI, TSC><<<grid_for(h->neb, 256 / BS_LANES), 256, 0, s>>>(h->neb, h->d_e_dims.ptr, h->d_eoffsets.ptr,
                                                                  h->d_boffsets_e.ptr, h->d_eliminated_ptr.ptr,
                                                                  h->d_entry_retained.ptr, h->d_coffsets.ptr,
                                                                  h->d_r_dims.ptr, h->d_boffsets_r.ptr, C_data, E_inv,
                                                                  z, x_r, s_r, s_e, x_e);
    BLS_HIP(hipGetLastError());
    return BLS_OK;
}

template <class T> int copyto(int64_t n, T *dest, const T *src, const int64_t *map, hipStream_t s) {
    if (n < 0) return fail(BLS_ERR_ARG, "copyto: negative length");
    if (n == 0) return BLS_OK;
    if (!dest || !src || !map) return fail(BLS_ERR_ARG, "copyto: a required array is NULL");
    k_gather<T><<<std::min<int64_t>(grid_for(n, 256), 8192), 256, 0, s>>>(dest, src, map, n);
    BLS_HIP(hipGetLastError());
    return BLS_OK;
}

}  // namespace

// ---- Float64 (SchurBlockHessian{Float64})
extern "C" int bls_reset_schur_sparse_f64(bls_schur *h, double *nzval, const double *R_nzval, const double *scaling_r,
                                          double lambda, const double *s_r, void *stream) {
    return reset_sparse<double, double>(h, nzval, R_nzval, scaling_r, lambda, s_r, (hipStream_t)stream);
}

extern "C" int bls_reset_schur_dense_f64(bls_schur *h, double *S, const double *R, const double *scaling_r,
                                         double lambda, const double *s_r, void *stream) {
    return reset_dense<double, double>(h, S, R, scaling_r, lambda, s_r, (hipStream_t)stream);
}

extern "C" int bls_reduce_schur_sparse_f64(bls_schur *h, double *nzval, double *g_S, const double *b_r,
                                           const double *b_e, const double *scaling_e, double lambda,
                                           const double *E_data, const double *C_data, double *E_inv, double *z,
                                           const double *s_r, const double *s_e, int64_t *failed_block, void *stream) {
    return reduce_sync<PrecF64>(h, 1, nzval, g_S, b_r, b_e, scaling_e, lambda, E_data, C_data, E_inv, z, s_r, s_e,
                                failed_block, (hipStream_t)stream);
}

extern "C" int bls_reduce_schur_dense_f64(bls_schur *h, double *S, double *g_S, const double *b_r, const double *b_e,
                                          const double *scaling_e, double lambda, const double *E_data,
                                          const double *C_data, double *E_inv, double *z, const double *s_r,
                                          const double *s_e, int64_t *failed_block, void *stream) {
    return reduce_sync<PrecF64>(h, 2, S, g_S, b_r, b_e, scaling_e, lambda, E_data, C_data, E_inv, z, s_r, s_e,
                                failed_block, (hipStream_t)stream);
}

extern "C" int bls_reduce_schur_sparse_f64_async(bls_schur *h, double *nzval, double *g_S, const double *b_r,
                                                 const double *b_e, const double *scaling_e, double lambda,
                                                 const double *E_data, const double *C_data, double *E_inv, double *z,
                                                 const double *s_r, const double *s_e, int64_t *d_failed_block,
                                                 void *stream) {
    return reduce_async<PrecF64>(h, 1, nzval, g_S, b_r, b_e, scaling_e, lambda, E_data, C_data, E_inv, z, s_r, s_e,
                                 d_failed_block, (hipStream_t)stream);
}

// ---- factor_all_eliminated! and the retained residual (the Float64 iterative refinement of the mixed path) -------
extern "C" int bls_schur_factor_f64(bls_schur *h, const double *E_data, const double *b_e, const double *scaling_e,
                                    double lambda, double *E_inv, double *z, const double *s_e, int64_t *failed_block,
                                    void *stream) {
    if (!h) return fail(BLS_ERR_ARG, "null handle");
    if (!(std::isfinite(lambda) && lambda >= 0))
        return fail(BLS_ERR_ARG, "lambda must be finite and non-negative, got " + std::to_string(lambda));
    if (h->neb && (!E_data || !b_e || !scaling_e || !E_inv || !z))
        return fail(BLS_ERR_ARG, "factor_all_eliminated: a required array is NULL");
    hipStream_t s = (hipStream_t)stream;
    h->flag_dirty = true;
    k_fill_u64<<<1, 1, 0, s>>>(h->d_flag.ptr, (unsigned long long)NO_FAILURE);
    int rc = phase1<PrecF64>(h, E_data, b_e, scaling_e, lambda, s_e, E_inv, z, h->d_flag.ptr, s);
    if (rc) return rc;
    k_finalize_flag<<<1, 1, 0, s>>>(h->d_flag.ptr, h->d_failed_out.ptr);
    h->flag_dirty = false;
    BLS_HIP(hipGetLastError());
    int64_t fb = 0;
    BLS_HIP(hipMemcpyAsync(&fb, h->d_failed_out.ptr, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    BLS_HIP(hipStreamSynchronize(s));
    if (failed_block) *failed_block = fb;
    return fb ? BLS_FAILED : BLS_OK;
}

template <class TC>
int retained_residual(bls_schur *h, double *out, const double *b_r, const double *x_e, const TC *C,
                      const double *s_r, const double *s_e, hipStream_t s) {
    if (!h) return fail(BLS_ERR_ARG, "null handle");
    if (!out || !b_r || (h->neb && (!x_e || !C))) return fail(BLS_ERR_ARG, "retained residual: a required array is NULL");
    if (h->wide)
        return fail(BLS_ERR_UNSUPPORTED, "retained residual: blocks up to d_e 8 and d_r " + std::to_string(MAXDR));
    const bool uni = h->uniform_de == 3 && h->uniform_dr_all == 6 && h->runs.size() == 1;
    if (h->nrb && uni)
        k_ct_xe<TC, 3, 6><<<(unsigned)h->nrb, CTX_THREADS, 0, s>>>(h->d_retained_ptr.ptr, h->d_rm_ce.ptr, h->d_e_dims.ptr,
                                                                   h->d_boffsets_e.ptr, h->d_coffsets.ptr, h->d_r_dims.ptr,
                                                                   h->d_boffsets_r.ptr, C, x_e, b_r, s_r, s_e, out);
    else if (h->nrb)
        k_ct_xe<TC, 0, 0><<<(unsigned)h->nrb, CTX_THREADS, 0, s>>>(h->d_retained_ptr.ptr, h->d_rm_ce.ptr, h->d_e_dims.ptr,
                                                                   h->d_boffsets_e.ptr, h->d_coffsets.ptr, h->d_r_dims.ptr,
                                                                   h->d_boffsets_r.ptr, C, x_e, b_r, s_r, s_e, out);
    BLS_HIP(hipGetLastError());
    return BLS_OK;
}

extern "C" int bls_schur_retained_residual_f64(bls_schur *h, double *out, const double *b_r, const double *x_e,
                                               const double *C_data, const double *s_r, const double *s_e,
                                               void *stream) {
    return retained_residual<double>(h, out, b_r, x_e, C_data, s_r, s_e, (hipStream_t)stream);
}

extern "C" int bls_schur_retained_residual_f32(bls_schur *h, double *out, const double *b_r, const double *x_e,
                                               const float *C_data, const double *s_r, const double *s_e,
                                               void *stream) {
    return retained_residual<float>(h, out, b_r, x_e, C_data, s_r, s_e, (hipStream_t)stream);
}


// ---- reduce_schur_dense! on the matrix cores (the dense configuration's SYRK) ------------------------------------
namespace {
// The landmark blocks E, their factorisation (E_inv, z) and S are Float64; the couplings C are TC and Y is TY.
template <class TC, class TY>
int reduce_dense_mfma(bls_schur *h, double *S, double *g, const double *b_r, const double *b_e, const double *sc,
                      double lambda, const double *E, const TC *C, double *E_inv, double *z,
                      int64_t *failed_block, int64_t nslices_req, hipStream_t s) {
    if (!h) return fail(BLS_ERR_ARG, "null handle");
    if (h->mode != 2) return fail(BLS_ERR_STATE, "handle is not dense (bls_schur_set_dense)");
    if (!(std::isfinite(lambda) && lambda >= 0))
        return fail(BLS_ERR_ARG, "lambda must be finite and non-negative, got " + std::to_string(lambda));
    if (!S || !g || !b_r || !sc || !E || !C || !E_inv || !z || (h->neb && !b_e))
        return fail(BLS_ERR_ARG, "reduce_schur_dense (MFMA): a required array is NULL");
    if (h->wide)
        return fail(BLS_ERR_UNSUPPORTED, "MFMA dense reduction: blocks up to d_e 8 and d_r " + std::to_string(MAXDR));
    // phase 1 (the reference's landmark factorisation and acceptance), then its verdict
    h->flag_dirty = true;
    k_fill_u64<<<1, 1, 0, s>>>(h->d_flag.ptr, (unsigned long long)NO_FAILURE);
    int rc = phase1<PrecF64>(h, E, b_e, sc, lambda, nullptr, E_inv, z, h->d_flag.ptr, s);
    if (rc) return rc;
    k_finalize_flag<<<1, 1, 0, s>>>(h->d_flag.ptr, h->d_failed_out.ptr);
    h->flag_dirty = false;
    int64_t fb = 0;
    BLS_HIP(hipMemcpyAsync(&fb, h->d_failed_out.ptr, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    BLS_HIP(hipStreamSynchronize(s));
    if (failed_block) *failed_block = fb;
    if (fb) return BLS_FAILED;
    const int64_t nr = h->nr, K = h->neb ? h->boffsets_e[h->neb - 1] + h->e_dims[h->neb - 1] : 0;   // rows of Y
    // the upper tiles of S and the K slices: enough workgroups to fill the device several times over
    const int nt1 = (int)((nr + SY_T - 1) / SY_T);
    if (h->sy_ntiles == 0 && nt1 > 0) {
        std::vector<int2> tiles;
        for (int j = 0; j < nt1; j++)
            for (int i = 0; i <= j; i++) tiles.push_back(make_int2(i, j));
        if ((rc = h->d_sy_tiles.upload(tiles))) return rc;
        h->sy_ntiles = (int64_t)tiles.size();
    }
    const int ntiles = (int)h->sy_ntiles;
    int64_t nslices = nslices_req > 0 ? nslices_req : std::max<int64_t>(1, (2048 + ntiles - 1) / std::max(ntiles, 1));
    nslices = std::min<int64_t>(nslices, std::max<int64_t>(1, (K + SY_KT - 1) / SY_KT));
    int64_t kc = (K + nslices - 1) / nslices;
    kc = (kc + SY_KT - 1) / SY_KT * SY_KT;
    nslices = kc > 0 ? (K + kc - 1) / kc : 1;
    const size_t ybytes = (size_t)K * (size_t)nr * sizeof(TY);
    const size_t pbytes = (size_t)nslices * ntiles * SY_T * SY_T * sizeof(TY);
    if (h->d_sy_Y.n * sizeof(double) < ybytes && (rc = h->d_sy_Y.alloc((ybytes + 7) / 8))) return rc;
    if (h->d_sy_partial.n * sizeof(double) < pbytes && (rc = h->d_sy_partial.alloc((pbytes + 7) / 8))) return rc;
    TY *Y = reinterpret_cast<TY *>(h->d_sy_Y.ptr);
    TY *part = reinterpret_cast<TY *>(h->d_sy_partial.ptr);
    h->timing.mark(1, s);
    if (K > 0 && nr > 0) {
        k_fill<TY><<<std::min<int64_t>(grid_for(K * nr, 256), 16384), 256, 0, s>>>(Y, K * nr, (TY)0);
        k_ybuild<TC, TY><<<grid_for(h->neb, 128), 128, 0, s>>>(
            h->neb, h->d_e_dims.ptr, h->d_eoffsets.ptr, h->d_boffsets_e.ptr, h->d_eliminated_ptr.ptr,
            h->d_entry_retained.ptr, h->d_coffsets.ptr, h->d_r_dims.ptr, h->d_boffsets_r.ptr, E_inv, C, nr, Y);
        h->timing.mark(2, s);
        k_syrk<TY><<<dim3((unsigned)ntiles, (unsigned)nslices), 256, 0, s>>>(Y, nr, K, kc, h->d_sy_tiles.ptr, ntiles, part);
        h->timing.mark(3, s);
        k_syrk_sum<TY><<<grid_for((int64_t)ntiles * SY_T * SY_T, 256), 256, 0, s>>>(part, (int)nslices, ntiles,
                                                                                     h->d_sy_tiles.ptr, nr, S);
    } else {
        h->timing.mark(2, s);
        h->timing.mark(3, s);
    }
    h->timing.mark(4, s);
    // g_S = b_r - sum_c C_c' z_e (the retained residual with x_e = z), Float64
    rc = retained_residual<TC>(h, g, b_r, z, C, nullptr, nullptr, s);
    if (rc) return rc;
    BLS_HIP(hipGetLastError());
    BLS_HIP(hipStreamSynchronize(s));
    if (h->timing.enabled)
        for (int i = 1; i < 4; i++) (void)hipEventElapsedTime(&h->timing.ms[i], h->timing.ev[i], h->timing.ev[i + 1]);
    return BLS_OK;
}
}  // namespace

extern "C" int bls_reduce_schur_dense_mfma_f64(bls_schur *h, double *S, double *g_S, const double *b_r,
                                               const double *b_e, const double *scaling_e, double lambda,
                                               const double *E_data, const double *C_data, double *E_inv, double *z,
                                               int64_t nslices, int64_t *failed_block, void *stream) {
    return reduce_dense_mfma<double, double>(h, S, g_S, b_r, b_e, scaling_e, lambda, E_data, C_data, E_inv, z,
                                             failed_block, nslices, (hipStream_t)stream);
}

extern "C" int bls_reduce_schur_dense_mfma_mixed(bls_schur *h, double *S, double *g_S, const double *b_r,
                                                 const double *b_e, const double *scaling_e, double lambda,
                                                 const double *E_data, const float *C_data, double *E_inv, double *z,
                                                 int64_t nslices, int64_t *failed_block, void *stream) {
    return reduce_dense_mfma<float, float>(h, S, g_S, b_r, b_e, scaling_e, lambda, E_data, C_data, E_inv, z,
                                           failed_block, nslices, (hipStream_t)stream);
}

// ---- the reduction that allocates nothing ------------------------------------------------------------------
// Scratch of bls_reduce_schur_sparse_f64_noalloc: word 0 the failure flag, then the row sums of split items.
// Each GPU holds its own copy for the life of the process; host code uses it only through the address
// hipGetSymbolAddress returns on that device (the array's name is a host-side stand-in).
constexpr int64_t NOALLOC_SCRATCH_DOUBLES = int64_t(1) << 21;   // 16 MiB
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];

namespace {
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// The device address of `scratch` on the current device, looked up once per device.
int scratch_on_current_device(double **out) {
    int dev = 0;
    BLS_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fail(BLS_ERR_UNSUPPORTED, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        BLS_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) return fail(BLS_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
        g_scratch_ptr[dev] = static_cast<double *>(ptr);
    }
    *out = g_scratch_ptr[dev];
    return BLS_OK;
}
}  // namespace

extern "C" int bls_reduce_schur_sparse_f64_noalloc(bls_schur *h, double *nzval, double *g_S, const double *b_r,
                                                   const double *b_e, const double *scaling_e, double lambda,
                                                   const double *E_data, const double *C_data, double *E_inv,
                                                   double *z, int64_t *d_failed_block, void *stream) {
    if (!h) return fail(BLS_ERR_ARG, "null handle");
    const int64_t need = 1 + (int64_t)h->d_partial.n;
    if (need > NOALLOC_SCRATCH_DOUBLES)
        return fail(BLS_ERR_ARG, "reduce_schur (no allocation): the work list's split items need " +
                                     std::to_string(need) + " doubles of scratch, more than the " +
                                     std::to_string(NOALLOC_SCRATCH_DOUBLES) +
                                     " the library holds; use longer parts (bls_schur_set_tuning) or "
                                     "bls_reduce_schur_sparse_f64_async");
    double *p = nullptr;
    int rc = scratch_on_current_device(&p);
    if (rc) return rc;
    const P2Scratch ws{reinterpret_cast<unsigned long long *>(p), p + 1};
    return reduce_async<PrecF64>(h, 1, nzval, g_S, b_r, b_e, scaling_e, lambda, E_data, C_data, E_inv, z, nullptr,
                                 nullptr, d_failed_block, (hipStream_t)stream, &ws);
}

extern "C" int bls_reduce_schur_dense_f64_async(bls_schur *h, double *S, double *g_S, const double *b_r,
                                                const double *b_e, const double *scaling_e, double lambda,
                                                const double *E_data, const double *C_data, double *E_inv, double *z,
                                                const double *s_r, const double *s_e, int64_t *d_failed_block,
                                                void *stream) {
    return reduce_async<PrecF64>(h, 2, S, g_S, b_r, b_e, scaling_e, lambda, E_data, C_data, E_inv, z, s_r, s_e,
                                 d_failed_block, (hipStream_t)stream);
}

extern "C" int bls_backsubstitute_schur_f64(bls_schur *h, double *x_e, const double *x_r, const double *C_data,
                                            const double *E_inv, const double *z, const double *s_r,
                                            const double *s_e, void *stream) {
    return backsub<double, double, double>(h, x_e, x_r, C_data, E_inv, z, s_r, s_e, (hipStream_t)stream);
}

extern "C" int bls_copyto_csc_f64(int64_t n, double *dest, const double *src, const int64_t *map, void *stream) {
    return copyto<double>(n, dest, src, map, (hipStream_t)stream);
}

// ---- Float32 (SchurBlockHessian{Float32})
extern "C" int bls_reset_schur_sparse_f32(bls_schur *h, float *nzval, const float *R_nzval, const double *scaling_r,
                                          double lambda, const double *s_r, void *stream) {
    return reset_sparse<float, float>(h, nzval, R_nzval, scaling_r, lambda, s_r, (hipStream_t)stream);
}

extern "C" int bls_reset_schur_dense_f32(bls_schur *h, float *S, const float *R, const double *scaling_r,
                                         double lambda, const double *s_r, void *stream) {
    return reset_dense<float, float>(h, S, R, scaling_r, lambda, s_r, (hipStream_t)stream);
}

extern "C" int bls_reduce_schur_sparse_f32(bls_schur *h, float *nzval, double *g_S, const double *b_r,
                                           const double *b_e, const double *scaling_e, double lambda,
                                           const float *E_data, const float *C_data, float *E_inv, double *z,
                                           const double *s_r, const double *s_e, int64_t *failed_block, void *stream) {
    return reduce_sync<PrecF32>(h, 1, nzval, g_S, b_r, b_e, scaling_e, lambda, E_data, C_data, E_inv, z, s_r, s_e,
                                failed_block, (hipStream_t)stream);
}

extern "C" int bls_reduce_schur_dense_f32(bls_schur *h, float *S, double *g_S, const double *b_r, const double *b_e,
                                          const double *scaling_e, double lambda, const float *E_data,
                                          const float *C_data, float *E_inv, double *z, const double *s_r,
                                          const double *s_e, int64_t *failed_block, void *stream) {
    return reduce_sync<PrecF32>(h, 2, S, g_S, b_r, b_e, scaling_e, lambda, E_data, C_data, E_inv, z, s_r, s_e,
                                failed_block, (hipStream_t)stream);
}

extern "C" int bls_reduce_schur_sparse_f32_async(bls_schur *h, float *nzval, double *g_S, const double *b_r,
                                                 const double *b_e, const double *scaling_e, double lambda,
                                                 const float *E_data, const float *C_data, float *E_inv, double *z,
                                                 const double *s_r, const double *s_e, int64_t *d_failed_block,
                                                 void *stream) {
    return reduce_async<PrecF32>(h, 1, nzval, g_S, b_r, b_e, scaling_e, lambda, E_data, C_data, E_inv, z, s_r, s_e,
                                 d_failed_block, (hipStream_t)stream);
}

extern "C" int bls_backsubstitute_schur_f32(bls_schur *h, double *x_e, const double *x_r, const float *C_data,
                                            const float *E_inv, const double *z, const double *s_r,
                                            const double *s_e, void *stream) {
    return backsub<float, float, float>(h, x_e, x_r, C_data, E_inv, z, s_r, s_e, (hipStream_t)stream);
}

extern "C" int bls_copyto_csc_f32(int64_t n, float *dest, const float *src, const int64_t *map, void *stream) {
    return copyto<float>(n, dest, src, map, (hipStream_t)stream);
}

// ---- mixed: Float32 blocks in, Float64 factorisation and reduced system
extern "C" int bls_reset_schur_sparse_mixed(bls_schur *h, double *nzval, const float *R_nzval,
                                            const double *scaling_r, double lambda, const double *s_r, void *stream) {
    return reset_sparse<double, float>(h, nzval, R_nzval, scaling_r, lambda, s_r, (hipStream_t)stream);
}

extern "C" int bls_reset_schur_dense_mixed(bls_schur *h, double *S, const float *R, const double *scaling_r,
                                           double lambda, const double *s_r, void *stream) {
    return reset_dense<double, float>(h, S, R, scaling_r, lambda, s_r, (hipStream_t)stream);
}

extern "C" int bls_reduce_schur_sparse_mixed(bls_schur *h, double *nzval, double *g_S, const double *b_r,
                                             const double *b_e, const double *scaling_e, double lambda,
                                             const float *E_data, const float *C_data, double *E_inv, double *z,
                                             const double *s_r, const double *s_e, int64_t *failed_block,
                                             void *stream) {
    return reduce_sync<PrecMixed>(h, 1, nzval, g_S, b_r, b_e, scaling_e, lambda, E_data, C_data, E_inv, z, s_r, s_e,
                                  failed_block, (hipStream_t)stream);
}

extern "C" int bls_reduce_schur_dense_mixed(bls_schur *h, double *S, double *g_S, const double *b_r,
                                            const double *b_e, const double *scaling_e, double lambda,
                                            const float *E_data, const float *C_data, double *E_inv, double *z,
                                            const double *s_r, const double *s_e, int64_t *failed_block,
                                            void *stream) {
    return reduce_sync<PrecMixed>(h, 2, S, g_S, b_r, b_e, scaling_e, lambda, E_data, C_data, E_inv, z, s_r, s_e,
                                  failed_block, (hipStream_t)stream);
}

extern "C" int bls_reduce_schur_sparse_mixed_async(bls_schur *h, double *nzval, double *g_S, const double *b_r,
                                                   const double *b_e, const double *scaling_e, double lambda,
                                                   const float *E_data, const float *C_data, double *E_inv,
                                                   double *z, const double *s_r, const double *s_e,
                                                   int64_t *d_failed_block, void *stream) {
    return reduce_async<PrecMixed>(h, 1, nzval, g_S, b_r, b_e, scaling_e, lambda, E_data, C_data, E_inv, z, s_r, s_e,
                                   d_failed_block, (hipStream_t)stream);
}

extern "C" int bls_backsubstitute_schur_mixed(bls_schur *h, double *x_e, const double *x_r, const float *C_data,
                                              const double *E_inv, const double *z, const double *s_r,
                                              const double *s_e, void *stream) {
    return backsub<float, double, double>(h, x_e, x_r, C_data, E_inv, z, s_r, s_e, (hipStream_t)stream);
}

extern "C" int bls_scatter_map_create(int64_t n, const int64_t *scatter_map, int64_t **d_map) {
    if (n < 0 || !d_map) return fail(BLS_ERR_ARG, "bls_scatter_map_create: bad arguments");
    *d_map = nullptr;
    if (n == 0) return BLS_OK;
    BLS_HIP(hipMalloc((void **)d_map, n * sizeof(int64_t)));
    BLS_HIP(hipMemcpy(*d_map, scatter_map, n * sizeof(int64_t), hipMemcpyHostToDevice));
    return BLS_OK;
}

extern "C" int bls_schur_set_timing(bls_schur *h, int enable) {
    if (!h) return fail(BLS_ERR_ARG, "null handle");
    if (enable) return h->timing.enable();
    h->timing.enabled = false;
    return BLS_OK;
}

extern "C" int bls_schur_kernel_times(bls_schur *h, float *times, int n) {
    if (!h || !times) return fail(BLS_ERR_ARG, "bad arguments");
    if (h->timing.enabled) (void)hipEventElapsedTime(&h->timing.ms[0], h->timing.ev[0], h->timing.ev[1]);
    for (int i = 0; i < n && i < 4; i++) times[i] = h->timing.ms[i];
    return BLS_OK;
}

// ---- which phase-2 kernel runs --------------------------------------------------------------------------------------
extern "C" int bls_schur_set_phase2_mode(bls_schur *h, int mode) {
    if (!h) return fail(BLS_ERR_ARG, "null handle");
    if (mode < 0 || mode > 2) return fail(BLS_ERR_ARG, "phase-2 mode must be 0 (automatic), 1 (work list) or 2 (tiles)");
    if (mode == 2 && !(h->mode == 1 && !h->wide && h->tiles.nunits > 0))
        return fail(BLS_ERR_UNSUPPORTED, "the tiled phase 2 does not fit this structure: " +
                                             (h->mode != 1 ? std::string("no sparse layout") : h->tiles_why));
    h->p2_mode = mode;
    return BLS_OK;
}

// info[0]: the kernel in use (1 work list, 2 tiles); info[1..3]: the tile plan's tiles, landmark entries and MFMAs
extern "C" int bls_schur_phase2_info(bls_schur *h, int64_t *info) {
    if (!h || !info) return fail(BLS_ERR_ARG, "bad arguments");
    info[0] = h->use_tiles() ? 2 : 1;
    info[1] = h->tiles.nunits;
    info[2] = h->tiles.nentries;
    info[3] = h->tiles.nmfma;
    return BLS_OK;
}

// ---- the shape of the phase-2 work list ------------------------------------------------------------------------
extern "C" int bls_schur_set_tuning(bls_schur *h, int64_t part, int64_t lanes) {
    if (!h) return fail(BLS_ERR_ARG, "null handle");
    if (part < 16 || part > (1 << 20) || lanes < 16 || lanes > P2_THREADS)
        return fail(BLS_ERR_ARG, "tuning: part must be 16..2^20 landmarks and lanes 16..256");
    h->p2_part = part;
    h->p2_lanes = lanes;
    if (h->mode == 1) return build_sparse_phase2(h);
    if (h->mode == 2) return build_dense_phase2(h);
    return BLS_OK;
}

extern "C" int bls_schur_get_tuning(bls_schur *h, int64_t *part, int64_t *lanes) {
    if (!h) return fail(BLS_ERR_ARG, "null handle");
    if (part) *part = h->p2_part;
    if (lanes) *lanes = h->p2_lanes;
    return BLS_OK;
}

// Time three reductions (after one untimed) per candidate shape on the caller's own data and keep the shape with the
// fastest median. Every shape computes the same sums in a fixed order of its own, so any of them is deterministic.
extern "C" int bls_schur_autotune_f64(bls_schur *h, int64_t ncand, const int64_t *parts, const int64_t *lanes,
                                      double *S, double *g_S, const double *b_r, const double *b_e,
                                      const double *scaling_e, double lambda, const double *E_data,
                                      const double *C_data, double *E_inv, double *z, const double *s_r,
                                      const double *s_e, float *best_ms, void *stream) {
    if (!h || ncand < 1 || !parts || !lanes) return fail(BLS_ERR_ARG, "bls_schur_autotune_f64: bad arguments");
    if (h->mode != 1 && h->mode != 2) return fail(BLS_ERR_STATE, "handle has no reduced-system layout");
    hipStream_t st = (hipStream_t)stream;
    hipEvent_t e0, e1;
    BLS_HIP(hipEventCreate(&e0));
    BLS_HIP(hipEventCreate(&e1));
    float best = -1.0f;
    int64_t bp = h->p2_part, bl = h->p2_lanes;
    int rc = BLS_OK;
    // the median of three timed reductions after an untimed one
    auto time_one = [&](float *ms) {
        float t[3] = {0.0f, 0.0f, 0.0f};
        for (int rep = 0; rep < 4 && rc == BLS_OK; rep++) {
            if (rep > 0 && hipEventRecord(e0, st) != hipSuccess) rc = fail(BLS_ERR_HIP, "hipEventRecord");
            if (rc == BLS_OK)
                rc = reduce_async<PrecF64>(h, h->mode, S, g_S, b_r, b_e, scaling_e, lambda, E_data, C_data, E_inv, z,
                                           s_r, s_e, h->d_failed_out.ptr, st);
            if (rc == BLS_OK && rep > 0 &&
                (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                 hipEventElapsedTime(&t[rep - 1], e0, e1) != hipSuccess))
                rc = fail(BLS_ERR_HIP, "autotune timing");
        }
        std::sort(t, t + 3);
        *ms = t[1];
    };
    // the work-list shapes, then (when the structure fits them) the tiles
    const int mode0 = h->p2_mode;
    h->p2_mode = 1;
    for (int64_t c = 0; c < ncand && rc == BLS_OK; c++) {
        if ((rc = bls_schur_set_tuning(h, parts[c], lanes[c]))) break;
        float ms = 0.0f;
        time_one(&ms);
        if (rc == BLS_OK && (best < 0 || ms < best)) {
            best = ms;
            bp = parts[c];
            bl = lanes[c];
        }
    }
    int mode = mode0 == 2 ? 2 : 1;
    if (rc == BLS_OK && mode0 != 1 && h->mode == 1 && !h->wide && h->tiles.nunits > 0) {
        h->p2_mode = 2;
        float ms = 0.0f;
        time_one(&ms);
        if (rc == BLS_OK && ms < best) {
            best = ms;
            mode = mode0;
        }
    }
    h->p2_mode = mode0;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    if (best_ms) *best_ms = best;
    h->p2_mode = mode;
    return bls_schur_set_tuning(h, bp, bl);
}
