// Block-Jacobi preconditioned CG on the reduced camera system(s), batched over independent problems.
//
// The reference has no PCG (its Krylov extension runs LSMR on the QR path with the BlockJacobi preconditioner of
// ext/KrylovExt.jl:88-95, from factor_diagonal_blocks! / solve_diagonal_blocks!, src/diagonal_blocks.jl:6-61);
// this is textbook PCG (Saad, Iterative Methods for Sparse Linear Systems, 2nd ed., Alg. 9.1) with that
// preconditioner, judged by the true residual ||S x - g|| / ||g|| in FP64.
//
// Layout: S arrives in the upper-block CSC of build_schur_pattern; at each solve it is gathered once (the copyto!
// kernel) into a symmetric block-row layout (BSR: every block (i, j) of the full matrix, row-major, blocks of a
// block row contiguous). An iteration is three kernels, each skipping converged problems:
//   k_bsr_spmv_u / k_bsr_spmv_v  q = S p, one wavefront per block row (lane (b, r) takes row r of every
//                 (64 / d_i)-th block from the b-th: contiguous runs of the row's values; fixed-order sums), and the
//                 block row's partial p.q (u: one block dim throughout, v: dims that differ);
//   k_pcg_step1   per group of up to 256 blocks of one problem: alpha = rz / (sum of the problem's p.q partials,
//                 in a fixed order), x += alpha p, r -= alpha q, z = M^-1 r per block, partials r.z and r.r;
//   k_pcg_step2   beta = rz_new / rz from the same fixed-order sums, p = z + beta p; the problem's first group
//                 records rz_new, the iteration count and the convergence test.
// Every group of a problem forms the same sums in the same order, so the iteration is deterministic. rz is kept
// per iteration parity (rz2[it & 1] read, rz2[(it + 1) & 1] written) so that no group overwrites a value another
// still reads. A fixed even number of iterations is captured once in a hipGraph on the handle's own stream and
// replayed until every problem has converged.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "bls_internal.h"

using namespace bls;

namespace {

constexpr int UPD_THREADS = 512;
constexpr int GRP_BLOCKS = 256;     // blocks per group of k_pcg_step1 / k_pcg_step2 (one thread each)
constexpr int ITERS_PER_GRAPH = 8;  // even: the rz parity of every problem is 0 at each replay's start
constexpr int MAXD = 16;
constexpr int PC_GMAX = 4;          // blocks per preconditioner group of the learned correction

struct ProbState {
    double rz2[2], gnorm, relres;
    int64_t iters;
    int32_t active, failed;
};

struct PcgArgs {
    int64_t nr;
    const int32_t *row_block;          // block row of each scalar row
    const int32_t *bsr_rowptr, *bsr_col;
    const int64_t *bsr_voff;
    const int32_t *dims;
    const int64_t *bo;                 // 0-based scalar offset of each block
    const int32_t *prob_of_block;
    const int64_t *prob_bptr;          // block range of each problem (0-based, nprob + 1)
    const double *vals;                // BSR values
    const double *L;                   // preconditioner factors (lower, column-major d x d per block)
    const int64_t *loff;               // offset of each block's factor
    // preconditioner groups (the learned correction): gsize[b] = the number of consecutive blocks of the group that
    // block b starts (1 .. PC_GMAX), 0 for the other blocks of a group; gloff[b] the group's factor offset.
    // nullptr: every block alone
    const int32_t *gsize;
    const int64_t *gloff;
    double *x, *r, *z, *p, *q, *g;
    double *pq_part, *rz_part, *rr_part;   // per block
    const int32_t *grp_prob;           // problem of each group
    const int64_t *grp_b0, *grp_b1;    // block range of each group
    int64_t nrb;
    ProbState *st;
    int *nactive;
    double tol;
    int64_t maxit;
};

// sum over the 64 lanes; lane 0's result is formed in a fixed order (the only one used)
__device__ inline double wave_sum64(double v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

// Block dims that differ between block rows (d_i, d_j <= D): one wavefront per block row i; lane (b, r) =
// (lane / d_i, lane % d_i) takes row r of every (64 / d_i)-th block from the b-th, reading that row's d_j contiguous
// values, so a wavefront streams contiguous runs of the row's blocks; the per-lane sums are added in a fixed order.
template <int D>
__global__ __launch_bounds__(256) void k_bsr_spmv_v(PcgArgs a, const double *xin, double *yout, double *pq_part,
                                                     int only_active) {
    __shared__ double red[4][64];
    const int w = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int64_t i = (int64_t)blockIdx.x * 4 + w;
    if (i >= a.nrb) return;
    if (only_active && !a.st[a.prob_of_block[i]].active) return;
    const int di = a.dims[i];
    const int nb = 64 / di;   // blocks per pass
    const int bsub = lane / di, r = lane - bsub * di;
    const int k0 = a.bsr_rowptr[i], k1 = a.bsr_rowptr[i + 1];
    double acc = 0.0;
    if (bsub < nb) {
        for (int k = k0 + bsub; k < k1; k += nb) {
            const int j = a.bsr_col[k];
            const int dj = a.dims[j];
            const double *B = a.vals + a.bsr_voff[k] + (int64_t)r * dj;   // row r of the d_i x d_j block
            const double *xj = xin + a.bo[j];
#pragma unroll
            for (int c = 0; c < D; c++)
                if (c < dj) acc += B[c] * xj[c];
        }
    }
    red[w][lane] = acc;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    double y = 0.0, pq = 0.0;
    const int64_t b = a.bo[i];
    if (lane < di) {
        for (int q = 0; q < nb; q++) y += red[w][q * di + lane];
        yout[b + lane] = y;
        pq = xin[b + lane] * y;
    }
    if (pq_part) {
        pq = wave_sum64(pq);
        if (lane == 0) pq_part[i] = pq;
    }
}

// The same product when every block is D x D (the camera systems): lane (b, r) of a wavefront takes row r of the
// b-th block of the pass (64 / D blocks per pass), so a pass reads a contiguous run of the block row's values, row
// segments of D doubles as 16-byte loads; the lanes of one row r are then summed in fixed lane order through LDS.
template <int D>
__global__ __launch_bounds__(256) void k_bsr_spmv_u(PcgArgs a, const double *xin, double *yout, double *pq_part,
                                                     int only_active) {
    constexpr int NB = 64 / D;          // blocks per pass
    __shared__ double red[4][64];
    const int w = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int64_t i = (int64_t)blockIdx.x * 4 + w;
    if (i >= a.nrb) return;
    if (only_active && !a.st[a.prob_of_block[i]].active) return;
    const int bsub = lane / D, r = lane - bsub * D;
    const int k0 = a.bsr_rowptr[i], k1 = a.bsr_rowptr[i + 1];
    const double *V = a.vals + a.bsr_voff[k0];   // the row's blocks are contiguous, D*D values each
    double acc = 0.0;
    if (bsub < NB) {
        for (int k = k0 + bsub; k < k1; k += NB) {
            const double *B = V + (int64_t)(k - k0) * (D * D) + r * D;
            const double *xj = xin + a.bo[a.bsr_col[k]];
            if constexpr (D % 2 == 0) {
#pragma unroll
                for (int c = 0; c < D; c += 2) {
                    const double2 bv = *reinterpret_cast<const double2 *>(B + c);
                    const double2 xv = *reinterpret_cast<const double2 *>(xj + c);
                    acc += bv.x * xv.x;
                    acc += bv.y * xv.y;
                }
            } else {
#pragma unroll
                for (int c = 0; c < D; c++) acc += B[c] * xj[c];
            }
        }
    }
    red[w][lane] = acc;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    double y = 0.0, pq = 0.0;
    const int64_t b = a.bo[i];
    if (lane < D) {
#pragma unroll
        for (int q = 0; q < NB; q++) y += red[w][q * D + lane];
        yout[b + lane] = y;
        pq = xin[b + lane] * y;
    }
    if (pq_part) {
        pq = wave_sum64(pq);
        if (lane == 0) pq_part[i] = pq;
    }
}

// z_b = (L L')^-1 r_b for one block of d <= DM (the kernels' UD): the loops unrolled to DM and predicated on d, so
// y stays in registers; the same operations in the same order as block_solve
template <int DM> __device__ inline void block_solve_upto(const double *L, int d, const double *rb, double *zb) {
    double y[DM];
#pragma unroll
    for (int i = 0; i < DM; i++) y[i] = i < d ? rb[i] : 0.0;
#pragma unroll
    for (int j = 0; j < DM; j++) {
        if (j < d) {
            const double yj = y[j] / L[j * d + j];
            y[j] = yj;
#pragma unroll
            for (int i = j + 1; i < DM; i++)
                if (i < d) y[i] -= L[j * d + i] * yj;
        }
    }
#pragma unroll
    for (int j = DM - 1; j >= 0; j--) {
        if (j < d) {
            double t = y[j];
#pragma unroll
            for (int i = j + 1; i < DM; i++)
                if (i < d) t -= L[j * d + i] * y[i];
            y[j] = t / L[j * d + j];
        }
    }
#pragma unroll
    for (int i = 0; i < DM; i++)
        if (i < d) zb[i] = y[i];
}

// z_b = (L L')^-1 r_b for one block (solve_diagonal_blocks!, src/diagonal_blocks.jl:33-61), any size up to a
// group of PC_GMAX blocks of MAXD (y is indexed at run time, so it lives in private memory)
__device__ inline void block_solve(const double *L, int d, const double *rb, double *zb) {
    double y[PC_GMAX * MAXD];
    for (int i = 0; i < d; i++) y[i] = rb[i];
    for (int j = 0; j < d; j++) {
        const double yj = y[j] / L[j * d + j];
        y[j] = yj;
        for (int i = j + 1; i < d; i++) y[i] -= L[j * d + i] * yj;
    }
    for (int j = d - 1; j >= 0; j--) {
        double t = y[j];
        for (int i = j + 1; i < d; i++) t -= L[j * d + i] * y[i];
        y[j] = t / L[j * d + j];
    }
    for (int i = 0; i < d; i++) zb[i] = y[i];
}

// Dimension of the preconditioner group that block b starts (0 when b is not a group's first block) and its factor.
__device__ inline int group_dim(const PcgArgs &a, int64_t b) {
    if (!a.gsize) return a.dims[b];
    const int gs = a.gsize[b];
    int d = 0;
    for (int k = 0; k < gs; k++) d += a.dims[b + k];
    return d;
}
__device__ inline const double *group_L(const PcgArgs &a, int64_t b) { return a.gsize ? a.L + a.gloff[b] : a.L + a.loff[b]; }

// fixed-order workgroup sum (deterministic)
__device__ inline double wg_sum(double v, double *red) {
    const int t = threadIdx.x;
    red[t] = v;
    __syncthreads();
    for (int s = UPD_THREADS / 2; s > 0; s >>= 1) {
        if (t < s) red[t] += red[t + s];
        __syncthreads();
    }
    const double out = red[0];
    __syncthreads();
    return out;
}

// fixed-order sum over the 256 threads of a group
__device__ inline double grp_sum(double v, double *red) {
    const int t = threadIdx.x;
    red[t] = v;
    __syncthreads();
    for (int s = GRP_BLOCKS / 2; s > 0; s >>= 1) {
        if (t < s) red[t] += red[t + s];
        __syncthreads();
    }
    const double out = red[0];
    __syncthreads();
    return out;
}

// sum of a per-block partial over the blocks [b0, b1) of one problem, in a fixed order
__device__ inline double problem_sum(const double *part, int64_t b0, int64_t b1, double *red) {
    double s = 0.0;
    for (int64_t b = b0 + threadIdx.x; b < b1; b += GRP_BLOCKS) s += part[b];
    return grp_sum(s, red);
}

// (Re)start of CG from the current x, per group of up to GRP_BLOCKS blocks of one problem (the groups of
// k_pcg_step1): r = g - S x (q holds S x; first = 1: x = 0, r = g), z = M^-1 r, p = z, and per block the partials
// g.g (in pq_part, free between iterations), r.r and r.z; k_pcg_restart_fin forms the problem's sums in a fixed
// order and decides from the true relative residual whether the problem (re)enters the iteration.
// UD > 0: every block has dimension at most UD and there are no groups (the solve in registers); 0: any
template <int UD>
__global__ __launch_bounds__(GRP_BLOCKS) void k_pcg_restart_grp(PcgArgs a, int first) {
    const int pb = a.grp_prob[blockIdx.x];
    const bool ok = !a.st[pb].failed;
    const int64_t b = a.grp_b0[blockIdx.x] + threadIdx.x;
    const bool in = b < a.grp_b1[blockIdx.x];
    const int64_t o = in ? a.bo[b] : 0;
    const int d = in ? a.dims[b] : 0;
    double gg = 0.0, rr = 0.0;
    for (int i = 0; i < d; i++) {
        const double gv = a.g[o + i];
        if (first) a.x[o + i] = 0.0;
        const double rv = first ? gv : gv - a.q[o + i];
        a.r[o + i] = rv;
        gg += gv * gv;
        rr += rv * rv;
    }
    if (a.gsize) __syncthreads();   // a group's rows are complete before its first block solves them
    if constexpr (UD > 0) {
        if (ok && in) block_solve_upto<UD>(a.L + a.loff[b], d, a.r + o, a.z + o);
    } else {
        const int gd = in ? group_dim(a, b) : 0;
        if (ok && gd) block_solve(group_L(a, b), gd, a.r + o, a.z + o);
    }
    if (a.gsize) __syncthreads();   // and its z before the other blocks read theirs
    if (!in) return;
    double rz = 0.0;
    for (int i = 0; i < d; i++) {
        const double zv = ok ? a.z[o + i] : 0.0;
        a.p[o + i] = zv;
        rz += a.r[o + i] * zv;
    }
    a.pq_part[b] = gg;
    a.rr_part[b] = rr;
    a.rz_part[b] = rz;
}

__global__ __launch_bounds__(GRP_BLOCKS) void k_pcg_restart_fin(PcgArgs a, int first) {
    __shared__ double red[GRP_BLOCKS];
    const int pb = blockIdx.x;
    const int64_t b0 = a.prob_bptr[pb], b1 = a.prob_bptr[pb + 1];
    ProbState &ps = a.st[pb];
    const double gg = problem_sum(a.pq_part, b0, b1, red);
    const double rr = problem_sum(a.rr_part, b0, b1, red);
    const double rz = problem_sum(a.rz_part, b0, b1, red);
    if (threadIdx.x == 0) {
        const bool ok = !ps.failed;
        if (first) { ps.gnorm = sqrt(gg); ps.iters = 0; }
        ps.rz2[0] = rz;
        ps.relres = ps.gnorm > 0 ? sqrt(rr) / ps.gnorm : 0.0;
        const int was = first ? 0 : ps.active;
        ps.active = (ok && ps.relres > a.tol && ps.iters < a.maxit && isfinite(ps.relres)) ? 1 : 0;
        if (ps.active && !was) atomicAdd(a.nactive, 1);
        if (!ps.active && was) atomicSub(a.nactive, 1);
    }
}

// PCG iteration, after q = S p (Saad Alg. 9.1, lines 3-5): alpha, x, r, z = M^-1 r, partial r.z and r.r. UD as in
// k_pcg_restart_grp.
template <int UD>
__global__ __launch_bounds__(GRP_BLOCKS) void k_pcg_step1(PcgArgs a, int parity) {
    __shared__ double red[GRP_BLOCKS];
    const int pb = a.grp_prob[blockIdx.x];
    const ProbState &ps = a.st[pb];
    if (!ps.active) return;
    const double pq = problem_sum(a.pq_part, a.prob_bptr[pb], a.prob_bptr[pb + 1], red);
    const double alpha = ps.rz2[parity] / pq;
    const int64_t b = a.grp_b0[blockIdx.x] + threadIdx.x;
    const bool in = b < a.grp_b1[blockIdx.x];
    const int64_t o = in ? a.bo[b] : 0;
    double rr = 0.0;
    if (in) {
        const int d = a.dims[b];
        for (int i = 0; i < d; i++) {
            a.x[o + i] += alpha * a.p[o + i];
            const double rv = a.r[o + i] - alpha * a.q[o + i];
            a.r[o + i] = rv;
            rr += rv * rv;
        }
    }
    // a group's other blocks are updated by the next threads (preconditioner groups never cross an update group of
    // GRP_BLOCKS: they start at multiples of their size, a divisor of GRP_BLOCKS, from their problem's first block)
    if (a.gsize) __syncthreads();
    if (!in) return;
    const int gd = UD > 0 ? a.dims[b] : group_dim(a, b);
    if (gd) {
        if constexpr (UD > 0) block_solve_upto<UD>(a.L + a.loff[b], gd, a.r + o, a.z + o);
        else block_solve(group_L(a, b), gd, a.r + o, a.z + o);
        // r.z per block, as the plain preconditioner sums it (the group's other blocks' partials by this thread)
        const int gs = a.gsize ? a.gsize[b] : 1;
        int64_t oo = o;
        for (int k = 0; k < gs; k++) {
            const int dk = a.dims[b + k];
            double rz = 0.0;
            for (int i = 0; i < dk; i++) rz += a.r[oo + i] * a.z[oo + i];
            a.rz_part[b + k] = rz;
            oo += dk;
        }
    }
    a.rr_part[b] = rr;
}

// PCG iteration (lines 6-7): beta, p = z + beta p; the problem's first group records rz, the iteration count and
// the convergence test on the recurrence residual (confirmed with the true residual after the iterations)
__global__ __launch_bounds__(GRP_BLOCKS) void k_pcg_step2(PcgArgs a, int parity) {
    __shared__ double red[GRP_BLOCKS];
    const int pb = a.grp_prob[blockIdx.x];
    ProbState &ps = a.st[pb];
    if (!ps.active) return;
    const int64_t b0 = a.prob_bptr[pb], b1 = a.prob_bptr[pb + 1];
    const double rz = problem_sum(a.rz_part, b0, b1, red);
    const double rr = problem_sum(a.rr_part, b0, b1, red);
    const double beta = rz / ps.rz2[parity];
    const int64_t b = a.grp_b0[blockIdx.x] + threadIdx.x;
    if (b < a.grp_b1[blockIdx.x]) {
        const int d = a.dims[b];
        const int64_t o = a.bo[b];
        for (int i = 0; i < d; i++) a.p[o + i] = a.z[o + i] + beta * a.p[o + i];
    }
    if (a.grp_b0[blockIdx.x] == b0 && threadIdx.x == 0) {
        ps.rz2[parity ^ 1] = rz;
        ps.iters += 1;
        ps.relres = ps.gnorm > 0 ? sqrt(rr) / ps.gnorm : 0.0;
        if (!(ps.relres > a.tol) || ps.iters >= a.maxit || !isfinite(ps.relres)) {
            ps.active = 0;
            atomicSub(a.nactive, 1);
        }
    }
}

// Preconditioner blocks: the diagonal blocks of S (upper triangle mirrored), optionally replaced by a corrected
// block, then factored in place (factor_diagonal_blocks!, src/diagonal_blocks.jl:6-30). A block that is not
// positive definite marks its problem failed.
__global__ void k_pcg_factor(PcgArgs a, int64_t nblocks, const int64_t *diag_voff, double *Lout,
                             const double *Mblocks) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nblocks) return;
    const int d = a.dims[b];
    double *L = Lout + a.loff[b];
    if (Mblocks) {
        for (int q = 0; q < d * d; q++) L[q] = Mblocks[a.loff[b] + q];
    } else {
        const double *B = a.vals + diag_voff[b];   // row-major full block
        for (int c = 0; c < d; c++)
            for (int r = 0; r < d; r++) L[c * d + r] = B[r * d + c];
    }
    for (int j = 0; j < d; j++) {
        double s = L[j * d + j];
        for (int k = 0; k < j; k++) s -= L[k * d + j] * L[k * d + j];
        if (!(s > 0.0) || !isfinite(s)) {
            a.st[a.prob_of_block[b]].failed = 1;
            return;
        }
        const double ljj = sqrt(s);
        L[j * d + j] = ljj;
        for (int i = j + 1; i < d; i++) {
            double t = L[j * d + i];
            for (int k = 0; k < j; k++) t -= L[k * d + i] * L[k * d + j];
            L[j * d + i] = t / ljj;
        }
    }
}

// ---------------------------------------------------------------------------------------------------------
// Preconditioner groups: factor each group's full block (Mblocks at gloff, gd x gd column-major) into Lout at gloff.
__global__ void k_pcg_factor_groups(PcgArgs a, int64_t nblocks, double *Lout, const double *Mblocks) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nblocks) return;
    const int d = group_dim(a, b);
    if (!d) return;
    double *L = Lout + a.gloff[b];
    for (int q = 0; q < d * d; q++) L[q] = Mblocks[a.gloff[b] + q];
    for (int j = 0; j < d; j++) {
        double s = L[j * d + j];
        for (int k = 0; k < j; k++) s -= L[k * d + j] * L[k * d + j];
        if (!(s > 0.0) || !isfinite(s)) {
            a.st[a.prob_of_block[b]].failed = 1;
            return;
        }
        const double ljj = sqrt(s);
        L[j * d + j] = ljj;
        for (int i = j + 1; i < d; i++) {
            double t = L[j * d + i];
            for (int k = 0; k < j; k++) t -= L[k * d + i] * L[k * d + j];
            L[j * d + i] = t / ljj;
        }
    }
}

// The same, one wavefront per group (blocks b that start one; the others return): lane c holds column c of the
// upper triangle in registers, and at step j the pivot and row j of U are broadcast by readlane; each entry receives
// its updates in the order the column-by-column form above subtracts them, and is divided by its pivot, so the
// factor is bitwise that of k_pcg_factor_groups (and of k_pcg_factor's for a block-diagonal group).
constexpr int PCW = PC_GMAX * MAXD;   // largest group dimension
__global__ __launch_bounds__(64) void k_pcg_factor_groups_w(PcgArgs a, int64_t nblocks, double *Lout,
                                                           const double *Mblocks) {
    const int64_t b = blockIdx.x;
    if (b >= nblocks) return;
    const int gd = group_dim(a, b);
    if (!gd) return;
    const int lane = threadIdx.x;
    const double *M = Mblocks + a.gloff[b];
    double col[PCW];
#pragma unroll
    for (int r = 0; r < PCW; r++) col[r] = (lane < gd && r <= lane) ? M[lane * gd + r] : 0.0;
    bool ok = true;
#pragma unroll
    for (int j = 0; j < PCW; j++) {
        if (j < gd && ok) {   // uniform
            const double piv = __shfl(col[j], j, 64);
            if (!(piv > 0.0) || !isfinite(piv)) {
                ok = false;
            } else {
                const double ljj = sqrt(piv);
                const double lij = lane == j ? ljj : col[j] / ljj;   // L(lane, j) for lanes > j
                col[j] = lij;
#pragma unroll
                for (int r = j + 1; r < PCW; r++)
                    if (r < gd) col[r] -= __shfl(lij, r, 64) * lij;   // entry (r, lane): L(r, j) L(lane, j)
            }
        }
    }
    if (!ok) {
        if (lane == 0) a.st[a.prob_of_block[b]].failed = 1;
        return;
    }
    double *L = Lout + a.gloff[b];
#pragma unroll
    for (int j = 0; j < PCW; j++)
        if (lane < gd && j <= lane) L[j * gd + lane] = col[j];
}

// The group correction of block-Jacobi (the learned correction's family; no reference counterpart): the group of
// consecutive blocks b .. b + gs - 1 that block b starts gets the block matrix with S's diagonal blocks on its
// diagonal and w S_ij (w = omega[b] in [0, 1]; zero where the pattern has no block (i, j)) off it: w = 0 is plain
// block-Jacobi, w = 1 the group's exact principal block submatrix of S. For w in [0, 1] it is a convex combination
// of the two SPD matrices, so it stays SPD.
__global__ void k_pcg_group_blocks(PcgArgs a, const double *S, const int64_t *map, const int64_t *diag_voff,
                                   const float *omega, double *M) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= a.nrb) return;
    const int gd = group_dim(a, b);
    if (!gd) return;
    const int gs = a.gsize[b];
    double *out = M + a.gloff[b];
    for (int q = 0; q < gd * gd; q++) out[q] = 0.0;
    const double w = fmin(1.0, fmax(0.0, (double)omega[b]));
    int ri = 0;   // first row of block i in the group
    for (int i = 0; i < gs; i++) {
        const int64_t bi = b + i;
        const int di = a.dims[bi];
        int cj = 0;
        for (int j = 0; j < gs; j++) {
            const int64_t bj = b + j;
            const int dj = a.dims[bj];
            if (i == j) {
                for (int c = 0; c < di; c++)
                    for (int r = 0; r < di; r++) out[(ri + c) * gd + ri + r] = S[map[diag_voff[bi] + r * di + c] - 1];
            } else if (w != 0.0) {
                int64_t voff = -1;
                for (int k = a.bsr_rowptr[bi]; k < a.bsr_rowptr[bi + 1]; k++)
                    if (a.bsr_col[k] == bj) voff = a.bsr_voff[k];
                if (voff >= 0)
                    for (int r = 0; r < di; r++)
                        for (int c = 0; c < dj; c++) out[(cj + c) * gd + ri + r] = w * S[map[voff + r * dj + c] - 1];
            }
            cj += dj;
        }
        ri += di;
    }
}

// The same, one wavefront per group: lanes 0 .. gs^2 - 1 find the pattern's blocks (i, j) of the group in their
// block rows, then every lane writes entries of the gd x gd block (the values as k_pcg_group_blocks writes them).
__global__ __launch_bounds__(64) void k_pcg_group_blocks_w(PcgArgs a, const double *S, const int64_t *map,
                                                          const int64_t *diag_voff, const float *omega, double *M) {
    __shared__ int64_t voff[PC_GMAX * PC_GMAX];
    __shared__ int boff[PC_GMAX + 1];
    const int64_t b = blockIdx.x;
    if (b >= a.nrb) return;
    const int gd = group_dim(a, b);
    if (!gd) return;
    const int gs = a.gsize[b];
    const int lane = threadIdx.x;
    const double w = fmin(1.0, fmax(0.0, (double)omega[b]));
    if (lane < gs * gs) {
        const int i = lane / gs, j = lane % gs;
        int64_t v = -1;
        if (i == j) {
            v = diag_voff[b + i];
        } else if (w != 0.0) {
            for (int k = a.bsr_rowptr[b + i]; k < a.bsr_rowptr[b + i + 1]; k++)
                if (a.bsr_col[k] == b + j) v = a.bsr_voff[k];
        }
        voff[lane] = v;
    }
    if (lane == 0) {
        int o = 0;
        for (int i = 0; i < gs; i++) { boff[i] = o; o += a.dims[b + i]; }
        boff[gs] = o;
    }
    __syncthreads();
    double *out = M + a.gloff[b];
    for (int q = lane; q < gd * gd; q += 64) {
        const int row = q % gd, cl = q / gd;
        int i = 0, j = 0;
        while (row >= boff[i + 1]) i++;
        while (cl >= boff[j + 1]) j++;
        const int r = row - boff[i], c = cl - boff[j];
        const int64_t v = voff[i * gs + j];
        double x = 0.0;
        if (i == j) x = S[map[v + r * a.dims[b + i] + c] - 1];
        else if (v >= 0) x = w * S[map[v + r * a.dims[b + j] + c] - 1];
        out[q] = x;
    }
}

// The learned correction of the block-Jacobi preconditioner (no reference counterpart): per retained block,
// features for a small MLP, and the corrected blocks M_b = D_b + mu_b m_b I (D_b the diagonal block of S, m_b the
// mean of its diagonal), read from S through the handle's BSR map.
// ---------------------------------------------------------------------------------------------------------
constexpr int PC_NF = 26;

// Features of block row i (dimension 6; other dimensions get zeros): the 21 upper entries of D / m, log10 m,
// log(1 + number of off-diagonal blocks), log10 lambda, log10(m / lambda), log10(off-diagonal Frobenius mass of the
// block row / ||D||_F). One wavefront per block row; the off-diagonal sum is a fixed-order wave reduction.
__global__ __launch_bounds__(256) void k_pcg_features(PcgArgs a, const double *vals, const int64_t *diag_voff,
                                                       double lambda, float *F) {
    const int64_t i = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= a.nrb) return;
    const int di = a.dims[i];
    // the block row's values lie one after another in the gathered block-row copy: the off-diagonal ones, the lanes
    // over the run (coalesced), each lane adding its values in order
    const int kl = a.bsr_rowptr[i + 1] - 1;
    const int64_t v0 = a.bsr_voff[a.bsr_rowptr[i]], v1 = a.bsr_voff[kl] + (int64_t)di * a.dims[a.bsr_col[kl]];
    const int64_t d0 = diag_voff[i], d1 = d0 + (int64_t)di * di;
    double off = 0.0;
#pragma unroll 4
    for (int64_t v = v0 + lane; v < v1; v += 64) {
        const double x = vals[v];
        off += (v >= d0 && v < d1) ? 0.0 : x * x;
    }
    off = wave_sum64(off);
    if (lane != 0) return;
    float *f = F + i * PC_NF;
    for (int q = 0; q < PC_NF; q++) f[q] = 0.0f;
    if (di != 6) return;
    double D[36], m = 0.0, dd = 0.0;
    for (int q = 0; q < 36; q++) {
        D[q] = vals[d0 + q];   // row-major
        dd += D[q] * D[q];
    }
    for (int r = 0; r < 6; r++) m += D[r * 6 + r];
    m /= 6.0;
    const double mm = m > 0 ? m : 1e-300;
    int q = 0;
    for (int c = 0; c < 6; c++)
        for (int r = 0; r <= c; r++) f[q++] = (float)(D[r * 6 + c] / mm);
    f[21] = (float)log10(mm);
    f[22] = (float)log(1.0 + (double)(a.bsr_rowptr[i + 1] - a.bsr_rowptr[i] - 1));
    f[23] = (float)log10(lambda > 0 ? lambda : 1e-300);
    f[24] = (float)log10(mm / (lambda > 0 ? lambda : 1e-300));
    f[25] = (float)log10((sqrt(off) + 1e-300) / (sqrt(dd) + 1e-300));
}

// M_b = D_b + mu_b m_b I, full d x d column-major at the preconditioner offsets (the precond_blocks of the solve)
__global__ void k_pcg_shift_blocks(PcgArgs a, const double *S, const int64_t *map, const int64_t *diag_voff,
                                   const float *mu, double *M) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= a.nrb) return;
    const int d = a.dims[b];
    double m = 0.0;
    for (int r = 0; r < d; r++) m += S[map[diag_voff[b] + r * d + r] - 1];
    m /= d;
    const double shift = (double)mu[b] * m;
    double *out = M + a.loff[b];
    for (int c = 0; c < d; c++)
        for (int r = 0; r < d; r++) out[c * d + r] = S[map[diag_voff[b] + r * d + c] - 1] + (r == c ? shift : 0.0);
}

__global__ void k_gather_f64(double *dest, const double *src, const int64_t *map, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dest[i] = src[map[i] - 1];
}

__global__ void k_copy_f64(double *dst, const double *src, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// ||g - q|| / ||g|| per problem (q = S x), one workgroup per problem, fixed-order sums
__global__ __launch_bounds__(UPD_THREADS) void k_pcg_relres(PcgArgs a, const double *g, double *out) {
    __shared__ double red[UPD_THREADS];
    const int pb = blockIdx.x;
    const int64_t b0 = a.prob_bptr[pb], b1 = a.prob_bptr[pb + 1];
    const int64_t s0 = a.bo[b0], s1 = a.bo[b1 - 1] + a.dims[b1 - 1];
    double gg = 0.0, rr = 0.0;
    for (int64_t s = s0 + threadIdx.x; s < s1; s += UPD_THREADS) {
        const double gv = g[s], rv = gv - a.q[s];
        gg += gv * gv;
        rr += rv * rv;
    }
    gg = wg_sum(gg, red);
    rr = wg_sum(rr, red);
    if (threadIdx.x == 0) out[pb] = gg > 0 ? sqrt(rr) / sqrt(gg) : sqrt(rr);
}

// r = t - S x (q holds S x) and, per problem, ||r|| / ||gref|| (gref: the right-hand side the residual is judged
// against); one workgroup per problem, fixed-order sums.
__global__ __launch_bounds__(UPD_THREADS) void k_pcg_residual(PcgArgs a, const double *t, const double *gref,
                                                              double *r, double *out) {
    __shared__ double red[UPD_THREADS];
    const int pb = blockIdx.x;
    const int64_t b0 = a.prob_bptr[pb], b1 = a.prob_bptr[pb + 1];
    const int64_t s0 = a.bo[b0], s1 = a.bo[b1 - 1] + a.dims[b1 - 1];
    double gg = 0.0, rr = 0.0;
    for (int64_t s = s0 + threadIdx.x; s < s1; s += UPD_THREADS) {
        const double rv = t[s] - a.q[s], gv = gref[s];
        r[s] = rv;
        gg += gv * gv;
        rr += rv * rv;
    }
    gg = wg_sum(gg, red);
    rr = wg_sum(rr, red);
    if (threadIdx.x == 0) out[pb] = gg > 0 ? sqrt(rr) / sqrt(gg) : sqrt(rr);
}

__global__ void k_pcg_init_state(ProbState *st, int nprob, int *nactive) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) *nactive = 0;
    if (i < nprob) {
        st[i].rz2[0] = 0; st[i].rz2[1] = 0; st[i].gnorm = 0; st[i].relres = 0; st[i].iters = 0; st[i].active = 0; st[i].failed = 0;
    }
}

}  // namespace

struct bls_pcg {
    int64_t nr = 0, nrb = 0, nprob = 0, nvals = 0, nbsr = 0, nL = 0;
    std::vector<int64_t> prob_bptr;
    DevArray<int32_t> d_row_block, d_rowptr, d_col, d_dims, d_prob_of_block;
    DevArray<int64_t> d_voff, d_bo, d_prob_bptr, d_map, d_loff, d_diag_voff;
    DevArray<double> d_vals, d_L, d_x, d_r, d_z, d_p, d_q, d_g, d_pq, d_rzp, d_rrp;
    DevArray<int32_t> d_grp_prob;
    DevArray<int64_t> d_grp_b0, d_grp_b1;
    int64_t ngroups = 0;
    int maxd = 1;
    int uniform_d = 0;   // the block dimension when every block has it, else 0
    DevArray<ProbState> d_st;
    DevArray<int> d_nactive;
    hipStream_t stream = nullptr;
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    double graph_tol = -1;
    int64_t graph_maxit = -1;
    // preconditioner groups of `group` consecutive blocks (bls_pcg_set_groups) and their iteration graph
    DevArray<int32_t> d_gsize;
    DevArray<int64_t> d_gloff;
    DevArray<double> d_Lp;
    int64_t nLp = 0;
    int group = 0;
    hipGraph_t graph_p = nullptr;
    hipGraphExec_t exec_p = nullptr;
    double graph_p_tol = -1;
    int64_t graph_p_maxit = -1;
    ~bls_pcg() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (exec_p) (void)hipGraphExecDestroy(exec_p);
        if (graph_p) (void)hipGraphDestroy(graph_p);
        if (ev_in) (void)hipEventDestroy(ev_in);
        if (ev_out) (void)hipEventDestroy(ev_out);
        if (stream) (void)hipStreamDestroy(stream);
    }
    PcgArgs args(double tol, int64_t maxit, bool groups = false) {
        PcgArgs a{};
        if (groups) {
            a.gsize = d_gsize.ptr;
            a.gloff = d_gloff.ptr;
        }
        a.nr = nr;
        a.row_block = d_row_block.ptr;
        a.bsr_rowptr = d_rowptr.ptr;
        a.bsr_col = d_col.ptr;
        a.bsr_voff = d_voff.ptr;
        a.dims = d_dims.ptr;
        a.bo = d_bo.ptr;
        a.prob_of_block = d_prob_of_block.ptr;
        a.prob_bptr = d_prob_bptr.ptr;
        a.vals = d_vals.ptr;
        a.L = groups ? d_Lp.ptr : d_L.ptr;
        a.loff = d_loff.ptr;
        a.x = d_x.ptr; a.r = d_r.ptr; a.z = d_z.ptr; a.p = d_p.ptr; a.q = d_q.ptr; a.g = d_g.ptr;
        a.pq_part = d_pq.ptr; a.rz_part = d_rzp.ptr; a.rr_part = d_rrp.ptr;
        a.grp_prob = d_grp_prob.ptr; a.grp_b0 = d_grp_b0.ptr; a.grp_b1 = d_grp_b1.ptr;
        a.nrb = nrb;
        a.st = d_st.ptr;
        a.nactive = d_nactive.ptr;
        a.tol = tol;
        a.maxit = maxit;
        return a;
    }
};

// S x for the handle's blocks: the row-segment kernel when every block has one dimension 6 or 9, else the generic one
static void launch_spmv(const bls_pcg *h, const PcgArgs &a, const double *xin, double *yout, double *pq,
                        int only_active, hipStream_t s) {
    const unsigned grid = (unsigned)((h->nrb + 3) / 4);
    if (h->uniform_d == 6) k_bsr_spmv_u<6><<<grid, 256, 0, s>>>(a, xin, yout, pq, only_active);
    else if (h->uniform_d == 9) k_bsr_spmv_u<9><<<grid, 256, 0, s>>>(a, xin, yout, pq, only_active);
    else if (h->maxd <= 6) k_bsr_spmv_v<6><<<grid, 256, 0, s>>>(a, xin, yout, pq, only_active);
    else if (h->maxd <= 9) k_bsr_spmv_v<9><<<grid, 256, 0, s>>>(a, xin, yout, pq, only_active);
    else k_bsr_spmv_v<16><<<grid, 256, 0, s>>>(a, xin, yout, pq, only_active);
}

extern "C" int bls_pcg_create(int64_t nr, int64_t nrblocks, const int64_t *r_dims, const int64_t *boffsets_r,
                              int64_t nnz, const int64_t *S_colptr, int64_t nnz_blocks, const int64_t *block_colptr,
                              const int64_t *block_rowval, const int64_t *block_row_offsets,
                              const int64_t *col_heights, int64_t nprob, const int64_t *prob_block_ptr,
                              bls_pcg **out) {
    if (!out || nr < 0 || nrblocks < 0 || nprob < 1) return fail(BLS_ERR_ARG, "bls_pcg_create: bad arguments");
    *out = nullptr;
    for (int64_t r = 0; r < nrblocks; r++)
        if (r_dims[r] < 1 || r_dims[r] > MAXD) return fail(BLS_ERR_UNSUPPORTED, "retained block dimension outside 1..16");
    if (prob_block_ptr[0] != 1 || prob_block_ptr[nprob] != nrblocks + 1)
        return fail(BLS_ERR_ARG, "prob_block_ptr must run from 1 to nrblocks + 1");
    for (int64_t p = 0; p < nprob; p++)
        if (prob_block_ptr[p + 1] <= prob_block_ptr[p]) return fail(BLS_ERR_ARG, "every problem needs a retained block");
    auto *h = new bls_pcg();
    h->nr = nr;
    h->nrb = nrblocks;
    h->nprob = nprob;
    // full symmetric block-row lists: (col, slot, transposed)
    struct Ent { int64_t col, slot; bool tr; };
    std::vector<std::vector<Ent>> rows(nrblocks);
    for (int64_t j = 0; j < nrblocks; j++)
        for (int64_t s = block_colptr[j] - 1; s < block_colptr[j + 1] - 1; s++) {
            const int64_t i = block_rowval[s] - 1;
            rows[i].push_back({j, s, false});
            if (i < j) rows[j].push_back({i, s, true});
        }
    std::vector<int32_t> rowptr{0}, col, dims(nrblocks), row_block(nr), prob_of_block(nrblocks);
    std::vector<int64_t> voff, map, bo(nrblocks), loff(nrblocks), diag_voff(nrblocks, -1), pbp(nprob + 1);
    for (int64_t r = 0; r < nrblocks; r++) {
        dims[r] = (int32_t)r_dims[r];
        bo[r] = boffsets_r[r] - 1;
        for (int64_t a = 0; a < r_dims[r]; a++) row_block[bo[r] + a] = (int32_t)r;
    }
    for (int64_t p = 0; p <= nprob; p++) pbp[p] = prob_block_ptr[p] - 1;
    for (int64_t p = 0; p < nprob; p++)
        for (int64_t b = pbp[p]; b < pbp[p + 1]; b++) prob_of_block[b] = (int32_t)p;
    int64_t nv = 0, nl = 0;
    for (int64_t i = 0; i < nrblocks; i++) {
        std::sort(rows[i].begin(), rows[i].end(), [](const Ent &x, const Ent &y) { return x.col < y.col; });
        const int64_t di = r_dims[i];
        for (const Ent &en : rows[i]) {
            const int64_t j = en.col, dj = r_dims[j];
            if (prob_of_block[j] != prob_of_block[i]) { delete h; return fail(BLS_ERR_ARG, "a block couples two problems"); }
            col.push_back((int32_t)j);
            voff.push_back(nv);
            if (j == i) diag_voff[i] = nv;
            for (int64_t a = 0; a < di; a++)
                for (int64_t c = 0; c < dj; c++) {
                    int64_t idx;   // 1-based index into S.nzval of S[(i,a), (j,c)]
                    const int64_t ro = block_row_offsets[en.slot];
                    if (!en.tr) {   // stored block (i, j), i <= j, in column j
                        if (i == j && a > c) idx = S_colptr[bo[i] + a] + ro + c;
                        else idx = S_colptr[bo[j] + c] + ro + a;
                    } else {        // i > j: stored (j, i) in column i; element (c, a)
                        idx = S_colptr[bo[i] + a] + ro + c;
                    }
                    map.push_back(idx);
                    nv++;
                }
        }
        rowptr.push_back((int32_t)col.size());
        loff[i] = nl;
        nl += di * di;
        if (diag_voff[i] < 0) { delete h; return fail(BLS_ERR_ARG, "the pattern has no diagonal block " + std::to_string(i + 1)); }
    }
    (void)nnz; (void)nnz_blocks; (void)col_heights;
    // update groups: up to GRP_BLOCKS consecutive blocks of one problem each
    std::vector<int32_t> grp_prob;
    std::vector<int64_t> grp_b0, grp_b1;
    for (int64_t p = 0; p < nprob; p++)
        for (int64_t b = pbp[p]; b < pbp[p + 1]; b += GRP_BLOCKS) {
            grp_prob.push_back((int32_t)p);
            grp_b0.push_back(b);
            grp_b1.push_back(std::min<int64_t>(b + GRP_BLOCKS, pbp[p + 1]));
        }
    h->ngroups = (int64_t)grp_prob.size();
    for (int64_t r = 0; r < nrblocks; r++) h->maxd = std::max(h->maxd, (int)r_dims[r]);
    h->uniform_d = nrblocks ? (int)r_dims[0] : 0;
    for (int64_t r = 0; r < nrblocks; r++)
        if (r_dims[r] != h->uniform_d) h->uniform_d = 0;
    h->nvals = nv;
    h->nbsr = (int64_t)col.size();
    h->nL = nl;
    h->prob_bptr = pbp;
    int rc;
    if ((rc = h->d_row_block.upload(row_block)) || (rc = h->d_rowptr.upload(rowptr)) || (rc = h->d_col.upload(col)) ||
        (rc = h->d_dims.upload(dims)) || (rc = h->d_prob_of_block.upload(prob_of_block)) || (rc = h->d_voff.upload(voff)) ||
        (rc = h->d_bo.upload(bo)) || (rc = h->d_prob_bptr.upload(pbp)) || (rc = h->d_map.upload(map)) ||
        (rc = h->d_loff.upload(loff)) || (rc = h->d_diag_voff.upload(diag_voff)) || (rc = h->d_vals.alloc(nv)) ||
        (rc = h->d_L.alloc(nl)) || (rc = h->d_x.alloc(nr)) || (rc = h->d_r.alloc(nr)) || (rc = h->d_z.alloc(nr)) ||
        (rc = h->d_p.alloc(nr)) || (rc = h->d_q.alloc(nr)) || (rc = h->d_g.alloc(nr)) || (rc = h->d_st.alloc(nprob)) ||
        (rc = h->d_nactive.alloc(1)) || (rc = h->d_pq.alloc(nrblocks)) || (rc = h->d_rzp.alloc(nrblocks)) ||
        (rc = h->d_rrp.alloc(nrblocks)) || (rc = h->d_grp_prob.upload(grp_prob)) || (rc = h->d_grp_b0.upload(grp_b0)) ||
        (rc = h->d_grp_b1.upload(grp_b1))) {
        delete h;
        return rc;
    }
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&h->ev_in, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&h->ev_out, hipEventDisableTiming) != hipSuccess) {
        delete h;
        return fail(BLS_ERR_HIP, "stream / event creation failed");
    }
    *out = h;
    return BLS_OK;
}

extern "C" void bls_pcg_destroy(bls_pcg *h) { delete h; }

extern "C" int bls_pcg_features_f32(bls_pcg *h, const double *S_nzval, double lambda, float *features,
                                    void *stream) {
    if (!h || !S_nzval || !features) return fail(BLS_ERR_ARG, "bls_pcg_features_f32: bad arguments");
    PcgArgs a = h->args(1.0, 0);
    hipStream_t s = (hipStream_t)stream;
    // S gathered into the block-row copy (as a solve starts), then read row by row
    k_gather_f64<<<std::min<int64_t>(grid_for(h->nvals, 256), 16384), 256, 0, s>>>(h->d_vals.ptr, S_nzval,
                                                                                 h->d_map.ptr, h->nvals);
    k_pcg_features<<<(unsigned)((h->nrb + 3) / 4), 256, 0, s>>>(a, h->d_vals.ptr, h->d_diag_voff.ptr, lambda,
                                                               features);
    BLS_HIP(hipGetLastError());
    return BLS_OK;
}

extern "C" int bls_pcg_shift_blocks_f64(bls_pcg *h, const double *S_nzval, const float *mu, double *blocks,
                                        void *stream) {
    if (!h || !S_nzval || !mu || !blocks) return fail(BLS_ERR_ARG, "bls_pcg_shift_blocks_f64: bad arguments");
    PcgArgs a = h->args(1.0, 0);
    k_pcg_shift_blocks<<<grid_for(h->nrb, 256), 256, 0, (hipStream_t)stream>>>(a, S_nzval, h->d_map.ptr,
                                                                            h->d_diag_voff.ptr, mu, blocks);
    BLS_HIP(hipGetLastError());
    return BLS_OK;
}

extern "C" int bls_pcg_values_size(bls_pcg *h, int64_t *nvals, int64_t *nfactor) {
    if (!h) return fail(BLS_ERR_ARG, "null handle");
    if (nvals) *nvals = h->nvals;
    if (nfactor) *nfactor = h->nL;
    return BLS_OK;
}

// Preconditioner groups of every problem: blocks (b0 + G j, ..., b0 + G j + G - 1) from the problem's first block b0
// (the last group may be shorter), and factor offsets for the group blocks.
static int build_groups(bls_pcg *h, int group) {
    if (group < 1 || group > PC_GMAX || GRP_BLOCKS % group != 0)
        return fail(BLS_ERR_ARG, "preconditioner groups of 1 to 4 blocks");
    if (h->group == group) return BLS_OK;
    std::vector<int32_t> gsize(h->nrb, 0), dims(h->nrb);
    std::vector<int64_t> gloff(h->nrb, 0);
    BLS_HIP(hipMemcpy(dims.data(), h->d_dims.ptr, sizeof(int32_t) * h->nrb, hipMemcpyDeviceToHost));
    int64_t nl = 0;
    for (int64_t p = 0; p < h->nprob; p++)
        for (int64_t b = h->prob_bptr[p]; b < h->prob_bptr[p + 1]; b += group) {
            const int64_t e = std::min<int64_t>(b + group, h->prob_bptr[p + 1]);
            gsize[b] = (int32_t)(e - b);
            gloff[b] = nl;
            int64_t gd = 0;
            for (int64_t k = b; k < e; k++) gd += dims[k];
            nl += gd * gd;
        }
    int rc;
    if ((rc = h->d_gsize.upload(gsize)) || (rc = h->d_gloff.upload(gloff)) || (rc = h->d_Lp.alloc((size_t)std::max<int64_t>(1, nl))))
        return rc;
    h->nLp = nl;
    h->group = group;
    if (h->exec_p) { (void)hipGraphExecDestroy(h->exec_p); h->exec_p = nullptr; }
    if (h->graph_p) { (void)hipGraphDestroy(h->graph_p); h->graph_p = nullptr; }
    return BLS_OK;
}

static int pcg_solve(bls_pcg *h, const double *S_nzval, const double *g, double *x, double tol, int64_t maxit,
                     const double *precond_blocks, int64_t *iters, double *relres, bool groups, void *stream) {
    if (!h || !S_nzval || !g || !x) return fail(BLS_ERR_ARG, "bls_pcg_solve_f64: bad arguments");
    if (!(tol > 0) || maxit < 0) return fail(BLS_ERR_ARG, "tol must be positive and maxit non-negative");
    if (groups && (!precond_blocks || h->group == 0))
        return fail(BLS_ERR_STATE, "the group-corrected solve needs bls_pcg_set_groups and its group blocks");
    hipStream_t cs = (hipStream_t)stream;
    hipStream_t s = h->stream;
    BLS_HIP(hipEventRecord(h->ev_in, cs));
    BLS_HIP(hipStreamWaitEvent(s, h->ev_in, 0));
    PcgArgs a = h->args(tol, maxit, groups);
    const int nprob = (int)h->nprob;
    auto spmv = [&](const double *xin, double *yout, double *pq, int only_active) {
        launch_spmv(h, a, xin, yout, pq, only_active, s);
    };
    const unsigned ngroups = (unsigned)h->ngroups;
    // the register-resident block solve, unrolled to the largest block, when there are no groups
    const int ud = groups ? 0 : h->maxd <= 3 ? 3 : h->maxd <= 6 ? 6 : h->maxd <= 9 ? 9 : h->maxd <= 16 ? 16 : 0;
    auto restart = [&](int first) {
        if (ud == 3) k_pcg_restart_grp<3><<<ngroups, GRP_BLOCKS, 0, s>>>(a, first);
        else if (ud == 6) k_pcg_restart_grp<6><<<ngroups, GRP_BLOCKS, 0, s>>>(a, first);
        else if (ud == 9) k_pcg_restart_grp<9><<<ngroups, GRP_BLOCKS, 0, s>>>(a, first);
        else if (ud == 16) k_pcg_restart_grp<16><<<ngroups, GRP_BLOCKS, 0, s>>>(a, first);
        else k_pcg_restart_grp<0><<<ngroups, GRP_BLOCKS, 0, s>>>(a, first);
    };
    auto step1 = [&](int parity) {
        if (ud == 3) k_pcg_step1<3><<<ngroups, GRP_BLOCKS, 0, s>>>(a, parity);
        else if (ud == 6) k_pcg_step1<6><<<ngroups, GRP_BLOCKS, 0, s>>>(a, parity);
        else if (ud == 9) k_pcg_step1<9><<<ngroups, GRP_BLOCKS, 0, s>>>(a, parity);
        else if (ud == 16) k_pcg_step1<16><<<ngroups, GRP_BLOCKS, 0, s>>>(a, parity);
        else k_pcg_step1<0><<<ngroups, GRP_BLOCKS, 0, s>>>(a, parity);
    };
    k_gather_f64<<<std::min<int64_t>(grid_for(h->nvals, 256), 16384), 256, 0, s>>>(h->d_vals.ptr, S_nzval, h->d_map.ptr, h->nvals);
    k_copy_f64<<<std::min<int64_t>(grid_for(h->nr, 256), 4096), 256, 0, s>>>(h->d_g.ptr, g, h->nr);
    k_pcg_init_state<<<grid_for(nprob, 256), 256, 0, s>>>(h->d_st.ptr, nprob, h->d_nactive.ptr);
    if (groups)
        k_pcg_factor_groups_w<<<(unsigned)h->nrb, 64, 0, s>>>(a, h->nrb, h->d_Lp.ptr, precond_blocks);
    else
        k_pcg_factor<<<grid_for(h->nrb, 128), 128, 0, s>>>(a, h->nrb, h->d_diag_voff.ptr, h->d_L.ptr, precond_blocks);
    restart(1);
    k_pcg_restart_fin<<<nprob, GRP_BLOCKS, 0, s>>>(a, 1);
    BLS_HIP(hipGetLastError());
    // the iteration graph (handle-owned buffers only; rebuilt when tol / maxit change; one per preconditioner kind)
    hipGraph_t &graph = groups ? h->graph_p : h->graph;
    hipGraphExec_t &exec = groups ? h->exec_p : h->exec;
    double &gtol = groups ? h->graph_p_tol : h->graph_tol;
    int64_t &gmaxit = groups ? h->graph_p_maxit : h->graph_maxit;
    if (!exec || gtol != tol || gmaxit != maxit) {
        if (exec) { (void)hipGraphExecDestroy(exec); exec = nullptr; }
        if (graph) { (void)hipGraphDestroy(graph); graph = nullptr; }
        BLS_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        for (int it = 0; it < ITERS_PER_GRAPH; it++) {
            spmv(h->d_p.ptr, h->d_q.ptr, h->d_pq.ptr, 1);
            step1(it & 1);
            k_pcg_step2<<<ngroups, GRP_BLOCKS, 0, s>>>(a, it & 1);
        }
        hipError_t e = hipStreamEndCapture(s, &graph);
        if (e != hipSuccess) return fail(BLS_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        BLS_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        gtol = tol;
        gmaxit = maxit;
    }
    // iterate; then confirm every problem with its true residual, restarting those still above tol
    int nact = 0;
    for (int round = 0; round < 4; round++) {
        BLS_HIP(hipMemcpyAsync(&nact, h->d_nactive.ptr, sizeof(int), hipMemcpyDeviceToHost, s));
        BLS_HIP(hipStreamSynchronize(s));
        int64_t launched = 0;
        while (nact > 0 && launched <= maxit + ITERS_PER_GRAPH) {
            BLS_HIP(hipGraphLaunch(exec, s));
            launched += ITERS_PER_GRAPH;
            BLS_HIP(hipMemcpyAsync(&nact, h->d_nactive.ptr, sizeof(int), hipMemcpyDeviceToHost, s));
            BLS_HIP(hipStreamSynchronize(s));
        }
        spmv(h->d_x.ptr, h->d_q.ptr, nullptr, 0);
        restart(0);
        k_pcg_restart_fin<<<nprob, GRP_BLOCKS, 0, s>>>(a, 0);
        BLS_HIP(hipGetLastError());
        BLS_HIP(hipMemcpyAsync(&nact, h->d_nactive.ptr, sizeof(int), hipMemcpyDeviceToHost, s));
        BLS_HIP(hipStreamSynchronize(s));
        if (nact == 0) break;
    }
    std::vector<ProbState> stv(nprob);
    BLS_HIP(hipMemcpyAsync(stv.data(), h->d_st.ptr, sizeof(ProbState) * nprob, hipMemcpyDeviceToHost, s));
    BLS_HIP(hipMemcpyAsync(x, h->d_x.ptr, sizeof(double) * h->nr, hipMemcpyDeviceToDevice, s));
    BLS_HIP(hipEventRecord(h->ev_out, s));
    BLS_HIP(hipStreamWaitEvent(cs, h->ev_out, 0));
    BLS_HIP(hipStreamSynchronize(s));
    bool all = true;
    for (int p = 0; p < nprob; p++) {
        if (relres) relres[p] = stv[p].relres;
        if (iters) iters[p] = stv[p].failed ? -1 : stv[p].iters;
        if (stv[p].failed || !(stv[p].relres <= tol)) all = false;
    }
    return all ? BLS_OK : BLS_FAILED;
}

extern "C" int bls_pcg_solve_f64(bls_pcg *h, const double *S_nzval, const double *g, double *x, double tol,
                                 int64_t maxit, const double *precond_blocks, int64_t *iters, double *relres,
                                 void *stream) {
    return pcg_solve(h, S_nzval, g, x, tol, maxit, precond_blocks, iters, relres, false, stream);
}

extern "C" int bls_pcg_set_groups(bls_pcg *h, int group) {
    if (!h) return fail(BLS_ERR_ARG, "null handle");
    return build_groups(h, group);
}

extern "C" int bls_pcg_solve_groups_f64(bls_pcg *h, const double *S_nzval, const double *g, double *x, double tol,
                                        int64_t maxit, const double *group_blocks, int64_t *iters, double *relres,
                                        void *stream) {
    return pcg_solve(h, S_nzval, g, x, tol, maxit, group_blocks, iters, relres, true, stream);
}

extern "C" int bls_pcg_group_blocks_size(bls_pcg *h, int64_t *n) {
    if (!h || !n) return fail(BLS_ERR_ARG, "bls_pcg_group_blocks_size: bad arguments");
    if (h->group == 0) return fail(BLS_ERR_STATE, "no preconditioner groups (bls_pcg_set_groups)");
    *n = h->nLp;
    return BLS_OK;
}

extern "C" int bls_pcg_group_blocks_f64(bls_pcg *h, const double *S_nzval, const float *omega, double *blocks,
                                        void *stream) {
    if (!h || !S_nzval || !omega || !blocks) return fail(BLS_ERR_ARG, "bls_pcg_group_blocks_f64: bad arguments");
    if (h->group == 0) return fail(BLS_ERR_STATE, "no preconditioner groups (bls_pcg_set_groups)");
    PcgArgs a = h->args(1.0, 0, true);
    k_pcg_group_blocks_w<<<(unsigned)h->nrb, 64, 0, (hipStream_t)stream>>>(a, S_nzval, h->d_map.ptr,
                                                                        h->d_diag_voff.ptr, omega, blocks);
    BLS_HIP(hipGetLastError());
    return BLS_OK;
}

__global__ void k_axpy_f64(int64_t n, double alpha, const double *x, double *y) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] += alpha * x[i];
}

extern "C" int bls_axpy_f64(int64_t n, double alpha, const double *x, double *y, void *stream) {
    if (n < 0 || (n && (!x || !y))) return fail(BLS_ERR_ARG, "bls_axpy_f64: bad arguments");
    if (n) k_axpy_f64<<<std::min<int64_t>(grid_for(n, 256), 4096), 256, 0, (hipStream_t)stream>>>(n, alpha, x, y);
    BLS_HIP(hipGetLastError());
    return BLS_OK;
}

extern "C" int bls_pcg_residual_f64(bls_pcg *h, const double *S_nzval, const double *t, const double *x,
                                    const double *gref, double *r, double *relres, void *stream) {
    if (!h || !S_nzval || !t || !x || !gref || !r || !relres) return fail(BLS_ERR_ARG, "bls_pcg_residual_f64: bad arguments");
    hipStream_t s = (hipStream_t)stream;
    PcgArgs a = h->args(1.0, 0);
    k_gather_f64<<<std::min<int64_t>(grid_for(h->nvals, 256), 16384), 256, 0, s>>>(h->d_vals.ptr, S_nzval, h->d_map.ptr, h->nvals);
    k_copy_f64<<<std::min<int64_t>(grid_for(h->nr, 256), 4096), 256, 0, s>>>(h->d_z.ptr, x, h->nr);
    launch_spmv(h, a, h->d_z.ptr, h->d_q.ptr, nullptr, 0, s);
    k_pcg_residual<<<(unsigned)h->nprob, UPD_THREADS, 0, s>>>(a, t, gref, r, h->d_pq.ptr);
    BLS_HIP(hipGetLastError());
    BLS_HIP(hipMemcpyAsync(relres, h->d_pq.ptr, sizeof(double) * h->nprob, hipMemcpyDeviceToHost, s));
    BLS_HIP(hipStreamSynchronize(s));
    return BLS_OK;
}

extern "C" int bls_pcg_relres_f64(bls_pcg *h, const double *S_nzval, const double *g, const double *x,
                                  double *relres, void *stream) {
    if (!h || !S_nzval || !g || !x || !relres) return fail(BLS_ERR_ARG, "bls_pcg_relres_f64: bad arguments");
    hipStream_t s = (hipStream_t)stream;
    PcgArgs a = h->args(1.0, 0);
    k_gather_f64<<<std::min<int64_t>(grid_for(h->nvals, 256), 16384), 256, 0, s>>>(h->d_vals.ptr, S_nzval, h->d_map.ptr, h->nvals);
    // the row-segment product reads x with 16-byte loads: copy the caller's x into an aligned handle buffer
    k_copy_f64<<<std::min<int64_t>(grid_for(h->nr, 256), 4096), 256, 0, s>>>(h->d_z.ptr, x, h->nr);
    launch_spmv(h, a, h->d_z.ptr, h->d_q.ptr, nullptr, 0, s);
    k_pcg_relres<<<(unsigned)h->nprob, UPD_THREADS, 0, s>>>(a, g, h->d_pq.ptr);
    BLS_HIP(hipGetLastError());
    BLS_HIP(hipMemcpyAsync(relres, h->d_pq.ptr, sizeof(double) * h->nprob, hipMemcpyDeviceToHost, s));
    BLS_HIP(hipStreamSynchronize(s));
    return BLS_OK;
}
