// The dense reduced solve on gfx950: S x = g for the dense reduced system S (nr x nr, its upper triangle stored
// column-major, as reduce_schur_dense! writes it), by the Cholesky factorization S = U'U of a copy (S is left as
// built), behind the C ABI of include/bls_amd.h. This is the :cholesky branch of solve_reduced_dense!
// (src/schur_linear_systems.jl:312-319), whose factorization the reference takes from DampedLinearSolve (LAPACK's
// potrf); a matrix that is not positive definite returns BLS_FAILED with the first column whose pivot is not
// positive (NotPositiveDefinite), the caller's cue to fall back to an indefinite factorization.
//
// Blocked right-looking factorization, panels of NB = 64 rows of U:
//   k_chol_diag   one workgroup factors the diagonal block in LDS (unblocked, row by row of U);
//   k_chol_trsm   the row panel U(k, j) = U_kk^-T S(k, j), one wavefront per column (a lane per row), U_kk in LDS;
//   k_chol_syrk   the trailing upper triangle S(i, j) -= U(k, i)' U(k, j), one workgroup per 64 x 64 tile (tiles
//                 below the diagonal skipped), the two row-panel slices staged in LDS and multiplied on the FP64
//                 matrix instruction (16 x 16 x 4), each wavefront one 32 x 32 quadrant.
// The solve runs U'y = g then U x = y, NB rows at a time: one workgroup solves the diagonal block, then every other
// thread updates the rest of the vector with the block's panel. Every sum is taken in a fixed order (no atomics):
// bitwise reproducible. A failure found by k_chol_diag is read by every later kernel, which then does nothing.
#include <algorithm>
#include <cmath>

#include "bls_internal.h"

using namespace bls;

namespace {
constexpr int NB = 64;
constexpr int SLD = 72;   // LDS row stride of the staged panels (doubles): two-way bank use, the minimum for FP64
typedef double cv4d __attribute__((ext_vector_type(4)));

// U_kk from the block's rows k0 .. k0 + kb of A (upper triangle, column-major, leading dimension n): for each row j
// in turn, the pivot's square root, the row divided by it, and the trailing upper triangle of the block updated.
// A pivot that is not positive (or NaN) records column k0 + j + 1 in *info and stops.
__global__ __launch_bounds__(256) void k_chol_diag(int64_t n, int64_t k0, int kb, double *__restrict__ A,
                                                   int64_t *__restrict__ info) {
    __shared__ double T[NB][NB + 1];
    __shared__ int failed;
    if (*info) return;
    for (int w = threadIdx.x; w < kb * kb; w += blockDim.x) {
        const int c = w / kb, r = w - c * kb;
        if (r <= c) T[r][c] = A[(k0 + r) + (k0 + c) * n];
    }
    if (threadIdx.x == 0) failed = 0;
    __syncthreads();
    for (int j = 0; j < kb; j++) {
        if (threadIdx.x == 0) {
            const double piv = T[j][j];
            if (!(piv > 0.0)) {
                failed = 1;
                *info = k0 + j + 1;
            } else {
                T[j][j] = sqrt(piv);
            }
        }
        __syncthreads();
        if (failed) return;
        const double djj = T[j][j];
        for (int c = j + 1 + (int)threadIdx.x; c < kb; c += blockDim.x) T[j][c] = T[j][c] / djj;
        __syncthreads();
        const int m = kb - j - 1;   // the trailing m x m upper triangle, as pairs (r <= c)
        for (int w = threadIdx.x; w < m * m; w += blockDim.x) {
            const int c = j + 1 + w / m, r = j + 1 + w % m;
            if (r <= c) T[r][c] -= T[j][r] * T[j][c];
        }
        __syncthreads();
    }
    for (int w = threadIdx.x; w < kb * kb; w += blockDim.x) {
        const int c = w / kb, r = w - c * kb;
        if (r <= c) A[(k0 + r) + (k0 + c) * n] = T[r][c];
    }
}

// U(k, j) = U_kk^-T A(k, j) for the columns j >= k0 + kb, TRSM_COLS columns per workgroup, one wavefront per
// column at a time: lane i holds row i; at step i, y_i = v_i / U_ii is broadcast and every lane t > i subtracts
// U(i, t) y_i -- forward substitution, each v_t reduced in row order. Rows past kb are padded (unit diagonal, zero
// coupling).
constexpr int TRSM_COLS = 64;
__global__ __launch_bounds__(256) void k_chol_trsm(int64_t n, int64_t k0, int kb, double *__restrict__ A,
                                                   const int64_t *__restrict__ info) {
    __shared__ double U[NB][NB + 1];
    if (*info) return;
    for (int w = threadIdx.x; w < NB * NB; w += blockDim.x) {
        const int c = w / NB, r = w - c * NB;
        double v = (r == c) ? 1.0 : 0.0;
        if (r < kb && c < kb && r <= c) v = A[(k0 + r) + (k0 + c) * n];
        U[r][c] = v;
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int q = wave; q < TRSM_COLS; q += 4) {
        const int64_t col = k0 + kb + (int64_t)blockIdx.x * TRSM_COLS + q;
        if (col >= n) break;   // the same for the whole wavefront
        double *a = A + k0 + col * n;
        double v = lane < kb ? a[lane] : 0.0;
        for (int i = 0; i < kb; i++) {
            const double yi = __shfl(v, i, 64) / U[i][i];
            if (lane == i) v = yi;
            else if (lane > i) v -= U[i][lane] * yi;
        }
        if (lane < kb) a[lane] = v;
    }
}

// A(i, j) -= U(k, i)' U(k, j) over the trailing upper triangle, one workgroup per 64 x 64 tile (ti <= tj) of the
// trailing columns, the tiles listed column by column. The row-panel slices of both tiles are staged in LDS as
// P[p][c] (p the panel row); wavefront w takes quadrant (w & 1, w >> 1), 2 x 2 tiles of 16 x 16, k in steps of 4:
// lane l supplies A(l % 16, l / 16) = P[k + l / 16][i] and B(l / 16, l % 16) = P[k + l / 16][j], and accumulator v
// holds D(l / 16 + 4 v, l % 16). The accumulators start from A(i, j) and take the negated products.
__global__ __launch_bounds__(256) void k_chol_syrk(int64_t n, int64_t k0, int kb, double *__restrict__ A,
                                                   const int64_t *__restrict__ info) {
    __shared__ double Pi[NB][SLD], Pj[NB][SLD];
    if (*info) return;
    const int64_t t0 = k0 + kb;                       // first trailing row / column
    int64_t tj = 0, rem = blockIdx.x;
    while (rem > tj) {                                // tile (ti, tj): column tj holds tj + 1 tiles
        rem -= tj + 1;
        tj++;
    }
    const int64_t ti = rem;
    const int64_t r0 = t0 + NB * ti, c0 = t0 + NB * tj;
    for (int w = threadIdx.x; w < NB * NB; w += blockDim.x) {
        const int c = w / NB, p = w - c * NB;         // p fastest: the panel rows of one column are contiguous
        const bool ok = p < kb;
        Pi[p][c] = (ok && r0 + c < n) ? A[(k0 + p) + (r0 + c) * n] : 0.0;
        Pj[p][c] = (ok && c0 + c < n) ? A[(k0 + p) + (c0 + c) * n] : 0.0;
    }
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
    const int qi = wave & 1, qj = wave >> 1;
    if (ti == tj && qi > qj) return;                  // entirely below the diagonal
    cv4d acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int64_t r = r0 + 32 * qi + 16 * a + lk + 4 * v, c = c0 + 32 * qj + 16 * b + li;
                acc[a][b][v] = (r < n && c < n && r <= c) ? A[r + c * n] : 0.0;
            }
    for (int k = 0; k < kb; k += 4) {
        const double a0 = -Pi[k + lk][32 * qi + li], a1 = -Pi[k + lk][32 * qi + 16 + li];
        const double b0 = Pj[k + lk][32 * qj + li], b1 = Pj[k + lk][32 * qj + 16 + li];
        acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
    }
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int64_t r = r0 + 32 * qi + 16 * a + lk + 4 * v, c = c0 + 32 * qj + 16 * b + li;
                if (r < n && c < n && r <= c) A[r + c * n] = acc[a][b][v];
            }
}

// U'y = g, block k: solve the diagonal block (rows k0 .. k0 + kb of x, already updated by the earlier blocks)
__global__ __launch_bounds__(NB) void k_chol_fwd_diag(int64_t n, int64_t k0, int kb, const double *__restrict__ A,
                                                      double *__restrict__ x, const int64_t *__restrict__ info) {
    __shared__ double y[NB];
    if (*info) return;
    const int t = threadIdx.x;
    if (t < kb) y[t] = x[k0 + t];
    __syncthreads();
    for (int i = 0; i < kb; i++) {
        if (t == i) y[i] = y[i] / A[(k0 + i) + (k0 + i) * n];
        __syncthreads();
        if (t > i && t < kb) y[t] -= A[(k0 + i) + (k0 + t) * n] * y[i];
        __syncthreads();
    }
    if (t < kb) x[k0 + t] = y[t];
}

// x_j -= U(k, j)' y_k for the rows j >= k0 + kb, one thread per row, the block's values added in row order
__global__ __launch_bounds__(256) void k_chol_fwd_update(int64_t n, int64_t k0, int kb, const double *__restrict__ A,
                                                         double *__restrict__ x, const int64_t *__restrict__ info) {
    if (*info) return;
    const int64_t j = k0 + kb + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const double *u = A + k0 + j * n;
    double s = 0.0;
    for (int p = 0; p < kb; p++) s += u[p] * x[k0 + p];
    x[j] -= s;
}

// U x = y, block k (from the last): solve the diagonal block, rows k0 .. k0 + kb
__global__ __launch_bounds__(NB) void k_chol_bwd_diag(int64_t n, int64_t k0, int kb, const double *__restrict__ A,
                                                      double *__restrict__ x, const int64_t *__restrict__ info) {
    __shared__ double y[NB];
    if (*info) return;
    const int t = threadIdx.x;
    if (t < kb) y[t] = x[k0 + t];
    __syncthreads();
    for (int i = kb - 1; i >= 0; i--) {
        if (t == i) y[i] = y[i] / A[(k0 + i) + (k0 + i) * n];
        __syncthreads();
        if (t < i) y[t] -= A[(k0 + t) + (k0 + i) * n] * y[i];
        __syncthreads();
    }
    if (t < kb) x[k0 + t] = y[t];
}

// y_i -= U(i, k) x_k for the rows i < k0, one thread per row (consecutive rows: coalesced down each column)
__global__ __launch_bounds__(256) void k_chol_bwd_update(int64_t n, int64_t k0, int kb, const double *__restrict__ A,
                                                         double *__restrict__ x, const int64_t *__restrict__ info) {
    if (*info) return;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= k0) return;
    double s = 0.0;
    for (int c = 0; c < kb; c++) s += A[i + (k0 + c) * n] * x[k0 + c];
    x[i] -= s;
}

__global__ void k_chol_copy(int64_t len, const double *__restrict__ src, double *__restrict__ dst) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

__global__ void k_chol_zero(int64_t *p) { p[0] = 0; }
}  // namespace

struct bls_chol {
    int64_t n = 0;
    bool factored = false;
    DevArray<double> U;       // the factor (upper triangle of an n x n column-major array)
    DevArray<int64_t> info;   // 0, or the first column (1-based) whose pivot is not positive
};

extern "C" int bls_chol_create(int64_t n, bls_chol **out) {
    if (!out || n < 0) return fail(BLS_ERR_ARG, "bls_chol_create: bad arguments");
    if (n > (int64_t)1 << 20) return fail(BLS_ERR_UNSUPPORTED, "bls_chol_create: n above 2^20");
    *out = nullptr;
    auto h = new bls_chol();
    h->n = n;
    int rc;
    if ((rc = h->U.alloc((size_t)std::max<int64_t>(1, n * n))) || (rc = h->info.alloc(1))) {
        delete h;
        return rc;
    }
    *out = h;
    return BLS_OK;
}

extern "C" void bls_chol_destroy(bls_chol *h) { delete h; }

extern "C" int bls_chol_factor_f64_async(bls_chol *h, const double *S, void *stream) {
    if (!h) return fail(BLS_ERR_ARG, "null handle");
    if (h->n && !S) return fail(BLS_ERR_ARG, "bls_chol_factor_f64: S is NULL");
    hipStream_t s = (hipStream_t)stream;
    const int64_t n = h->n;
    double *A = h->U.ptr;
    int64_t *info = h->info.ptr;
    k_chol_zero<<<1, 1, 0, s>>>(info);
    if (n) k_chol_copy<<<(unsigned)std::min<int64_t>(8192, grid_for(n * n, 256)), 256, 0, s>>>(n * n, S, A);
    for (int64_t k0 = 0; k0 < n; k0 += NB) {
        const int kb = (int)std::min<int64_t>(NB, n - k0);
        k_chol_diag<<<1, 256, 0, s>>>(n, k0, kb, A, info);
        const int64_t rest = n - k0 - kb;
        if (rest > 0) {
            k_chol_trsm<<<grid_for(rest, TRSM_COLS), 256, 0, s>>>(n, k0, kb, A, info);
            const int64_t nt = (rest + NB - 1) / NB;
            k_chol_syrk<<<(unsigned)(nt * (nt + 1) / 2), 256, 0, s>>>(n, k0, kb, A, info);
        }
    }
    BLS_HIP(hipGetLastError());
    h->factored = true;
    return BLS_OK;
}

extern "C" int bls_chol_factor_f64(bls_chol *h, const double *S, int64_t *failed_col, void *stream) {
    int rc = bls_chol_factor_f64_async(h, S, stream);
    if (rc) return rc;
    int64_t v = 0;
    BLS_HIP(hipMemcpyAsync(&v, h->info.ptr, sizeof(int64_t), hipMemcpyDeviceToHost, (hipStream_t)stream));
    BLS_HIP(hipStreamSynchronize((hipStream_t)stream));
    if (failed_col) *failed_col = v;
    return v ? BLS_FAILED : BLS_OK;
}

extern "C" int bls_chol_solve_f64(bls_chol *h, const double *g, double *x, void *stream) {
    if (!h) return fail(BLS_ERR_ARG, "null handle");
    if (!h->factored) return fail(BLS_ERR_STATE, "bls_chol_solve_f64: no factorization (bls_chol_factor_f64)");
    if (h->n && (!g || !x)) return fail(BLS_ERR_ARG, "bls_chol_solve_f64: g and x are required");
    hipStream_t s = (hipStream_t)stream;
    const int64_t n = h->n;
    const double *A = h->U.ptr;
    const int64_t *info = h->info.ptr;
    if (n) k_chol_copy<<<(unsigned)std::min<int64_t>(1024, grid_for(n, 256)), 256, 0, s>>>(n, g, x);
    for (int64_t k0 = 0; k0 < n; k0 += NB) {
        const int kb = (int)std::min<int64_t>(NB, n - k0);
        k_chol_fwd_diag<<<1, NB, 0, s>>>(n, k0, kb, A, x, info);
        if (n - k0 - kb > 0) k_chol_fwd_update<<<grid_for(n - k0 - kb, 256), 256, 0, s>>>(n, k0, kb, A, x, info);
    }
    for (int64_t k0 = ((n - 1) / NB) * NB; n && k0 >= 0; k0 -= NB) {
        const int kb = (int)std::min<int64_t>(NB, n - k0);
        k_chol_bwd_diag<<<1, NB, 0, s>>>(n, k0, kb, A, x, info);
        if (k0 > 0) k_chol_bwd_update<<<grid_for(k0, 256), 256, 0, s>>>(n, k0, kb, A, x, info);
    }
    BLS_HIP(hipGetLastError());
    return BLS_OK;
}

extern "C" int bls_chol_factor_copy_f64(bls_chol *h, double *U, void *stream) {
    if (!h || !U) return fail(BLS_ERR_ARG, "bls_chol_factor_copy_f64: bad arguments");
    if (!h->factored) return fail(BLS_ERR_STATE, "bls_chol_factor_copy_f64: no factorization");
    if (h->n)
        k_chol_copy<<<(unsigned)std::min<int64_t>(8192, grid_for(h->n * h->n, 256)), 256, 0, (hipStream_t)stream>>>(
            h->n * h->n, h->U.ptr, U);
    BLS_HIP(hipGetLastError());
    return BLS_OK;
}
