// The dense reduced solve on gfx950: S x = g for the dense reduced system S (nr x nr, its upper triangle stored
// column-major, as reduce_schur_dense! writes it), by the Cholesky factorization S = U'U of a copy (S is left as
// built), behind the C ABI of include/bls_amd.h. This is the :cholesky branch of solve_reduced_dense!
// (src/schur_linear_systems.jl:312-319), whose factorization the reference takes from DampedLinearSolve (LAPACK's
// potrf); a matrix that is not positive definite returns BLS_FAILED with the first column whose pivot is not
// positive (NotPositiveDefinite), the caller's cue to fall back to an indefinite factorization.
//
// Blocked right-looking factorization with a look-ahead of one panel, panels of NB = 64 rows of U:
//   k_chol_diag   one wavefront factors the diagonal block, a column per lane in registers (unblocked, row by row of U);
//   k_chol_trsm   the row panel U(k, j) = U_kk^-T S(k, j), one wavefront per column (a lane per row), U_kk in LDS;
//   k_chol_syrk   the trailing upper triangle S(i, j) -= U(k, i)' U(k, j), one workgroup per 64 x 64 tile (tiles
//                 below the diagonal skipped), the two row-panel slices staged in LDS and multiplied on the FP64
//                 matrix instruction (16 x 16 x 4), each wavefront one 32 x 32 quadrant.
// The solve runs U'y = g then U x = y, NB rows at a time: one wavefront solves the diagonal block, then the rest of
// the vector is updated with the block's panel (a wavefront per entry going forward, a thread per entry going back). Every sum is taken in a fixed order (no atomics):
// bitwise reproducible. A failure found by k_chol_diag is read by every later kernel, which then does nothing.
#include <algorithm>
#include <cmath>

#include "bls_internal.h"

using namespace bls;

namespace {
constexpr int NB = 64;
constexpr int SLD = 72;   // LDS row stride of the staged panels (doubles): two-way bank use, the minimum for FP64
constexpr int SYRK_ALL = 0, SYRK_FIRST = 1, SYRK_REST = 2;   // which tiles of the trailing update
typedef double cv4d __attribute__((ext_vector_type(4)));

// lane i's v, to every lane (i the same for the whole wavefront)
__device__ inline double bcast(double v, int i) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), i);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), i);
    return __hiloint2double(hi, lo);
}

// The 64 x 64 diagonal block's upper triangle (rows and columns past kb: zero) into LDS, the global loads of each
// thread all issued before its LDS stores (16 per thread)
__device__ inline void stage_diag_block(int64_t n, int64_t k0, int kb, const double *__restrict__ A,
                                        double (*T)[NB + 1], double pad) {
    double v[NB * NB / 256];
#pragma unroll
    for (int q = 0; q < NB * NB / 256; q++) {
        const int w = threadIdx.x + 256 * q, c = w >> 6, r = w & 63;
        v[q] = (r < kb && c < kb && r <= c) ? A[(k0 + r) + (k0 + c) * n] : (r == c ? pad : 0.0);
    }
#pragma unroll
    for (int q = 0; q < NB * NB / 256; q++) {
        const int w = threadIdx.x + 256 * q;
        T[w & 63][w >> 6] = v[q];
    }
}

// U_kk from the block's rows k0 .. k0 + kb of A (upper triangle, column-major, leading dimension n) on one wavefront:
// lane c keeps column c of the block in registers (staged through LDS with coalesced loads). At row j every lane reads
// the pivot u = T(j, j) from lane j, forms sqrt and its reciprocal, scales its entry of row j (as LAPACK's dpotf2 scales
// the row), and updates T(r, c) -= u_jr u_jc for every r > j with u_jr read from lane r; entries below the diagonal
// are updated too but never read or stored. A pivot that is not positive (or NaN) records column k0 + j + 1 in *info
// and stops.
__global__ __launch_bounds__(64) void k_chol_diag(int64_t n, int64_t k0, int kb, double *__restrict__ A,
                                                  int64_t *__restrict__ info) {
    __shared__ double T[NB][NB + 1];
    if (*info) return;
    const int lane = threadIdx.x;
    for (int c = 0; c < NB; c += 16) {   // lane r reads row r of 16 columns (coalesced), loads before stores
        double v[16];
#pragma unroll
        for (int q = 0; q < 16; q++)
            v[q] = (lane < kb && c + q < kb && lane <= c + q) ? A[(k0 + lane) + (k0 + c + q) * n] : 0.0;
#pragma unroll
        for (int q = 0; q < 16; q++) T[lane][c + q] = v[q];
    }
    __syncthreads();
    double col[NB];
#pragma unroll
    for (int r = 0; r < NB; r++) col[r] = T[r][lane];
    bool ok = true;
#pragma unroll
    for (int j = 0; j < NB; j++) {
        if (j < kb && ok) {   // uniform: kb and the pivots are the same on every lane
            const double piv = bcast(col[j], j);
            if (!(piv > 0.0)) {
                if (lane == 0) *info = k0 + j + 1;
                ok = false;
            } else {
                const double ujj = sqrt(piv), rinv = 1.0 / ujj;
                col[j] = lane == j ? ujj : col[j] * rinv;   // row j of U (lanes c > j); lanes c < j hold unused values
#pragma unroll
                for (int r = j + 1; r < NB; r++) col[r] -= bcast(col[j], r) * col[j];
            }
        }
    }
    if (!ok) return;
#pragma unroll
    for (int r = 0; r < NB; r++) T[r][lane] = col[r];
    __syncthreads();
    for (int c = 0; c < kb; c++)
        if (lane <= c && lane < kb) A[(k0 + lane) + (k0 + c) * n] = T[lane][c];
}

// U(k, j) = U_kk^-T A(k, j) for the columns j >= k0 + kb, TRSM_COLS columns per workgroup, four per wavefront
// side by side: lane i holds row i; at step i, y_i = v_i (1 / U_ii) is broadcast and every lane t > i subtracts
// U(i, t) y_i -- forward substitution, each v_t reduced in row order. Rows past kb are padded (unit diagonal, zero
// coupling).
constexpr int TRSM_COLS = 16;
__global__ __launch_bounds__(256) void k_chol_trsm(int64_t n, int64_t k0, int kb, double *__restrict__ A,
                                                   const int64_t *__restrict__ info) {
    __shared__ double U[NB][NB + 1];
    __shared__ double dinv[NB];
    if (*info) return;
    stage_diag_block(n, k0, kb, A, U, 1.0);
    __syncthreads();
    if (threadIdx.x < NB) dinv[threadIdx.x] = 1.0 / U[threadIdx.x][threadIdx.x];
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t cb = k0 + kb + (int64_t)blockIdx.x * TRSM_COLS + 4 * wave;   // four columns per wavefront
    double v[4];
#pragma unroll
    for (int u = 0; u < 4; u++) v[u] = (cb + u < n && lane < kb) ? A[k0 + lane + (cb + u) * n] : 0.0;
    for (int i = 0; i < kb; i++) {   // the four substitutions interleaved
        const double di = dinv[i], uil = U[i][lane];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const double yi = bcast(v[u], i) * di;
            v[u] = lane == i ? yi : (lane > i ? v[u] - uil * yi : v[u]);
        }
    }
#pragma unroll
    for (int u = 0; u < 4; u++)
        if (cb + u < n && lane < kb) A[k0 + lane + (cb + u) * n] = v[u];
}

// A(i, j) -= U(k, i)' U(k, j) over the trailing upper triangle, one workgroup per 64 x 64 tile (ti <= tj) of the
// trailing columns, the tiles listed column by column: all of them, the first tile row only (the next panel's rows),
// or the others. The row-panel slices of both tiles are staged in LDS as
// P[p][c] (p the panel row); wavefront w takes quadrant (w & 1, w >> 1), 2 x 2 tiles of 16 x 16, k in steps of 4:
// lane l supplies A(l % 16, l / 16) = P[k + l / 16][i] and B(l / 16, l % 16) = P[k + l / 16][j], and accumulator v
// holds D(l / 16 + 4 v, l % 16). The accumulators start from A(i, j) and take the negated products.
__global__ __launch_bounds__(256) void k_chol_syrk(int64_t n, int64_t k0, int kb, double *__restrict__ A,
                                                   const int64_t *__restrict__ info, int rows) {
    __shared__ double Pi[NB][SLD], Pj[NB][SLD];
    if (*info) return;
    const int64_t t0 = k0 + kb;                       // first trailing row / column
    int64_t ti, tj;
    if (rows == SYRK_FIRST) {                         // the first tile row: tile (0, blockIdx)
        ti = 0;
        tj = blockIdx.x;
    } else if (rows == SYRK_REST) {                   // the others: column tj >= 1 holds tiles 1 .. tj
        int64_t rem = blockIdx.x;
        tj = 1;
        while (rem >= tj) {
            rem -= tj;
            tj++;
        }
        ti = rem + 1;
    } else {                                          // all: column tj holds tiles 0 .. tj
        int64_t rem = blockIdx.x;
        tj = 0;
        while (rem > tj) {
            rem -= tj + 1;
            tj++;
        }
        ti = rem;
    }
    const int64_t r0 = t0 + NB * ti, c0 = t0 + NB * tj;
    {   // p fastest: the panel rows of one column are contiguous; all loads issued before the LDS stores
        double vi[NB * NB / 256], vj[NB * NB / 256];
#pragma unroll
        for (int q = 0; q < NB * NB / 256; q++) {
            const int w = threadIdx.x + 256 * q, c = w >> 6, p = w & 63;
            const bool ok = p < kb;
            vi[q] = (ok && r0 + c < n) ? A[(k0 + p) + (r0 + c) * n] : 0.0;
            vj[q] = (ok && c0 + c < n) ? A[(k0 + p) + (c0 + c) * n] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < NB * NB / 256; q++) {
            const int w = threadIdx.x + 256 * q;
            Pi[w & 63][w >> 6] = vi[q];
            Pj[w & 63][w >> 6] = vj[q];
        }
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

// The diagonal block's rows into LDS by one wavefront, lane t reading row t of every column (coalesced), in groups
// of 16 loads issued before their stores
__device__ inline void stage_rows(int64_t n, int64_t k0, int kb, const double *__restrict__ A, double (*U)[NB + 1]) {
    const int t = threadIdx.x;
    for (int c0 = 0; c0 < NB; c0 += 16) {
        double v[16];
#pragma unroll
        for (int q = 0; q < 16; q++) {
            const int c = c0 + q;
            v[q] = (t < kb && c < kb && t <= c) ? A[(k0 + t) + (k0 + c) * n] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < 16; q++) U[t][c0 + q] = v[q];
    }
}

// U'y = g, block k: solve the diagonal block (rows k0 .. k0 + kb of x, already updated by the earlier blocks)
__global__ __launch_bounds__(NB) void k_chol_fwd_diag(int64_t n, int64_t k0, int kb, const double *__restrict__ A,
                                                      double *__restrict__ x, const int64_t *__restrict__ info) {
    __shared__ double U[NB][NB + 1];
    __shared__ double dinv[NB];
    if (*info) return;
    const int t = threadIdx.x;
    stage_rows(n, k0, kb, A, U);
    double v = t < kb ? x[k0 + t] : 0.0;
    __syncthreads();
    dinv[t] = t < kb ? 1.0 / U[t][t] : 1.0;
    __syncthreads();
    for (int i = 0; i < kb; i++) {   // lane t holds row t: y_i broadcast, then every t > i updated
        const double yi = bcast(v, i) * dinv[i];
        v = t == i ? yi : (t > i ? v - U[i][t] * yi : v);
    }
    if (t < kb) x[k0 + t] = v;
}

// x_j -= U(k, j)' y_k for the rows j >= k0 + kb, one wavefront per row j: lane p takes U(k0 + p, j) y_p (the
// column's kb values, contiguous), added in a fixed tree
__global__ __launch_bounds__(256) void k_chol_fwd_update(int64_t n, int64_t k0, int kb, const double *__restrict__ A,
                                                         double *__restrict__ x, const int64_t *__restrict__ info) {
    if (*info) return;
    const int64_t j = k0 + kb + (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int p = threadIdx.x & 63;
    if (j >= n) return;   // the whole wavefront
    double s = p < kb ? A[k0 + p + j * n] * x[k0 + p] : 0.0;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o, 64);
    if (p == 0) x[j] -= s;
}

// U x = y, block k (from the last): solve the diagonal block, rows k0 .. k0 + kb
__global__ __launch_bounds__(NB) void k_chol_bwd_diag(int64_t n, int64_t k0, int kb, const double *__restrict__ A,
                                                      double *__restrict__ x, const int64_t *__restrict__ info) {
    __shared__ double U[NB][NB + 1];
    __shared__ double dinv[NB];
    if (*info) return;
    const int t = threadIdx.x;
    stage_rows(n, k0, kb, A, U);
    double v = t < kb ? x[k0 + t] : 0.0;
    __syncthreads();
    dinv[t] = t < kb ? 1.0 / U[t][t] : 1.0;
    __syncthreads();
    for (int i = kb - 1; i >= 0; i--) {   // lane t holds row t: x_i broadcast, then every t < i updated
        const double xi = bcast(v, i) * dinv[i];
        v = t == i ? xi : (t < i ? v - U[t][i] * xi : v);
    }
    if (t < kb) x[k0 + t] = v;
}

// y_i -= U(i, k) x_k for the rows i < k0, one thread per row (consecutive rows: coalesced down each column)
__global__ __launch_bounds__(256) void k_chol_bwd_update(int64_t n, int64_t k0, int kb, const double *__restrict__ A,
                                                         double *__restrict__ x, const int64_t *__restrict__ info) {
    if (*info) return;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= k0) return;
    double s = 0.0;
    int c = 0;
    for (; c + 16 <= kb; c += 16) {   // sixteen loads in flight, added in column order
        double u[16];
#pragma unroll
        for (int q = 0; q < 16; q++) u[q] = A[i + (k0 + c + q) * n];
#pragma unroll
        for (int q = 0; q < 16; q++) s += u[q] * x[k0 + c + q];
    }
    for (; c < kb; c++) s += A[i + (k0 + c) * n] * x[k0 + c];
    x[i] -= s;
}

template <class T>
__global__ void k_chol_copy(int64_t len, const T *__restrict__ src, double *__restrict__ dst) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = (double)src[i];
}

__global__ void k_chol_zero(int64_t *p) { p[0] = 0; }
}  // namespace

struct bls_chol {
    int64_t n = 0;
    bool factored = false;
    DevArray<double> U;       // the factor (upper triangle of an n x n column-major array)
    DevArray<int64_t> info;   // 0, or the first column (1-based) whose pivot is not positive
    // the look-ahead: the bulk of each trailing update runs on a second stream (default priority)
    hipStream_t s2 = nullptr;
    hipEvent_t ev_panel = nullptr, ev_bulk = nullptr;
    ~bls_chol() {
        if (ev_panel) (void)hipEventDestroy(ev_panel);
        if (ev_bulk) (void)hipEventDestroy(ev_bulk);
        if (s2) (void)hipStreamDestroy(s2);
    }
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
    hipError_t e = hipStreamCreateWithFlags(&h->s2, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&h->ev_panel, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&h->ev_bulk, hipEventDisableTiming);
    if (e != hipSuccess) {
        delete h;
        return fail(BLS_ERR_HIP, std::string("bls_chol_create: ") + hipGetErrorString(e));
    }
    *out = h;
    return BLS_OK;
}

extern "C" void bls_chol_destroy(bls_chol *h) { delete h; }

namespace {
template <class T> int chol_factor(bls_chol *h, const T *S, hipStream_t s);
}

extern "C" int bls_chol_factor_f64_async(bls_chol *h, const double *S, void *stream) {
    return chol_factor<double>(h, S, (hipStream_t)stream);
}
extern "C" int bls_chol_factor_f32_async(bls_chol *h, const float *S, void *stream) {
    return chol_factor<float>(h, S, (hipStream_t)stream);
}

namespace {
template <class T> int chol_factor(bls_chol *h, const T *S, hipStream_t s) {
    if (!h) return fail(BLS_ERR_ARG, "null handle");
    if (h->n && !S) return fail(BLS_ERR_ARG, "bls_chol_factor: S is NULL");
    const int64_t n = h->n;
    double *A = h->U.ptr;
    int64_t *info = h->info.ptr;
    k_chol_zero<<<1, 1, 0, s>>>(info);
    if (n) k_chol_copy<T><<<(unsigned)std::min<int64_t>(8192, grid_for(n * n, 256)), 256, 0, s>>>(n * n, S, A);
    // Panel k on the caller's stream: its row solve, then the first tile row of its trailing update (the rows of
    // panel k + 1), then panel k + 1's diagonal block. The other tile rows of the update go to the second stream once
    // the row solve is done, and run beside panel k + 1's diagonal block and row solve; panel k + 1's own update waits
    // for them (they write the rows of panel k + 2 it reads and updates).
    bool bulk_pending = false;
    if (n) k_chol_diag<<<1, 64, 0, s>>>(n, 0, (int)std::min<int64_t>(NB, n), A, info);
    for (int64_t k0 = 0; k0 < n; k0 += NB) {
        const int kb = (int)std::min<int64_t>(NB, n - k0);
        const int64_t rest = n - k0 - kb;
        if (rest <= 0) break;
        k_chol_trsm<<<grid_for(rest, TRSM_COLS), 256, 0, s>>>(n, k0, kb, A, info);
        if (bulk_pending) BLS_HIP(hipStreamWaitEvent(s, h->ev_bulk, 0));   // panel k - 1's bulk update
        const int64_t nt = (rest + NB - 1) / NB;
        bulk_pending = nt > 1;
        if (bulk_pending) {
            BLS_HIP(hipEventRecord(h->ev_panel, s));
            BLS_HIP(hipStreamWaitEvent(h->s2, h->ev_panel, 0));
            k_chol_syrk<<<(unsigned)(nt * (nt - 1) / 2), 256, 0, h->s2>>>(n, k0, kb, A, info, SYRK_REST);
            BLS_HIP(hipEventRecord(h->ev_bulk, h->s2));
        }
        k_chol_syrk<<<(unsigned)nt, 256, 0, s>>>(n, k0, kb, A, info, SYRK_FIRST);
        k_chol_diag<<<1, 64, 0, s>>>(n, k0 + kb, (int)std::min<int64_t>(NB, rest), A, info);
    }
    if (bulk_pending) BLS_HIP(hipStreamWaitEvent(s, h->ev_bulk, 0));
    BLS_HIP(hipGetLastError());
    h->factored = true;
    return BLS_OK;
}

int chol_wait(bls_chol *h, int64_t *failed_col, void *stream);
}  // namespace

extern "C" int bls_chol_factor_f64(bls_chol *h, const double *S, int64_t *failed_col, void *stream) {
    int rc = bls_chol_factor_f64_async(h, S, stream);
    return rc ? rc : chol_wait(h, failed_col, stream);
}
extern "C" int bls_chol_factor_f32(bls_chol *h, const float *S, int64_t *failed_col, void *stream) {
    int rc = bls_chol_factor_f32_async(h, S, stream);
    return rc ? rc : chol_wait(h, failed_col, stream);
}

namespace {
int chol_wait(bls_chol *h, int64_t *failed_col, void *stream) {
    int64_t v = 0;
    BLS_HIP(hipMemcpyAsync(&v, h->info.ptr, sizeof(int64_t), hipMemcpyDeviceToHost, (hipStream_t)stream));
    BLS_HIP(hipStreamSynchronize((hipStream_t)stream));
    if (failed_col) *failed_col = v;
    return v ? BLS_FAILED : BLS_OK;
}
}  // namespace

namespace {
// r = t - S x for a symmetric S of order n stored by its upper triangle (column-major): one wavefront per row a,
// S[a, b] read from column max(a, b); the row's products summed by lanes in a fixed butterfly.
__global__ __launch_bounds__(256) void k_dense_residual(int64_t n, const double *S, const double *t, const double *x,
                                                        double *r) {
    const int64_t a = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (a >= n) return;
    double s = 0.0;
    for (int64_t b = lane; b < n; b += 64) s += (a <= b ? S[a + b * n] : S[b + a * n]) * x[b];
    for (int d = 32; d >= 1; d >>= 1) s += __shfl_xor(s, d, 64);
    if (lane == 0) r[a] = t[a] - s;
}
}  // namespace

extern "C" int bls_dense_residual_f64(int64_t n, const double *S, const double *t, const double *x, double *r,
                                      void *stream) {
    if (n < 0 || (n && (!S || !t || !x || !r))) return fail(BLS_ERR_ARG, "bls_dense_residual_f64: bad arguments");
    if (n) k_dense_residual<<<(unsigned)((n + 3) / 4), 256, 0, (hipStream_t)stream>>>(n, S, t, x, r);
    BLS_HIP(hipGetLastError());
    return BLS_OK;
}

extern "C" int bls_chol_solve_f64(bls_chol *h, const double *g, double *x, void *stream) {
    if (!h) return fail(BLS_ERR_ARG, "null handle");
    if (!h->factored) return fail(BLS_ERR_STATE, "bls_chol_solve_f64: no factorization (bls_chol_factor_f64)");
    if (h->n && (!g || !x)) return fail(BLS_ERR_ARG, "bls_chol_solve_f64: g and x are required");
    hipStream_t s = (hipStream_t)stream;
    const int64_t n = h->n;
    const double *A = h->U.ptr;
    const int64_t *info = h->info.ptr;
    if (n) k_chol_copy<double><<<(unsigned)std::min<int64_t>(1024, grid_for(n, 256)), 256, 0, s>>>(n, g, x);
    for (int64_t k0 = 0; k0 < n; k0 += NB) {
        const int kb = (int)std::min<int64_t>(NB, n - k0);
        k_chol_fwd_diag<<<1, NB, 0, s>>>(n, k0, kb, A, x, info);
        if (n - k0 - kb > 0) k_chol_fwd_update<<<grid_for(n - k0 - kb, 4), 256, 0, s>>>(n, k0, kb, A, x, info);
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
        k_chol_copy<double><<<(unsigned)std::min<int64_t>(8192, grid_for(h->n * h->n, 256)), 256, 0, (hipStream_t)stream>>>(
            h->n * h->n, h->U.ptr, U);
    BLS_HIP(hipGetLastError());
    return BLS_OK;
}
