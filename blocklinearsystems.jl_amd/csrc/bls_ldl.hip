// Bunch-Kaufman factorization of the dense reduced S and its solve: the :ldl fallback of the reference's dense
// reduced solve (solve_reduced_dense!, src/schur_linear_systems.jl:312-319, through DampedLinearSolve's
// bunchkaufman, i.e. LAPACK sytrf), behind the C ABI of include/bls_amd.h.
//
// The factorization is LAPACK's unblocked dsytf2 on the lower triangle: S = P L D L' P', alpha = (1 + sqrt 17) / 8,
// 1 x 1 and 2 x 2 pivots with symmetric interchanges, LAPACK's ipiv. Column k is one pair of launches: k_bk_pivot
// (one workgroup: the column and row maxima by fixed-order reductions, the pivot decision, the interchange, the
// 2 x 2 multipliers; it also finishes the previous step's columns of L) and k_bk_update (the rank-1 or rank-2
// update of the trailing lower triangle, 32 x 32 tiles over the whole GPU). Pivot decisions live in a device state
// record, so the host queues all 2n launches without waiting; a launch at a column that a 2 x 2 pivot already took
// does nothing. The solve is dsytrs ('L') in one workgroup.
#include <algorithm>
#include <cmath>

#include "bls_internal.h"

using namespace bls;

namespace {

constexpr int BK_THREADS = 1024;
constexpr int BK_TILE = 32;

struct BKState {
    int64_t next;      // first column not yet factored
    int64_t cur;       // column of the step the last pivot launch took
    int64_t info;      // first zero pivot (1-based), 0 none
    int32_t kstep;     // that step's trailing update: 1 (rank 1), 2 (rank 2), 0 none
    int32_t pending;   // its columns of L still to finish: 1 (scale by d11), 2 (copy the multipliers), 0 none
    double d11;        // 1 / D(k) of a 1 x 1 step
};

__global__ void k_bk_init(BKState *st) {
    st->next = 0;
    st->cur = -1;
    st->info = 0;
    st->kstep = 0;
    st->pending = 0;
    st->d11 = 0.0;
}

// A = the symmetric matrix of S's upper triangle, both triangles written
__global__ void k_bk_copy(const double *S, double *A, int64_t n) {
    const int64_t nn = n * n;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < nn; q += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = q % n, j = q / n;
        A[q] = i <= j ? S[q] : S[i * n + j];
    }
}

// Largest value of the workgroup with the smallest index among equal ones (IDAMAX takes the first).
__device__ inline void bk_argmax(double &v, int64_t &i, double *sv, int64_t *si) {
    for (int o = 32; o >= 1; o >>= 1) {
        const double v2 = __shfl_xor(v, o, 64);
        const int64_t i2 = __shfl_xor(i, o, 64);
        if (v2 > v || (v2 == v && i2 < i)) { v = v2; i = i2; }
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { sv[w] = v; si[w] = i; }
    __syncthreads();
    v = sv[0];
    i = si[0];
    for (int q = 1; q < BK_THREADS / 64; q++)
        if (sv[q] > v || (sv[q] == v && si[q] < i)) { v = sv[q]; i = si[q]; }
    __syncthreads();
}

#pragma clang fp contract(off)
__global__ __launch_bounds__(BK_THREADS) void k_bk_pivot(double *A, int64_t n, BKState *st, double *w,
                                                         int64_t *ipiv, int64_t c) {
    __shared__ double sv[BK_THREADS / 64];
    __shared__ int64_t si[BK_THREADS / 64];
    const int t = threadIdx.x;
    const BKState S = *st;
    // the previous step's columns of L, now that its trailing update has run
    if (S.pending == 1) {
        for (int64_t i = S.cur + 1 + t; i < n; i += BK_THREADS) A[S.cur * n + i] *= S.d11;
    } else if (S.pending == 2) {
        for (int64_t i = S.cur + 2 + t; i < n; i += BK_THREADS) {
            A[S.cur * n + i] = w[i];
            A[(S.cur + 1) * n + i] = w[n + i];
        }
    }
    __syncthreads();   // every thread has read *st
    if (c >= n || c != S.next) {   // past the end, or the second column of a 2 x 2 pivot
        if (t == 0) { st->kstep = 0; st->pending = 0; }
        return;
    }
    const int64_t k = c;
    const double alpha = (1.0 + sqrt(17.0)) / 8.0;
    double colmax = -1.0;
    int64_t imax = n;
    for (int64_t i = k + 1 + t; i < n; i += BK_THREADS) {
        const double v = fabs(A[k * n + i]);
        if (v > colmax) { colmax = v; imax = i; }
    }
    bk_argmax(colmax, imax, sv, si);
    if (k == n - 1) { colmax = 0.0; imax = k; }
    const double absakk = fabs(A[k * n + k]);
    int kstep = 1;
    int64_t kp = k;
    const bool zero = fmax(absakk, colmax) == 0.0 || isnan(absakk);
    if (!zero && !(absakk >= alpha * colmax)) {
        // the largest off-diagonal entry of row / column imax of the trailing part
        double rowmax = -1.0;
        int64_t jm = n;
        for (int64_t j = k + t; j < imax; j += BK_THREADS) {
            const double v = fabs(A[j * n + imax]);
            if (v > rowmax) { rowmax = v; jm = j; }
        }
        for (int64_t i = imax + 1 + t; i < n; i += BK_THREADS) {
            const double v = fabs(A[imax * n + i]);
            if (v > rowmax) { rowmax = v; jm = i; }
        }
        bk_argmax(rowmax, jm, sv, si);
        if (absakk >= alpha * colmax * (colmax / rowmax)) kp = k;
        else if (fabs(A[imax * n + imax]) >= alpha * rowmax) kp = imax;
        else { kp = imax; kstep = 2; }
    }
    const int64_t kk = k + kstep - 1;
    if (!zero && kp != kk) {   // interchange rows and columns kk and kp of the trailing lower triangle
        for (int64_t i = kp + 1 + t; i < n; i += BK_THREADS) {
            const double a = A[kk * n + i];
            A[kk * n + i] = A[kp * n + i];
            A[kp * n + i] = a;
        }
        for (int64_t j = kk + 1 + t; j < kp; j += BK_THREADS) {
            const double a = A[kk * n + j];
            A[kk * n + j] = A[j * n + kp];
            A[j * n + kp] = a;
        }
        if (t == 0) {
            const double a = A[kk * n + kk];
            A[kk * n + kk] = A[kp * n + kp];
            A[kp * n + kp] = a;
            if (kstep == 2) {
                const double b = A[k * n + k + 1];
                A[k * n + k + 1] = A[k * n + kp];
                A[k * n + kp] = b;
            }
        }
    }
    __syncthreads();
    int32_t upd = 0, pend = 0;
    double d11 = 0.0;
    if (!zero) {
        if (kstep == 1) {
            if (k < n - 1) { d11 = 1.0 / A[k * n + k]; upd = 1; pend = 1; }
        } else if (k < n - 2) {
            // the 2 x 2 multipliers W = A(k+2:n, k:k+1) D^-1 (dsytf2), kept aside until the update has run
            double d21 = A[k * n + k + 1];
            const double e11 = A[(k + 1) * n + k + 1] / d21;
            const double e22 = A[k * n + k] / d21;
            const double tt = 1.0 / (e11 * e22 - 1.0);
            d21 = tt / d21;
            for (int64_t j = k + 2 + t; j < n; j += BK_THREADS) {
                w[j] = d21 * (e11 * A[k * n + j] - A[(k + 1) * n + j]);
                w[n + j] = d21 * (e22 * A[(k + 1) * n + j] - A[k * n + j]);
            }
            upd = 2;
            pend = 2;
        }
    }
    if (t == 0) {
        if (kstep == 1) ipiv[k] = kp + 1;
        else ipiv[k] = ipiv[k + 1] = -(kp + 1);
        if (zero && S.info == 0) st->info = k + 1;
        st->next = k + kstep;
        st->cur = k;
        st->kstep = upd;
        st->pending = pend;
        st->d11 = d11;
    }
}

// The trailing update of the step at column c: one 32 x 32 tile of the lower triangle of rows / columns c + 1 ..
// n - 1 per workgroup (blockIdx.x numbers the tiles on and below the diagonal); columns before c + kstep skipped.
__global__ __launch_bounds__(256) void k_bk_update(double *A, int64_t n, const BKState *st, const double *w,
                                                   int64_t c) {
    const BKState S = *st;
    if (S.kstep <= 0 || S.cur != c) return;
    const int64_t k = c, j0 = k + S.kstep, base = k + 1;
    const int64_t tid = blockIdx.x;
    int64_t ti = (int64_t)((sqrt(8.0 * (double)tid + 1.0) - 1.0) * 0.5);
    while ((ti + 1) * (ti + 2) / 2 <= tid) ti++;
    while (ti * (ti + 1) / 2 > tid) ti--;
    const int64_t tj = tid - ti * (ti + 1) / 2;
    const int64_t i = base + ti * BK_TILE + (threadIdx.x % BK_TILE);
    if (i >= n) return;
    const double xi = A[k * n + i];
    const double yi = S.kstep == 2 ? A[(k + 1) * n + i] : 0.0;
    for (int q = threadIdx.x / BK_TILE; q < BK_TILE; q += 256 / BK_TILE) {
        const int64_t j = base + tj * BK_TILE + q;
        if (j < j0 || j > i) continue;
        if (S.kstep == 1) A[j * n + i] += xi * (-S.d11 * A[k * n + j]);
        else A[j * n + i] = A[j * n + i] - xi * w[j] - yi * w[n + j];
    }
}

// Sum over the workgroup in a fixed order (a butterfly per wavefront, then the wavefronts in order), to all threads.
__device__ inline double bk_sum(double v, double *sv) {
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) sv[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0.0;
    for (int q = 0; q < BK_THREADS / 64; q++) s += sv[q];
    __syncthreads();
    return s;
}

// dsytrs ('L'): x = S^-1 g from the factor, in one workgroup; x holds g on entry.
__global__ __launch_bounds__(BK_THREADS) void k_bk_solve(const double *A, int64_t n, const int64_t *ipiv, double *x) {
    __shared__ double sv[BK_THREADS / 64];
    const int t = threadIdx.x;
    int64_t k = 0;
    while (k < n) {   // L D y = P g
        const int64_t p = ipiv[k];
        if (p > 0) {
            if (t == 0 && p - 1 != k) { const double a = x[k]; x[k] = x[p - 1]; x[p - 1] = a; }
            __syncthreads();
            const double bk = x[k];
            for (int64_t i = k + 1 + t; i < n; i += BK_THREADS) x[i] -= A[k * n + i] * bk;
            __syncthreads();
            if (t == 0) x[k] *= 1.0 / A[k * n + k];
            k += 1;
        } else {
            const int64_t kp = -p - 1;
            if (t == 0 && kp != k + 1) { const double a = x[k + 1]; x[k + 1] = x[kp]; x[kp] = a; }
            __syncthreads();
            const double b0 = x[k], b1 = x[k + 1];
            for (int64_t i = k + 2 + t; i < n; i += BK_THREADS) {
                x[i] -= A[k * n + i] * b0;
                x[i] -= A[(k + 1) * n + i] * b1;
            }
            __syncthreads();
            if (t == 0) {
                const double akm1k = A[k * n + k + 1];
                const double akm1 = A[k * n + k] / akm1k;
                const double ak = A[(k + 1) * n + k + 1] / akm1k;
                const double denom = akm1 * ak - 1.0;
                const double bkm1 = b0 / akm1k;
                const double bkk = b1 / akm1k;
                x[k] = (ak * bkm1 - bkk) / denom;
                x[k + 1] = (akm1 * bkk - bkm1) / denom;
            }
            k += 2;
        }
        __syncthreads();
    }
    k = n - 1;
    while (k >= 0) {   // L' P' x = y
        const int64_t p = ipiv[k];
        double d0 = 0.0, d1 = 0.0;
        for (int64_t i = k + 1 + t; i < n; i += BK_THREADS) {
            d0 += A[k * n + i] * x[i];
            if (p < 0) d1 += A[(k - 1) * n + i] * x[i];
        }
        d0 = bk_sum(d0, sv);
        if (p < 0) d1 = bk_sum(d1, sv);
        if (t == 0) {
            x[k] -= d0;
            if (p < 0) x[k - 1] -= d1;
            const int64_t kp = (p > 0 ? p : -p) - 1;
            if (kp != k) { const double a = x[k]; x[k] = x[kp]; x[kp] = a; }
        }
        k -= p > 0 ? 1 : 2;
        __syncthreads();
    }
}
#pragma clang fp contract(fast)

// D's diagonal and subdiagonal, for the inertia
__global__ void k_bk_diag(const double *A, int64_t n, double *d) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    d[i] = A[i * n + i];
    d[n + i] = i + 1 < n ? A[i * n + i + 1] : 0.0;
}

}  // namespace

struct bls_ldl {
    int64_t n = 0;
    bool factored = false;
    DevArray<double> A, w;
    DevArray<int64_t> ipiv;
    DevArray<BKState> st;
};

extern "C" int bls_ldl_create(int64_t n, bls_ldl **out) {
    if (!out || n < 0) return fail(BLS_ERR_ARG, "bls_ldl_create: bad arguments");
    if (n > (int64_t)1 << 20) return fail(BLS_ERR_UNSUPPORTED, "bls_ldl_create: n above 2^20");
    *out = nullptr;
    auto h = new bls_ldl();
    h->n = n;
    int rc;
    if ((rc = h->A.alloc((size_t)std::max<int64_t>(1, n * n))) || (rc = h->w.alloc((size_t)std::max<int64_t>(2, 2 * n))) ||
        (rc = h->ipiv.alloc((size_t)std::max<int64_t>(1, n))) || (rc = h->st.alloc(1))) {
        delete h;
        return rc;
    }
    *out = h;
    return BLS_OK;
}

extern "C" void bls_ldl_destroy(bls_ldl *h) { delete h; }

extern "C" int bls_ldl_factor_f64(bls_ldl *h, const double *S, int64_t *zero_col, int64_t *negative, void *stream) {
    if (!h || (h->n && !S)) return fail(BLS_ERR_ARG, "bls_ldl_factor_f64: bad arguments");
    hipStream_t s = (hipStream_t)stream;
    const int64_t n = h->n;
    h->factored = false;
    if (zero_col) *zero_col = 0;
    if (negative) *negative = 0;
    if (n == 0) { h->factored = true; return BLS_OK; }
    k_bk_init<<<1, 1, 0, s>>>(h->st.ptr);
    k_bk_copy<<<std::min<int64_t>(grid_for(n * n, 256), 8192), 256, 0, s>>>(S, h->A.ptr, n);
    for (int64_t c = 0; c <= n; c++) {
        k_bk_pivot<<<1, BK_THREADS, 0, s>>>(h->A.ptr, n, h->st.ptr, h->w.ptr, h->ipiv.ptr, c);
        const int64_t T = n - c - 1;
        if (c < n && T > 0) {
            const int64_t nt = (T + BK_TILE - 1) / BK_TILE;
            k_bk_update<<<(unsigned)(nt * (nt + 1) / 2), 256, 0, s>>>(h->A.ptr, n, h->st.ptr, h->w.ptr, c);
        }
    }
    k_bk_diag<<<grid_for(n, 256), 256, 0, s>>>(h->A.ptr, n, h->w.ptr);
    BLS_HIP(hipGetLastError());
    BKState stv;
    std::vector<int64_t> ipiv((size_t)n);
    std::vector<double> d((size_t)(2 * n));
    BLS_HIP(hipMemcpyAsync(&stv, h->st.ptr, sizeof(BKState), hipMemcpyDeviceToHost, s));
    BLS_HIP(hipMemcpyAsync(ipiv.data(), h->ipiv.ptr, sizeof(int64_t) * n, hipMemcpyDeviceToHost, s));
    BLS_HIP(hipMemcpyAsync(d.data(), h->w.ptr, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, s));
    BLS_HIP(hipStreamSynchronize(s));
    if (stv.info) {
        if (zero_col) *zero_col = stv.info;
        return BLS_FAILED;
    }
    // inertia of D (that of S, by Sylvester's law): a 1 x 1 pivot counts if negative, a 2 x 2 block by its
    // determinant and trace
    int64_t neg = 0;
    for (int64_t k = 0; k < n;) {
        if (ipiv[k] > 0) {
            neg += d[k] < 0.0;
            k += 1;
        } else {
            const double a = d[k], b = d[n + k], cc = d[k + 1];
            const double det = a * cc - b * b, tr = a + cc;
            neg += det < 0.0 ? 1 : (tr < 0.0 ? 2 : 0);
            k += 2;
        }
    }
    if (negative) *negative = neg;
    h->factored = true;
    return BLS_OK;
}

extern "C" int bls_ldl_solve_f64(bls_ldl *h, const double *g, double *x, void *stream) {
    if (!h || (h->n && (!g || !x))) return fail(BLS_ERR_ARG, "bls_ldl_solve_f64: bad arguments");
    if (!h->factored) return fail(BLS_ERR_STATE, "bls_ldl_solve_f64: no factorization (bls_ldl_factor_f64 failed or not run)");
    if (h->n == 0) return BLS_OK;
    hipStream_t s = (hipStream_t)stream;
    if (x != g) BLS_HIP(hipMemcpyAsync(x, g, sizeof(double) * h->n, hipMemcpyDeviceToDevice, s));
    k_bk_solve<<<1, BK_THREADS, 0, s>>>(h->A.ptr, h->n, h->ipiv.ptr, x);
    BLS_HIP(hipGetLastError());
    return BLS_OK;
}

extern "C" int bls_ldl_factor_copy_f64(bls_ldl *h, double *A, int64_t *ipiv, void *stream) {
    if (!h || (h->n && (!A || !ipiv))) return fail(BLS_ERR_ARG, "bls_ldl_factor_copy_f64: bad arguments");
    if (h->n == 0) return BLS_OK;
    hipStream_t s = (hipStream_t)stream;
    BLS_HIP(hipMemcpyAsync(A, h->A.ptr, sizeof(double) * h->n * h->n, hipMemcpyDeviceToDevice, s));
    BLS_HIP(hipMemcpyAsync(ipiv, h->ipiv.ptr, sizeof(int64_t) * h->n, hipMemcpyDeviceToDevice, s));
    return BLS_OK;
}
