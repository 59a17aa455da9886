// Block-diagonal kernels on a flat array of square blocks (factor_diagonal_blocks! / solve_diagonal_blocks!,
// src/diagonal_blocks.jl:6-61): block b, of dimension d_b, column-major at the sum of the d^2 before it, as
// reduced_diagonal_blocks! writes them. One lane per block; the Cholesky runs in the element type T, without
// contraction, so a block is accepted or rejected as the reference's loop on T decides.
#include <cmath>

#include "bls_internal.h"

using namespace bls;

namespace {

#pragma clang fp contract(off)
template <class T>
__global__ void k_factor_blocks(int64_t nb, const int32_t *dims, const int64_t *off, T *D, unsigned long long *flag) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const int d = dims[b];
    T *B = D + off[b];
    for (int j = 0; j < d; j++) {
        T s = B[j * d + j];
        for (int k = 0; k < j; k++) {
            const T l = B[k * d + j];
            s -= l * l;
        }
        if (!(s > (T)0 && isfinite(s))) {
            atomicMin(flag, (unsigned long long)(b + 1));
            return;
        }
        const T ljj = sqrt(s);
        B[j * d + j] = ljj;
        for (int i = j + 1; i < d; i++) {
            T t = B[j * d + i];
            for (int k = 0; k < j; k++) t -= B[k * d + i] * B[k * d + j];
            B[j * d + i] = t / ljj;
        }
    }
}
#pragma clang fp contract(fast)

// y = (L L')^-1 x per block, in Float64 from the factors in T; y may be x
template <class T>
__global__ void k_solve_blocks(int64_t nb, const int32_t *dims, const int64_t *off, const int64_t *pos, const T *D,
                               const double *x, double *y) {
    const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const int d = dims[b];
    const T *L = D + off[b];
    double *yb = y + pos[b];
    const double *xb = x + pos[b];
    double v[16];
    for (int i = 0; i < d; i++) v[i] = xb[i];
    for (int j = 0; j < d; j++) {
        const double yj = v[j] / (double)L[j * d + j];
        v[j] = yj;
        for (int i = j + 1; i < d; i++) v[i] -= (double)L[j * d + i] * yj;
    }
    for (int j = d - 1; j >= 0; j--) {
        double t = v[j];
        for (int i = j + 1; i < d; i++) t -= (double)L[j * d + i] * v[i];
        v[j] = t / (double)L[j * d + j];
    }
    for (int i = 0; i < d; i++) yb[i] = v[i];
}

__global__ void k_diag_flag(unsigned long long *p, unsigned long long v) { *p = v; }

}  // namespace

struct bls_diag {
    int64_t nb = 0;
    DevArray<int32_t> d_dims;
    DevArray<int64_t> d_off, d_pos;
    DevArray<unsigned long long> d_flag;
};

extern "C" int bls_diag_create(int64_t nblocks, const int64_t *dims, bls_diag **out) {
    if (!out || nblocks < 0 || (nblocks && !dims)) return fail(BLS_ERR_ARG, "bls_diag_create: bad arguments");
    *out = nullptr;
    std::vector<int32_t> d(nblocks);
    std::vector<int64_t> off(nblocks), pos(nblocks);
    int64_t o = 0, p = 0;
    for (int64_t b = 0; b < nblocks; b++) {
        if (dims[b] < 1 || dims[b] > 16) return fail(BLS_ERR_UNSUPPORTED, "diagonal block dimension outside 1..16");
        d[b] = (int32_t)dims[b];
        off[b] = o;
        pos[b] = p;
        o += dims[b] * dims[b];
        p += dims[b];
    }
    auto *h = new bls_diag();
    h->nb = nblocks;
    int rc;
    if ((rc = h->d_dims.upload(d)) || (rc = h->d_off.upload(off)) || (rc = h->d_pos.upload(pos)) ||
        (rc = h->d_flag.alloc(1))) {
        delete h;
        return rc;
    }
    *out = h;
    return BLS_OK;
}

extern "C" void bls_diag_destroy(bls_diag *h) { delete h; }

namespace {
template <class T> int factor_blocks(bls_diag *h, T *D, int64_t *failed_block, hipStream_t s) {
    if (!h || (h->nb && !D)) return fail(BLS_ERR_ARG, "factor_diagonal_blocks: bad arguments");
    if (!h->nb) {
        if (failed_block) *failed_block = 0;
        return BLS_OK;
    }
    k_diag_flag<<<1, 1, 0, s>>>(h->d_flag.ptr, (unsigned long long)NO_FAILURE);
    k_factor_blocks<T><<<grid_for(h->nb, 128), 128, 0, s>>>(h->nb, h->d_dims.ptr, h->d_off.ptr, D, h->d_flag.ptr);
    BLS_HIP(hipGetLastError());
    unsigned long long f = 0;
    BLS_HIP(hipMemcpyAsync(&f, h->d_flag.ptr, sizeof(f), hipMemcpyDeviceToHost, s));
    BLS_HIP(hipStreamSynchronize(s));
    const int64_t fb = f == (unsigned long long)NO_FAILURE ? 0 : (int64_t)f;
    if (failed_block) *failed_block = fb;
    return fb ? BLS_FAILED : BLS_OK;
}

template <class T> int solve_blocks(bls_diag *h, double *y, const T *D, const double *x, hipStream_t s) {
    if (!h || (h->nb && (!y || !D || !x))) return fail(BLS_ERR_ARG, "solve_diagonal_blocks: bad arguments");
    if (h->nb)
        k_solve_blocks<T><<<grid_for(h->nb, 128), 128, 0, s>>>(h->nb, h->d_dims.ptr, h->d_off.ptr, h->d_pos.ptr, D, x, y);
    BLS_HIP(hipGetLastError());
    return BLS_OK;
}
}  // namespace

extern "C" int bls_factor_diagonal_blocks_f64(bls_diag *h, double *D, int64_t *failed_block, void *stream) {
    return factor_blocks<double>(h, D, failed_block, (hipStream_t)stream);
}
extern "C" int bls_factor_diagonal_blocks_f32(bls_diag *h, float *D, int64_t *failed_block, void *stream) {
    return factor_blocks<float>(h, D, failed_block, (hipStream_t)stream);
}
extern "C" int bls_solve_diagonal_blocks_f64(bls_diag *h, double *y, const double *D, const double *x, void *stream) {
    return solve_blocks<double>(h, y, D, x, (hipStream_t)stream);
}
extern "C" int bls_solve_diagonal_blocks_f32(bls_diag *h, double *y, const float *D, const double *x, void *stream) {
    return solve_blocks<float>(h, y, D, x, (hipStream_t)stream);
}
