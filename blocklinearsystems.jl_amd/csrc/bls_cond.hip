// The conditioning scale of an elimination on gfx950 (src/conditioning.jl:24-87) behind the C ABI of
// include/bls_amd.h: prepare_scale! over the device arrays s (the scale over all n variables) and d (the diagonal it
// is computed from), with EliminationConditioning's stale flag and recomputation count held by the handle.
//
// The scale of one variable is the power of two nearest, on a log scale, to 1 / (1e-12 + sqrt|d_j + lambda sc_j|)
// (non-finite: 1; at least floatmin of the system's element type). With a fresh fill (stale) every s_j is computed;
// otherwise the scale is recomputed, all of it, only if some variable's |d_j + lambda sc_j| s_j^2 has left
// [2^-16, 2^16] and its own scale would change -- decided on the device (per-workgroup flags, then one workgroup),
// so the host never waits. Every operation is IEEE-rounded and uncontracted (sqrt and the division are correctly
// rounded; frexp / ldexp exact), so s is bitwise that of the Float64 rule on the same d.
#include <algorithm>
#include <cmath>

#include "bls_internal.h"

using namespace bls;

namespace {
constexpr int COND_WG = 256;
constexpr int64_t COND_MAXWG = 2048;   // workgroups of the drift test (grid-stride beyond)

#pragma clang fp contract(off)
// conditioning_scale (src/conditioning.jl:29-35)
__device__ inline double cond_scale(double d, double fmin) {
    const double dd = fabs(d);
    if (!isfinite(dd)) return 1.0;
    int e;
    const double f = frexp(1.0 / (1e-12 + sqrt(dd)), &e);
    const double s = f >= 0x1.6a09e667f3bcdp-1 ? ldexp(1.0, e) : ldexp(1.0, e - 1);   // sqrt(0.5), rounded
    return s > fmin ? s : fmin;
}

__device__ inline double damped(const double *d, const double *sc, double lambda, int64_t j) {
    return d[j] + lambda * (sc ? sc[j] : 1.0);
}

// s = the scale of every variable; with `flag` set, only if *flag is nonzero (the drift test's verdict)
__global__ __launch_bounds__(COND_WG) void k_cond_compute(int64_t n, const double *__restrict__ d,
                                                          const double *__restrict__ sc, double lambda, double fmin,
                                                          const int32_t *__restrict__ flag, double *__restrict__ s) {
    if (flag && !*flag) return;
    for (int64_t j = (int64_t)blockIdx.x * COND_WG + threadIdx.x; j < n; j += (int64_t)gridDim.x * COND_WG)
        s[j] = cond_scale(damped(d, sc, lambda, j), fmin);
}

// the drift test (src/conditioning.jl:69-77): flags[g] = 1 if workgroup g finds a variable whose damped diagonal
// has left the band of its scale and whose scale would change
__global__ __launch_bounds__(COND_WG) void k_cond_drift(int64_t n, const double *__restrict__ d,
                                                        const double *__restrict__ sc, double lambda, double fmin,
                                                        const double *__restrict__ s, int32_t *__restrict__ flags) {
    int drifted = 0;
    for (int64_t j = (int64_t)blockIdx.x * COND_WG + threadIdx.x; j < n; j += (int64_t)gridDim.x * COND_WG) {
        const double dj = damped(d, sc, lambda, j), sj = s[j];
        const double v = fabs(dj) * (sj * sj);
        if (!(v >= 0x1p-16 && v <= 0x1p16) && cond_scale(dj, fmin) != sj) drifted = 1;
    }
    drifted = __syncthreads_or(drifted);
    if (threadIdx.x == 0) flags[blockIdx.x] = drifted;
}

// flags[nwg] = any of flags[0 .. nwg); a drift counts one recomputation
__global__ __launch_bounds__(COND_WG) void k_cond_any(int nwg, int32_t *__restrict__ flags, int64_t *__restrict__ nrescale) {
    int any = 0;
    for (int g = threadIdx.x; g < nwg; g += COND_WG) any |= flags[g];
    any = __syncthreads_or(any);
    if (threadIdx.x == 0) {
        flags[nwg] = any;
        if (any) nrescale[0] += 1;
    }
}

__global__ void k_cond_zero(int64_t *p) { p[0] = 0; }
__global__ void k_cond_flag_clear(int32_t *p) { p[0] = 0; }

// unscale! (src/conditioning.jl:90-95): x_j *= s_j; flags[g] = 1 if workgroup g's entries of the result hold a value
// that is not finite (all(isfinite, x), src/schur_linear_systems.jl:333-334)
__global__ __launch_bounds__(COND_WG) void k_unscale(int64_t n, double *__restrict__ x, const double *__restrict__ s,
                                                     int32_t *__restrict__ flags) {
    int bad = 0;
    for (int64_t j = (int64_t)blockIdx.x * COND_WG + threadIdx.x; j < n; j += (int64_t)gridDim.x * COND_WG) {
        const double v = s ? x[j] * s[j] : x[j];
        if (s) x[j] = v;
        bad |= !isfinite(v);
    }
    bad = __syncthreads_or(bad);
    if (threadIdx.x == 0) flags[blockIdx.x] = bad;
}

// out[0] = any of flags[0 .. nwg)
__global__ __launch_bounds__(COND_WG) void k_flags_any(int nwg, const int32_t *__restrict__ flags,
                                                       int32_t *__restrict__ out) {
    int any = 0;
    for (int g = threadIdx.x; g < nwg; g += COND_WG) any |= flags[g];
    any = __syncthreads_or(any);
    if (threadIdx.x == 0) out[0] = any;
}
}  // namespace

struct bls_cond {
    int64_t n = 0;
    bool stale = true;
    int nwg = 1;
    DevArray<int32_t> flags;      // per-workgroup drift flags, then the verdict
    DevArray<int64_t> nrescale;   // recomputations at a later damping (EliminationConditioning.nrescale)
};

extern "C" int bls_cond_create(int64_t n, bls_cond **out) {
    if (!out || n < 0) return fail(BLS_ERR_ARG, "bls_cond_create: bad arguments");
    *out = nullptr;
    auto c = new bls_cond();
    c->n = n;
    c->nwg = (int)std::max<int64_t>(1, std::min<int64_t>(COND_MAXWG, (n + COND_WG - 1) / COND_WG));
    int rc;
    if ((rc = c->flags.alloc((size_t)c->nwg + 1)) || (rc = c->nrescale.alloc(1))) {
        delete c;
        return rc;
    }
    k_cond_zero<<<1, 1>>>(c->nrescale.ptr);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        delete c;
        return fail(BLS_ERR_HIP, std::string("bls_cond_create: ") + hipGetErrorString(e));
    }
    *out = c;
    return BLS_OK;
}

extern "C" void bls_cond_destroy(bls_cond *c) { delete c; }

extern "C" int bls_cond_is_stale(const bls_cond *c) {
    if (!c) return fail(BLS_ERR_ARG, "null handle");
    return c->stale ? 1 : 0;
}

extern "C" int bls_cond_markstale(bls_cond *c) {
    if (!c) return fail(BLS_ERR_ARG, "null handle");
    c->stale = true;
    return BLS_OK;
}

extern "C" int bls_prepare_scale_f64(bls_cond *c, double *s, const double *d, const double *sc, double lambda,
                                     int elem_f32, void *stream) {
    if (!c) return fail(BLS_ERR_ARG, "null handle");
    if (c->n && (!s || !d)) return fail(BLS_ERR_ARG, "bls_prepare_scale_f64: s and d are required");
    hipStream_t st = (hipStream_t)stream;
    lambda = lambda < 0.0 ? 0.0 : lambda;   // max(lambda, 0.0); NaN stays NaN
    const double fmin = elem_f32 ? 0x1p-126 : 0x1p-1022;   // floatmin(Float32), floatmin(Float64)
    if (c->n) {
        if (c->stale) {
            k_cond_compute<<<(unsigned)c->nwg, COND_WG, 0, st>>>(c->n, d, sc, lambda, fmin, nullptr, s);
        } else {
            k_cond_drift<<<(unsigned)c->nwg, COND_WG, 0, st>>>(c->n, d, sc, lambda, fmin, s, c->flags.ptr);
            k_cond_any<<<1, COND_WG, 0, st>>>(c->nwg, c->flags.ptr, c->nrescale.ptr);
            k_cond_compute<<<(unsigned)c->nwg, COND_WG, 0, st>>>(c->n, d, sc, lambda, fmin, c->flags.ptr + c->nwg, s);
        }
        BLS_HIP(hipGetLastError());
    }
    c->stale = false;
    return BLS_OK;
}

extern "C" int bls_cond_nrescale(bls_cond *c, int64_t *out, void *stream) {
    if (!c || !out) return fail(BLS_ERR_ARG, "bls_cond_nrescale: bad arguments");
    BLS_HIP(hipStreamSynchronize((hipStream_t)stream));
    BLS_HIP(hipMemcpy(out, c->nrescale.ptr, sizeof(int64_t), hipMemcpyDeviceToHost));
    return BLS_OK;
}

extern "C" int bls_unscale_f64(bls_cond *c, double *x, const double *s, int32_t *nonfinite, void *stream) {
    if (!c || !nonfinite || (c->n && !x)) return fail(BLS_ERR_ARG, "bls_unscale_f64: bad arguments");
    hipStream_t st = (hipStream_t)stream;
    if (c->n) k_unscale<<<(unsigned)c->nwg, COND_WG, 0, st>>>(c->n, x, s, c->flags.ptr);
    else k_cond_flag_clear<<<1, 1, 0, st>>>(c->flags.ptr);
    k_flags_any<<<1, COND_WG, 0, st>>>(c->n ? c->nwg : 1, c->flags.ptr, nonfinite);
    BLS_HIP(hipGetLastError());
    return BLS_OK;
}
