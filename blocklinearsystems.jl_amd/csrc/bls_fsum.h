// A fixed-order sum of n doubles on the device, shared by the translation units that reduce per-item partials to
// one scalar (dot(x, H, y), ||J v||^2). Included inside each unit's anonymous namespace, after bls_internal.h.
// The partials are added in one fixed order, in two levels: workgroup g sums chunk g (thread t takes elements
// t, t + 256, ... of it, then a fixed tree) into a chunk sum; the same kernel with one workgroup then sums the chunk
// sums. No atomics: the result depends only on n and the values, never on scheduling.
#pragma once

constexpr int FSUM_CHUNKS = 1024;   // at most this many chunk sums: the scratch `sums` holds FSUM_CHUNKS doubles

__global__ __launch_bounds__(256) void k_fsum(int64_t n, int64_t chunk, const double *__restrict__ part,
                                              double *__restrict__ out) {
    __shared__ double red[256];
    const int64_t lo = (int64_t)blockIdx.x * chunk, hi = std::min<int64_t>(n, lo + chunk);
    double acc = 0.0;
    for (int64_t i = lo + threadIdx.x; i < hi; i += 256) acc += part[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o >= 1; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

// out[0] = the sum of part[0 .. n) (0 for n = 0); sums: FSUM_CHUNKS doubles of scratch
inline void fsum_launch(int64_t n, const double *part, double *sums, double *out, hipStream_t s) {
    const int64_t chunk = std::max<int64_t>(256, (n + FSUM_CHUNKS - 1) / FSUM_CHUNKS);
    const int64_t nchunk = (n + chunk - 1) / chunk;
    if (nchunk > 1) {
        k_fsum<<<(unsigned)nchunk, 256, 0, s>>>(n, chunk, part, sums);
        k_fsum<<<1, 256, 0, s>>>(nchunk, nchunk, sums, out);
    } else {
        k_fsum<<<1, 256, 0, s>>>(n, std::max<int64_t>(n, 1), part, out);
    }
}
