// Internal helpers shared by the HIP translation units of libblsamd.so (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>

#include "bls_amd.h"

namespace bls {

void set_error(const std::string &msg);
int fail(int code, const std::string &msg);

#define BLS_HIP(expr)                                                                                          \
    do {                                                                                                       \
        hipError_t _e = (expr);                                                                                \
        if (_e != hipSuccess)                                                                                  \
            return ::bls::fail(BLS_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e));              \
    } while (0)

// A device array owned by a handle.
template <typename T> struct DevArray {
    T *ptr = nullptr;
    size_t n = 0;
    DevArray() = default;
    DevArray(const DevArray &) = delete;
    DevArray &operator=(const DevArray &) = delete;
    ~DevArray() { reset(); }
    void reset() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        n = 0;
    }
    int upload(const std::vector<T> &v) {
        reset();
        n = v.size();
        if (n == 0) return BLS_OK;
        hipError_t e = hipMalloc((void **)&ptr, n * sizeof(T));
        if (e != hipSuccess) return fail(BLS_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
        e = hipMemcpy(ptr, v.data(), n * sizeof(T), hipMemcpyHostToDevice);
        if (e != hipSuccess) return fail(BLS_ERR_HIP, std::string("hipMemcpy: ") + hipGetErrorString(e));
        return BLS_OK;
    }
    int alloc(size_t count) {
        reset();
        n = count;
        if (n == 0) return BLS_OK;
        hipError_t e = hipMalloc((void **)&ptr, n * sizeof(T));
        if (e != hipSuccess) return fail(BLS_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
        return BLS_OK;
    }
};

constexpr int64_t NO_FAILURE = INT64_MAX;

// Blocks per grid for n items at `threads` per block.
inline unsigned grid_for(int64_t n, int threads) { return (unsigned)((n + threads - 1) / threads); }

// XCD-aware remap of the hardware block index: the dispatcher places block i on XCD i % 8, so consecutive
// logical items (which share landmark data) are given to blocks of the same XCD, sharing its L2.
__device__ inline int64_t xcd_remap(int64_t bid, int64_t nblocks) {
    const int64_t per = (nblocks + 7) / 8;
    return (bid % 8) * per + bid / 8;
}

// Event-based timing of the kernels of one call (enabled per handle).
struct Timing {
    bool enabled = false;
    hipEvent_t ev[8] = {};
    int nev = 0;
    float ms[8] = {};
    ~Timing() {
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
    }
    int enable() {
        if (!enabled)
            for (auto &e : ev)
                if (!e && hipEventCreate(&e) != hipSuccess) return fail(BLS_ERR_HIP, "hipEventCreate");
        enabled = true;
        return BLS_OK;
    }
    void mark(int i, hipStream_t s) {
        if (enabled) (void)hipEventRecord(ev[i], s);
    }
};

}  // namespace bls
