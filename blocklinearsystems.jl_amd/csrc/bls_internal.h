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

// Phase 2 on 16 x 16 tiles of S (bls_tiles.hip). A unit is a run of nt <= 4 tiles of tile row ti, tile columns
// tj .. tj + nt - 1; its landmarks are the first cnt[0] entries of the row's list (elo), cnt[t] of them reaching tile
// t, whose couplings lie in [cbase, cbase + spanc).
struct TileUnit {
    int32_t ti, tj, nt, rb0;   // rb0: first retained block of the tile row's rows
    int32_t cnt[4];
    int32_t spanc, pad;        // spanc: couplings the row's landmarks reach
    int64_t elo, cbase;
};
struct TileEntry {
    int32_t e;      // landmark (0-based)
    int32_t cw;     // its first coupling at or after block rb0, relative to cbase
    int32_t qfql;   // that coupling's block - rb0 (low 16 bits), its last coupling's block - rb0 (high 16 bits)
    int32_t last;   // the last tile column it reaches, relative to ti
};
struct TilePlan {
    DevArray<TileUnit> units;
    DevArray<TileEntry> entries;
    DevArray<int32_t> tabs;   // per unit and tile: offsets in S's columns of (row block q, column block p), 4 x 4, -1 none
    int64_t nunits = 0, nentries = 0, nmfma = 0;
    int de = 0, dr = 0;
};
// What the plan is built from (0-based host arrays of the handle and its registered sparse layout).
struct TileStructure {
    int64_t nr, nrb, neb;
    const std::vector<int64_t> &e_dims, &r_dims, &eoffsets, &boffsets_e, &boffsets_r, &coffsets, &eptr, &eret;
    const std::vector<int64_t> *bcp, *brv, *bro;
};
int tiles_build(TilePlan &P, const TileStructure &T, std::string *why);
// variant 0: Float64 (E_inv, C, S double); 1: Float32 (float E_inv, C, S); 2: mixed (double E_inv, float C, double S)
int tiles_launch(const TilePlan &P, int variant, const void *E_inv, const double *z, const void *C, const double *b_r,
                 const double *s_r, void *S, double *g, const int64_t *S_colptr, int64_t nr,
                 const unsigned long long *flag, unsigned *fin_counter, unsigned long long *fin_flag, int64_t *fin_out,
                 hipStream_t s);

// Blocks per grid for n items at `threads` per block.
inline unsigned grid_for(int64_t n, int threads) { return (unsigned)((n + threads - 1) / threads); }

// XCD-aware remap of the hardware block index, a bijection of [0, nblocks): the dispatcher places block i on XCD
// i % 8, so XCD x is given the contiguous logical range of its blocks (the first nblocks % 8 XCDs one block more),
// and neighbouring logical blocks, which read neighbouring data, share that XCD's L2.
__device__ inline int64_t xcd_remap(int64_t bid, int64_t nblocks) {
    const int64_t q = nblocks / 8, r = nblocks % 8, x = bid % 8, slot = bid / 8;
    return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + slot;
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
