// Error state, version and device-memory helpers of the C ABI.
#include "bls_internal.h"

namespace bls {
static thread_local std::string g_error;
void set_error(const std::string &msg) { g_error = msg; }
int fail(int code, const std::string &msg) {
    g_error = msg;
    return code;
}
}  // namespace bls

using namespace bls;

extern "C" const char *bls_last_error(void) { return g_error.c_str(); }
extern "C" const char *bls_version(void) { return "blsamd 0.1 (gfx950)"; }

extern "C" int bls_alloc(void **ptr, int64_t bytes) {
    if (!ptr || bytes < 0) return fail(BLS_ERR_ARG, "bls_alloc: bad arguments");
    *ptr = nullptr;
    if (bytes == 0) return BLS_OK;
    BLS_HIP(hipMalloc(ptr, (size_t)bytes));
    return BLS_OK;
}

extern "C" int bls_free(void *ptr) {
    if (ptr) BLS_HIP(hipFree(ptr));
    return BLS_OK;
}

extern "C" int bls_upload(void *dst, const void *src, int64_t bytes, void *stream) {
    if (bytes < 0) return fail(BLS_ERR_ARG, "bls_upload: negative size");
    if (bytes == 0) return BLS_OK;
    BLS_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyHostToDevice, (hipStream_t)stream));
    BLS_HIP(hipStreamSynchronize((hipStream_t)stream));
    return BLS_OK;
}

extern "C" int bls_download(void *dst, const void *src, int64_t bytes, void *stream) {
    if (bytes < 0) return fail(BLS_ERR_ARG, "bls_download: negative size");
    if (bytes == 0) return BLS_OK;
    BLS_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToHost, (hipStream_t)stream));
    BLS_HIP(hipStreamSynchronize((hipStream_t)stream));
    return BLS_OK;
}

extern "C" int bls_stream_synchronize(void *stream) {
    BLS_HIP(hipStreamSynchronize((hipStream_t)stream));
    return BLS_OK;
}
