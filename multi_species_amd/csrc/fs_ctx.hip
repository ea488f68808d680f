// Context, error reporting and device-buffer management of libfarrsight_hip.
#include <cstdarg>
#include <cstdio>
#include <cstring>

#include "fs_field_batch.h"
#include "fs_internal.h"

namespace fs {

namespace {
thread_local char g_last_error[512] = "";
}

int fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(g_last_error, sizeof g_last_error, fmt, ap);
    va_end(ap);
    return code;
}

int DevBuf::ensure(size_t bytes, hipStream_t stream) {
    if (bytes <= cap) return FS_OK;
    if (ptr) {
        FS_HIP_CHECK(hipStreamSynchronize(stream));
        FS_HIP_CHECK(hipFree(ptr));
        ptr = nullptr;
        cap = 0;
    }
    const size_t want = align_up(bytes, size_t(2) << 20);
    if (hipMalloc(&ptr, want) != hipSuccess) {
        ptr = nullptr;
        (void)hipGetLastError();
        return fail(FS_E_NOMEM, "hipMalloc of %zu bytes failed", want);
    }
    cap = want;
    return FS_OK;
}

void DevBuf::release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    cap = 0;
}

}  // namespace fs

extern "C" {

const char* fs_last_error(void) { return fs::g_last_error; }

int fs_ctx_create(fs_ctx** out, int device) {
    if (!out) return fs::fail(FS_E_INVALID, "fs_ctx_create: null out pointer");
    *out = nullptr;
    int n = 0;
    FS_HIP_CHECK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n)
        return fs::fail(FS_E_INVALID, "fs_ctx_create: device %d of %d", device, n);
    FS_HIP_CHECK(hipSetDevice(device));
    auto* c = new fs_ctx();
    c->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess && prop.multiProcessorCount > 0)
        c->num_cus = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fs::fail(FS_E_HIP, "hipStreamCreateWithFlags failed");
    }
    // the allocation-free entry point's work space on this device; a failure
    // is kept for that entry point to report and leaves the others usable
    if (fs::noalloc_init(device) != FS_OK) (void)hipGetLastError();
    *out = c;
    return FS_OK;
}

int fs_ctx_destroy(fs_ctx* ctx) {
    if (!ctx) return FS_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    ctx->scratch.release();
    ctx->io.release();
    ctx->aux.release();
    ctx->tree.release();
    if (ctx->api_batch) {
        ctx->api_batch->release();
        delete ctx->api_batch;
        ctx->api_batch = nullptr;
    }
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return FS_OK;
}

void* fs_ctx_stream(fs_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int fs_ctx_synchronize(fs_ctx* ctx) {
    if (!ctx) return fs::fail(FS_E_INVALID, "null context");
    FS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return FS_OK;
}

}  // extern "C"
