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

// Pooled bytes beyond which released blocks are freed.
constexpr size_t kPoolLimit = size_t(16) << 30;

int BlockPool::ensure(DevBuf& b, size_t bytes, hipStream_t stream) {
    (void)stream;  // every block here is used on the pool's stream only
    if (bytes <= b.cap) return FS_OK;
    const size_t want = align_up(bytes, size_t(2) << 20);
    {
        std::lock_guard<std::mutex> lk(mu);
        if (b.ptr) {
            blocks.emplace(b.cap, b.ptr);
            held += b.cap;
            b.ptr = nullptr;
            b.cap = 0;
        }
        auto it = blocks.lower_bound(want);
        if (it != blocks.end() && it->first <= 2 * want) {
            b.ptr = it->second;
            b.cap = it->first;
            held -= it->first;
            blocks.erase(it);
            return FS_OK;
        }
    }
    for (int attempt = 0; attempt < 2; ++attempt) {
        if (hipMalloc(&b.ptr, want) == hipSuccess) {
            b.cap = want;
            return FS_OK;
        }
        b.ptr = nullptr;
        (void)hipGetLastError();
        if (attempt == 0) {
            // give the pooled blocks back and try once more
            FS_HIP_CHECK(hipStreamSynchronize(stream));
            clear();
        }
    }
    return fail(FS_E_NOMEM, "hipMalloc of %zu bytes failed", want);
}

void BlockPool::put(DevBuf& b) {
    if (!b.ptr) return;
    std::vector<void*> drop;
    {
        std::lock_guard<std::mutex> lk(mu);
        blocks.emplace(b.cap, b.ptr);
        held += b.cap;
        // beyond the limit the largest blocks are freed
        while (held > kPoolLimit && !blocks.empty()) {
            auto it = std::prev(blocks.end());
            held -= it->first;
            drop.push_back(it->second);
            blocks.erase(it);
        }
    }
    b.ptr = nullptr;
    b.cap = 0;
    for (void* q : drop) (void)hipFree(q);
}

void BlockPool::clear() {
    std::lock_guard<std::mutex> lk(mu);
    for (auto& kv : blocks) (void)hipFree(kv.second);
    blocks.clear();
    held = 0;
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
    ctx->pool.clear();
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
