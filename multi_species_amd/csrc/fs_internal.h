// Internal declarations shared by the HIP translation units of
// libfarrsight_hip.so. Nothing here is part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <map>
#include <mutex>
#include <utility>
#include <vector>

#include "../../include/farrsight_hip.h"

namespace fs {

class FieldBatch;

// Records a message for fs_last_error() on this thread and returns `code`.
int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));

#define FS_HIP_CHECK(call)                                                        \
    do {                                                                          \
        hipError_t fs_e_ = (call);                                                \
        if (fs_e_ != hipSuccess)                                                  \
            return ::fs::fail(FS_E_HIP, "%s failed: %s (%s:%d)", #call,           \
                              hipGetErrorString(fs_e_), __FILE__, __LINE__);      \
    } while (0)

#define FS_TRY(expr)                 \
    do {                             \
        int fs_rc_ = (expr);         \
        if (fs_rc_ != FS_OK) return fs_rc_; \
    } while (0)

// Grow-only device allocation owned by a context. Growing first waits for the
// owning stream so no enqueued kernel still reads the old block.
struct DevBuf {
    void* ptr = nullptr;
    size_t cap = 0;
    int ensure(size_t bytes, hipStream_t stream);
    void release();
    template <class T>
    T* as(size_t byte_offset = 0) const {
        return reinterpret_cast<T*>(static_cast<char*>(ptr) + byte_offset);
    }
};

inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Device blocks released by buffers used on one stream only, kept for reuse
// by buffers used on the same stream (every use of a released block is
// enqueued before any use of its next owner): decks joining and leaving a
// batch, and layouts that resize it, then allocate and free nothing, and no
// hipFree synchronizes the device. Thread-safe.
struct BlockPool {
    std::mutex mu;
    std::multimap<size_t, void*> blocks;  // capacity -> block
    size_t held = 0;
    // Grows b to `bytes` like DevBuf::ensure (contents not kept), its old
    // block going to the pool and a pooled block of at most twice the need
    // reused when there is one.
    int ensure(DevBuf& b, size_t bytes, hipStream_t stream);
    // b's block into the pool (b emptied).
    void put(DevBuf& b);
    // Frees every pooled block (the stream's work on them must be done).
    void clear();
};

// Field-solver launch state kept per context.
struct FieldTuning {
    int targets_per_thread = 0;  // 0 = automatic
    int source_splits = 0;       // 0 = automatic
    // FP64 direct sum with targets = sources: 0 = each pair once (K odd)
    // from a size up, 1 = never, 2 = at every size
    int symmetric = 0;
    // bound on the pair-once form's partials in bytes (0 = the default,
    // 4 GiB); beyond it the evaluation runs in passes (same result)
    int64_t sym_budget = 0;
};

// FP64 direct sum with targets = sources (one array, nt == ns): whether the
// pair-once form (fs_sym.h) runs, by the size and the context's tuning.
bool use_sym(int64_t nt, int64_t ns, bool same, const FieldTuning& tun);

// Periodic direct sum on device arrays (FP64, compensated). `scratch` must
// hold field_scratch_bytes(nt, ns, same, ...) bytes, same = the targets are
// the sources (one array, nt == ns). Enqueued on `stream`.
size_t field_scratch_bytes(int64_t nt, int64_t ns, bool same, int num_cus,
                           const FieldTuning& tun);
int field_direct_f64(const double* d_tgt, int64_t nt, const double* d_src,
                     const double* d_w, int64_t ns, double L, double eps,
                     double* d_es, void* scratch, int num_cus,
                     const FieldTuning& tun, hipStream_t stream,
                     int64_t* pair_evals = nullptr);

// Stable radix sort by (segment of the original index, top 24 key bits)
// (fs_radix.hip). kin must not alias kout.
size_t radix_sort_bytes(int n);
int radix_sort_seg(const uint64_t* kin, int n, const int32_t* seg_of, int nseg, uint64_t* kout,
                   int32_t* iout, void* tmp, hipStream_t stream);
// Stable sort by (segment of the original index, all 64 key bits).
int radix_sort_seg_full(const uint64_t* kin, int n, const int32_t* seg_of, int nseg,
                        uint64_t* kout, int32_t* iout, void* tmp, hipStream_t stream);
// Stable sort by all 64 key bits (one segment).
int radix_sort_full(const uint64_t* kin, int n, uint64_t* kout, int32_t* iout, void* tmp,
                    hipStream_t stream);
// Stable sort by the top key_bytes bytes of the keys (one segment); iout is
// the permutation (original index of each sorted key).
int radix_sort_top(const uint64_t* kin, int n, int key_bytes, uint64_t* kout, int32_t* iout,
                   void* tmp, hipStream_t stream);

// Per-species diagnostics on device arrays (fs_diag.hip): FS_DIAG_COUNT
// values into out (host); work holds diag_work_bytes(n) bytes. Synchronizes
// `stream`.
namespace diag {
size_t work_bytes(int n);
int species(const double* x, const double* p, const double* f, const double* w,
            const double* e, int n, double q, double qm, double L, void* work, double* out,
            hipStream_t stream);
}  // namespace diag

// The reference's on-disk state (fs_io.cpp): one species' files, and the
// deck's reduced set. Host arrays.
namespace io {
int write_species(const char* sim_dir, const char* name, int iter, int64_t n, const double* xs,
                  const double* ps, const double* fs, const double* q_ws, const double* es,
                  int64_t n_panels, const int32_t* rec);
int read_species(const char* sim_dir, const char* name, int iter, std::vector<double>& xs,
                 std::vector<double>& ps, std::vector<double>& fs, std::vector<int32_t>& rec);
int write_reduced(const char* sim_dir, int iter, int64_t u, const double* rx, const double* re);
}  // namespace io

// The drive amplitude a(t) of ExternalTanh / ExternalSine / ExternalLogistic
// (fs_open.hip; the reference's expressions and constants).
double external_amplitude(int type, double t, double Am, double k);

// Fixed device work space of fs_emq_periodic_f64_noalloc (per device).
constexpr size_t kNoAllocBytes = size_t(256) << 20;
// Takes the current device's address of that work space (once per device;
// fs_ctx_create calls it with the device current).
int noalloc_init(int device);
void* noalloc_ptr(int device);

// Work space of the device treecode: the build (sized by the point count
// and the node capacity) and the interaction lists with the evaluation's
// arrays (sized after the build). node_cap persists between calls and grows
// when a build needs more nodes.
struct TreeWork {
    DevBuf base, lists;
    int node_cap = 0;
    void release() {
        base.release();
        lists.release();
    }
};

// Treecode on device arrays: sort, tree, lists and sums on the device, one
// host synchronization (the lists' sizes). evals (optional): far and near
// kernel evaluations.
int treecode_f64(const double* d_x, const double* d_w, int n, double L, double eps, double mac,
                 int degree, int max_source, double* d_es, TreeWork& work, hipStream_t stream,
                 int64_t* evals);

}  // namespace fs

struct fs_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int num_cus = 256;
    fs::DevBuf scratch;   // kernel work space
    fs::DevBuf io;        // staging for the host-pointer entry points
    fs::DevBuf aux;       // concatenated points of the mixed entry points
    fs::TreeWork tree;    // the treecode entry points' work space
    fs::FieldTuning tuning;
    fs::BlockPool pool;   // blocks of the simulations' buffers (all on `stream`)
    fs::FieldBatch* api_batch = nullptr;  // plan of the last mixed entry-point call
    double mixed_target = 0;              // > 0: mixed entry points hold this error bound
    int64_t mixed_min_pairs = FS_MIXED_MIN_PAIRS_DEFAULT;  // smaller calls: FP64 sum
    int64_t f64_pair_evals = 0;           // kernel evaluations of the last FP64 direct sum
                                          // (periodic or open)
    // autotuned launch shapes (tpt, workgroups per CU) of the mixed entry
    // points, by problem class
    std::map<uint64_t, std::pair<int, int>> shape_cache;
    // the same for the batches of the step loop and the server (by
    // precision, total points in half-octaves and deck count in octaves)
    std::map<uint64_t, std::pair<int, int>> batch_shape_cache;
    std::mutex cache_mu;  // the serving loop's worker thread shares the caches
};
