// Periodic regularized E-field, direct sum, FP64 (gfx950).
//
// Replaces E_MQ_DirectSum::operator() (reference src/FieldStructure.cpp:
// 106-156). With z = (x_i - x_j)/L wrapped to [-1/2, 1/2] and delta = eps/L
// the reference term
//     w_j * (-z L) * (1/sqrt(L^2 + 4 eps^2) - 1/sqrt(4 z^2 L^2 + 4 eps^2))
// equals (w_j/2) * z * (1/sqrt(z^2 + delta^2) - C), C = 2/sqrt(1 + 4 delta^2).
// The kernel accumulates w_j z (1/sqrt(z^2 + delta^2) - C) and halves the sum
// at the end (an exact scaling). Per pair: d = x_i - x_j (exact for nearby
// points, like the reference), t = d / L as d * (1/L), z = t - rint(t), one
// FMA for z^2 + delta^2, a refined reciprocal square root, and three more
// FMA-class operations: 13 FP64 VALU operations and one v_rsq_f64. C is
// computed on the device with the same refined rsqrt, so z = +-1/2 gives
// exactly 0 like the reference's skip; z = 0 gives 0 through the factor z.
//
// Mapping to CDNA4: the source loop index is uniform across a wavefront, so
// x_j and w_j are scalar loads into SGPRs feeding every lane's VALU operation
// directly (no LDS traffic, no bank conflicts; the scalar cache serves the
// whole wavefront). Each lane owns TPT targets for instruction-level
// parallelism. When there are too few targets to fill 256 CUs the source
// range is split across blockIdx.y and the per-split partials are summed by a
// second kernel in split order, so results are deterministic and
// independent of co-scheduled work. Accumulation: a plain FP64 sum over each
// run of kSub sources, folded into a TwoSum (hi, lo) pair per target.
#include <cstring>
#include <mutex>

#include "fs_internal.h"
#include "fs_sym.h"

namespace fs {
namespace {

constexpr int kBlock = 256;
constexpr int kSub = 256;

// 1/sqrt(a), a > 0: hardware estimate (v_rsq_f64, relative error about
// 2^-23) plus one third-order correction y0 (1 + e/2 + 3e^2/8), e = 1 - a y0^2,
// which leaves a truncation error of order e^3 below FP64 rounding.
__device__ __forceinline__ double rsq_f64(double a) {
    const double y0 = __builtin_amdgcn_rsq(a);
    const double t = a * y0;
    const double e = __builtin_fma(-t, y0, 1.0);
    const double r = y0 * e;
    const double p = __builtin_fma(e, 0.375, 0.5);
    return __builtin_fma(r, p, y0);
}

// s + lo += b, error-free (Knuth TwoSum).
__device__ __forceinline__ void two_sum_acc(double& s, double& lo, double b) {
    const double t = s + b;
    const double bp = t - s;
    const double ap = t - bp;
    lo += (s - ap) + (b - bp);
    s = t;
}

template <int TPT>
__global__ __launch_bounds__(kBlock) void emq_direct_f64_kernel(
    const double* __restrict__ tgt, int nt, const double* __restrict__ src,
    const double* __restrict__ w, int ns, int chunk, double invL, double delta2,
    double* __restrict__ out) {
    const int split = blockIdx.y;
    const int j0 = split * chunk;
    const int j1 = min(ns, j0 + chunk);
    const int i0 = blockIdx.x * (kBlock * TPT) + threadIdx.x;
    double xi[TPT], hi[TPT], lo[TPT];
#pragma unroll
    for (int t = 0; t < TPT; ++t) {
        const int i = i0 + t * kBlock;
        xi[t] = (i < nt) ? tgt[i] : 0.0;
        hi[t] = 0.0;
        lo[t] = 0.0;
    }
    const double C = rsq_f64(__builtin_fma(0.5, 0.5, delta2));
    for (int jb = j0; jb < j1; jb += kSub) {
        const int je = min(j1, jb + kSub);
        double acc[TPT];
#pragma unroll
        for (int t = 0; t < TPT; ++t) acc[t] = 0.0;
#pragma unroll 4
        for (int j = jb; j < je; ++j) {
            const double xj = src[j];
            const double wj = w[j];
#pragma unroll
            for (int t = 0; t < TPT; ++t) {
                const double s = (xi[t] - xj) * invL;
                const double z = s - __builtin_rint(s);
                const double y = rsq_f64(__builtin_fma(z, z, delta2));
                acc[t] = __builtin_fma(wj * z, y - C, acc[t]);
            }
        }
#pragma unroll
        for (int t = 0; t < TPT; ++t) two_sum_acc(hi[t], lo[t], acc[t]);
    }
#pragma unroll
    for (int t = 0; t < TPT; ++t) {
        const int i = i0 + t * kBlock;
        if (i < nt) out[(size_t)split * nt + i] = 0.5 * (hi[t] + lo[t]);
    }
}

__global__ __launch_bounds__(kBlock) void reduce_splits_kernel(
    const double* __restrict__ part, int nt, int splits, double* __restrict__ es) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= nt) return;
    double s = 0.0, lo = 0.0;
    for (int k = 0; k < splits; ++k) two_sum_acc(s, lo, part[(size_t)k * nt + i]);
    es[i] = s + lo;
}

__global__ __launch_bounds__(kBlock) void fill_kernel(double* __restrict__ a, int n, double v) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) a[i] = v;
}

struct Plan {
    int tpt;
    int splits;
    int chunk;
};

Plan plan_direct(int64_t nt, int64_t ns, int num_cus, const FieldTuning& tun) {
    Plan p{};
    // Aim for >= 16 wavefronts per CU (4 per SIMD) so FP64 latency is hidden.
    const int64_t want_waves = int64_t(16) * num_cus;
    if (tun.targets_per_thread == 1 || tun.targets_per_thread == 2 ||
        tun.targets_per_thread == 4) {
        p.tpt = tun.targets_per_thread;
    } else {
        p.tpt = (nt >= want_waves * 64 * 2) ? 2 : 1;
    }
    const int64_t waves = (nt + 64 * p.tpt - 1) / (64 * p.tpt);
    int64_t splits = tun.source_splits > 0 ? tun.source_splits
                                           : (want_waves + waves - 1) / std::max<int64_t>(waves, 1);
    const int64_t max_splits = std::max<int64_t>(1, ns / kSub);
    splits = std::min(std::max<int64_t>(splits, 1), max_splits);
    splits = std::min<int64_t>(splits, 65535);
    int64_t chunk = (ns + splits - 1) / splits;
    chunk = (chunk + kSub - 1) / kSub * kSub;  // split boundaries on sub-sum boundaries
    chunk = std::max<int64_t>(chunk, kSub);   // ns == 0 still gives a valid plan
    splits = std::max<int64_t>(1, (ns + chunk - 1) / chunk);
    p.splits = (int)splits;
    p.chunk = (int)chunk;
    return p;
}

}  // namespace

bool use_sym(int64_t nt, int64_t ns, bool same, const FieldTuning& tun) {
    if (!same || nt != ns || nt == 0 || tun.symmetric == 1) return false;
    return tun.symmetric == 2 || nt >= sym::kMinPoints;
}

size_t field_scratch_bytes(int64_t nt, int64_t ns, bool same, int num_cus,
                           const FieldTuning& tun) {
    if (use_sym(nt, ns, same, tun))
        return sym::scratch_bytes(sym::plan(nt, tun, tun.sym_budget, true)) + sym::sort_bytes(nt);
    const Plan p = plan_direct(nt, ns, num_cus, tun);
    return p.splits > 1 ? align_up(sizeof(double) * nt * p.splits, 256) : 0;
}

int field_direct_f64(const double* d_tgt, int64_t nt, const double* d_src, const double* d_w,
                     int64_t ns, double L, double eps, double* d_es, void* scratch,
                     int num_cus, const FieldTuning& tun, hipStream_t stream,
                     int64_t* pair_evals) {
    if (pair_evals) *pair_evals = 0;
    if (nt < 0 || ns < 0 || nt > INT32_MAX || ns > INT32_MAX)
        return fail(FS_E_INVALID, "direct sum: sizes nt=%lld ns=%lld out of range",
                    (long long)nt, (long long)ns);
    if (!(L > 0)) return fail(FS_E_INVALID, "direct sum: L must be > 0");
    if (nt == 0) return FS_OK;
    if (!d_tgt || !d_es || (ns > 0 && (!d_src || !d_w)))
        return fail(FS_E_INVALID, "direct sum: null pointer");
    const int nti = (int)nt, nsi = (int)ns;
    if (ns == 0) {
        fill_kernel<<<(nti + kBlock - 1) / kBlock, kBlock, 0, stream>>>(d_es, nti, 0.0);
        FS_HIP_CHECK(hipGetLastError());
        return FS_OK;
    }
    const double invL = 1.0 / L;
    const double delta = eps / L;
    const double delta2 = delta * delta;
    if (use_sym(nt, ns, d_tgt == d_src, tun)) {
        const sym::Plan P = sym::plan(nt, tun, tun.sym_budget, true);
        FS_TRY(sym::run_sorted(d_src, d_w, P, L, invL, delta2, scratch, d_es, stream));
        if (pair_evals) *pair_evals = sym::evals(P);
        return FS_OK;
    }
    const Plan p = plan_direct(nt, ns, num_cus, tun);
    double* part = static_cast<double*>(scratch);
    double* out = (p.splits > 1) ? part : d_es;
    const dim3 grid((unsigned)((nt + (int64_t)kBlock * p.tpt - 1) / ((int64_t)kBlock * p.tpt)),
                    (unsigned)p.splits);
    switch (p.tpt) {
        case 4:
            emq_direct_f64_kernel<4><<<grid, kBlock, 0, stream>>>(d_tgt, nti, d_src, d_w,
                                                                  nsi, p.chunk, invL, delta2, out);
            break;
        case 2:
            emq_direct_f64_kernel<2><<<grid, kBlock, 0, stream>>>(d_tgt, nti, d_src, d_w,
                                                                  nsi, p.chunk, invL, delta2, out);
            break;
        default:
            emq_direct_f64_kernel<1><<<grid, kBlock, 0, stream>>>(d_tgt, nti, d_src, d_w,
                                                                  nsi, p.chunk, invL, delta2, out);
            break;
    }
    FS_HIP_CHECK(hipGetLastError());
    if (pair_evals)
        *pair_evals = (int64_t)grid.x * kBlock * p.tpt * ns;
    if (p.splits > 1) {
        reduce_splits_kernel<<<(nti + kBlock - 1) / kBlock, kBlock, 0, stream>>>(part, nti,
                                                                               p.splits, d_es);
        FS_HIP_CHECK(hipGetLastError());
    }
    return FS_OK;
}

// ---- the allocation-free form --------------------------------------------
// The work space of fs_emq_periodic_f64_noalloc: a fixed array in device
// memory, one per device the library runs on, held until the process ends.
// Host code never uses the array's name as an address: each device's
// address comes from hipGetSymbolAddress at context creation.
static __device__ __align__(256) double scratch[kNoAllocBytes / sizeof(double)];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
void* g_noalloc_ptr[kMaxDevices] = {};
}  // namespace

int noalloc_init(int device) {
    if (device < 0 || device >= kMaxDevices)
        return fail(FS_E_INVALID, "no-alloc work space: device %d beyond %d", device, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (g_noalloc_ptr[device]) return FS_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        return fail(FS_E_HIP, "hipGetSymbolAddress of the no-alloc work space failed: %s",
                    hipGetErrorString(e));
    }
    g_noalloc_ptr[device] = p;
    return FS_OK;
}

void* noalloc_ptr(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    return g_noalloc_ptr[device];
}

}  // namespace fs

extern "C" {

int fs_emq_periodic_f64_noalloc(fs_ctx* ctx, double* d_es, const double* d_targets, int64_t nt,
                                const double* d_sources, const double* d_q_ws, int64_t ns,
                                double L, double epsilon, void* stream) {
    if (!ctx) return fs::fail(FS_E_INVALID, "null context");
    if (nt < 0 || ns < 0 || nt > INT32_MAX || ns > INT32_MAX)
        return fs::fail(FS_E_INVALID, "direct sum: sizes nt=%lld ns=%lld out of range",
                        (long long)nt, (long long)ns);
    void* ws = fs::noalloc_ptr(ctx->device);
    if (!ws) {  // context creation could not take the address: once more, reporting why
        FS_HIP_CHECK(hipSetDevice(ctx->device));
        FS_TRY(fs::noalloc_init(ctx->device));
        ws = fs::noalloc_ptr(ctx->device);
    }
    hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
    // the pair-once partials in passes within the fixed array, beside the
    // sort of the points
    fs::FieldTuning tun = ctx->tuning;
    tun.sym_budget = std::max<int64_t>(
        1, (int64_t)fs::kNoAllocBytes - (int64_t)fs::sym::sort_bytes(std::max<int64_t>(nt, 1)));
    const size_t need = fs::field_scratch_bytes(nt, ns, d_targets == d_sources, ctx->num_cus, tun);
    if (need > fs::kNoAllocBytes)
        return fs::fail(FS_E_RANGE,
                        "fs_emq_periodic_f64_noalloc: %lld points need %zu bytes of work space, "
                        "beyond its fixed %zu (the pair-once sum, its partials and the sort of "
                        "the points, takes about 150-210 B per point: at most about 1.6e6 "
                        "points)",
                        (long long)nt, need, fs::kNoAllocBytes);
    return fs::field_direct_f64(d_targets, nt, d_sources, d_q_ws, ns, L, epsilon, d_es, ws,
                                ctx->num_cus, tun, st, &ctx->f64_pair_evals);
}

int fs_emq_set_tuning(fs_ctx* ctx, int targets_per_thread, int source_splits) {
    if (!ctx) return fs::fail(FS_E_INVALID, "null context");
    if (targets_per_thread != 0 && targets_per_thread != 1 && targets_per_thread != 2 &&
        targets_per_thread != 4 && targets_per_thread != 8)
        return fs::fail(FS_E_INVALID, "targets_per_thread must be 0, 1, 2, 4 or 8");
    if (source_splits < 0) return fs::fail(FS_E_INVALID, "source_splits must be >= 0");
    ctx->tuning.targets_per_thread = targets_per_thread;
    ctx->tuning.source_splits = source_splits;
    return FS_OK;
}

int fs_emq_f64_pair_evals(fs_ctx* ctx, int64_t* evals) {
    if (!ctx || !evals) return fs::fail(FS_E_INVALID, "null pointer");
    *evals = ctx->f64_pair_evals;
    return FS_OK;
}

int fs_emq_set_pair_once_budget(fs_ctx* ctx, int64_t bytes) {
    if (!ctx) return fs::fail(FS_E_INVALID, "null context");
    if (bytes < 0) return fs::fail(FS_E_INVALID, "budget must be >= 0");
    ctx->tuning.sym_budget = bytes;
    return FS_OK;
}

int fs_emq_pair_once_plan(int64_t n, int64_t budget_bytes, int64_t* scratch_bytes, int* passes,
                          int64_t* evals) {
    if (n < 1 || n > INT32_MAX) return fs::fail(FS_E_INVALID, "n must be in [1, 2^31)");
    if (budget_bytes < 0) return fs::fail(FS_E_INVALID, "budget must be >= 0");
    const fs::sym::Plan P = fs::sym::plan(n, fs::FieldTuning{}, budget_bytes);
    if (scratch_bytes) *scratch_bytes = (int64_t)fs::sym::scratch_bytes(P);
    if (passes) *passes = P.npass;
    if (evals) *evals = fs::sym::evals(P);
    return FS_OK;
}

int fs_emq_set_symmetric(fs_ctx* ctx, int mode) {
    if (!ctx) return fs::fail(FS_E_INVALID, "null context");
    if (mode < 0 || mode > 2) return fs::fail(FS_E_INVALID, "symmetric mode must be 0, 1 or 2");
    ctx->tuning.symmetric = mode;
    return FS_OK;
}

int fs_emq_periodic_f64_dev(fs_ctx* ctx, double* d_es, const double* d_targets, int64_t nt,
                            const double* d_sources, const double* d_q_ws, int64_t ns,
                            double L, double epsilon, void* stream) {
    if (!ctx) return fs::fail(FS_E_INVALID, "null context");
    hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
    const size_t need = fs::field_scratch_bytes(nt, ns, d_targets == d_sources, ctx->num_cus,
                                                ctx->tuning);
    FS_TRY(ctx->scratch.ensure(need, st));
    return fs::field_direct_f64(d_targets, nt, d_sources, d_q_ws, ns, L, epsilon, d_es,
                                ctx->scratch.ptr, ctx->num_cus, ctx->tuning, st,
                                &ctx->f64_pair_evals);
}

int fs_emq_periodic_f64(fs_ctx* ctx, double* es, const double* targets, int64_t nt,
                        const double* sources, const double* q_ws, int64_t ns, double L,
                        double epsilon) {
    if (!ctx) return fs::fail(FS_E_INVALID, "null context");
    if (nt < 0 || ns < 0) return fs::fail(FS_E_INVALID, "negative size");
    if (nt == 0) return FS_OK;
    if (!es || !targets || (ns > 0 && (!sources || !q_ws)))
        return fs::fail(FS_E_INVALID, "null pointer");
    FS_HIP_CHECK(hipSetDevice(ctx->device));
    const size_t bt = fs::align_up(sizeof(double) * nt, 256);
    const size_t bs = fs::align_up(sizeof(double) * ns, 256);
    FS_TRY(ctx->io.ensure(2 * bt + 2 * bs, ctx->stream));
    double* d_t = ctx->io.as<double>(0);
    double* d_e = ctx->io.as<double>(bt);
    double* d_s = ctx->io.as<double>(2 * bt);
    double* d_w = ctx->io.as<double>(2 * bt + bs);
    // targets and sources the same positions (the same array, or two copies
    // of it, as the reference passes them): one device copy, and the
    // pair-once form of the sum
    const bool same = nt == ns && (targets == sources ||
                                   std::memcmp(targets, sources, sizeof(double) * nt) == 0);
    if (same)
        d_t = d_s;
    else
        FS_HIP_CHECK(hipMemcpyAsync(d_t, targets, sizeof(double) * nt, hipMemcpyHostToDevice,
                                    ctx->stream));
    if (ns > 0) {
        FS_HIP_CHECK(hipMemcpyAsync(d_s, sources, sizeof(double) * ns, hipMemcpyHostToDevice,
                                    ctx->stream));
        FS_HIP_CHECK(hipMemcpyAsync(d_w, q_ws, sizeof(double) * ns, hipMemcpyHostToDevice,
                                    ctx->stream));
    }
    FS_TRY(fs_emq_periodic_f64_dev(ctx, d_e, d_t, nt, d_s, d_w, ns, L, epsilon, ctx->stream));
    FS_HIP_CHECK(hipMemcpyAsync(es, d_e, sizeof(double) * nt, hipMemcpyDeviceToHost,
                                ctx->stream));
    FS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return FS_OK;
}

}  // extern "C"
