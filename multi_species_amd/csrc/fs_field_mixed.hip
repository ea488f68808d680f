// Periodic regularized E-field, direct sum, mixed precision (gfx950).
//
// Same field as E_MQ_DirectSum (reference src/FieldStructure.cpp:106-156),
// E_i = sum_j w_j K(z_ij), K(z) = (z/2) (1/sqrt(z^2 + delta^2) - C), with
// z = (x_i - x_j)/L wrapped to [-1/2, 1/2), delta = eps/L, C = 2/sqrt(1 + 4
// delta^2). The kernel is split as
//     K(z) = 1/2 [ (s(z) - C z) + g(z) ],   g(z) = z/sqrt(z^2 + delta^2) - s(z),
// s(z) = +1 for z >= 0, -1 for z < 0.
//  * S_i = sum_j w_j (s(z_ij) - C z_ij), the bare periodic Coulomb step plus
//    the neutralizing background, is piecewise linear in the positions: with
//    the sources sorted by position it is a handful of prefix sums per target
//    (FP64, O(N log N) in total).
//  * G_i = sum_j w_j g(z_ij) carries the rest. g is O(1) only within a few
//    delta of z = 0 and decays like delta^2/(2 z^2) beyond. The sorted sources
//    form 32-source tiles, 128-source groups and far-field units of 512 and
//    2048 sources, and each (wavefront's 256 consecutive sorted targets,
//    unit, group or tile) takes
//    one of:
//      - beyond the band B = max(2^-6, 16 delta, 4000 delta^2) (B >= 1/4 sends
//        everything to FP64), the largest unit, else the group, else the
//        tile whose sources lie within B/16 of its centre: the FP32
//        expansion of g's series -s (q/2)(1 - 3q/4 + 5q^2/8), q =
//        delta^2/z^2, in the distance from the centre, from 8 FP64 moments
//        rounded once (one reciprocal and 12 FMAs per target for up to 2048
//        sources);
//      - beyond the band otherwise: the series pair by pair in FP32;
//      - a group or tile on one side of all the targets, at least gap away,
//        whose sources lie within sqrt(gap^2 + delta^2)/16 of its centre:
//        the FP64 Legendre expansion sum_j w_j z_j/sqrt(z_j^2 + delta^2) =
//        sum_k P_k(-Z/r) r^-(k+1) (Z M_k + M_k+1), r = sqrt(Z^2 + delta^2),
//        12 terms from 13 moments (truncation below 16^-12), minus
//        s sum_j w_j;
//      - otherwise (the target's own neighbourhood, spread-out tiles, tiles
//        within 2^-24 of the antipode where the sign flips) pair by pair in
//        FP64;
//    and per target the pieces go into an FP64 TwoSum (hi, lo) pair.
//  * With an error target (fs_emq_set_mixed_target) a probe first measures
//    on up to 128 targets the error each of 8 candidate bands would give
//    against the exact FP64 field, and the band is the narrowest within the
//    target.
// Positions are wrapped in FP64 (a = x/L - floor(x/L)) and turned into 64-bit
// fixed point U = a 2^64, so every displacement U_i - U_j (mod 2^64) is the
// exact periodic difference of the wrapped positions: the FP64 paths use it
// whole, the FP32 paths its high word (|z| >= B, so the high word keeps 2^-26
// relative precision).
//
// The FP64 method stays the drop-in (fs_field.hip); this one is measured
// against it: relative error max_i |E_i - E_i^FP64| / max_i |E_i^FP64| below
// 1e-8 on the decks the tests run (lattice decks, where rounding errors add
// coherently, are the hardest case), and at FP64 accuracy when every pair is
// in the band.
//
// Work on the device, in order (kernels in fs_field_batch.hip, shared with
// the batched decks): wrap + fixed point (prep), radix sort of the sources
// (and of the targets when they differ, as a second segment), gather of the
// sorted weights, tile and group expansions, a deterministic prefix scan,
// the probe and band choice when governed, the persistent pair kernel over
// (target block, source chunk) items, and a finalize kernel that evaluates
// S_i, adds the chunk partials in order and scatters E back to the caller's
// order. Every reduction runs in a fixed order, so results do not depend on
// scheduling.
//
// This file holds the C ABI of the single-call form.
#include <cstring>
#include <vector>

#include "fs_field_batch.h"

namespace fs {
namespace {

FieldBatch& api_batch(fs_ctx* ctx) {
    if (!ctx->api_batch) ctx->api_batch = new FieldBatch();
    return *ctx->api_batch;
}

// Launch shape for this problem class (sizes by powers of two): autotuned on
// the first call of a class, then reused. es is written by the tuning run.
int shape(fs_ctx* ctx, FieldBatch& fb, int64_t nt, int64_t ns, bool same, const double* x,
          const double* w, double* es, hipStream_t st) {
    if (ctx->tuning.targets_per_thread != 0) return FS_OK;  // explicit tuning wins
    auto lg = [](int64_t v) { uint64_t b = 0; while ((int64_t(1) << b) < v && b < 62) ++b; return b; };
    const uint64_t key = (lg(nt) << 8) | (lg(ns) << 1) | (same ? 1 : 0);
    std::pair<int, int> known{0, 0};
    {
        std::lock_guard<std::mutex> lk(ctx->cache_mu);
        auto it = ctx->shape_cache.find(key);
        if (it != ctx->shape_cache.end()) known = it->second;
    }
    if (known.first) return fb.set_shape(known.first, known.second, st);
    int t = 2, g = 8;
    FS_TRY(fb.autotune(x, w, es, st, &t, &g));
    std::lock_guard<std::mutex> lk(ctx->cache_mu);
    ctx->shape_cache[key] = std::make_pair(t, g);
    return FS_OK;
}

}  // namespace
}  // namespace fs

extern "C" {

int fs_emq_periodic_mixed_dev(fs_ctx* ctx, double* d_es, const double* d_targets, int64_t nt,
                              const double* d_sources, const double* d_q_ws, int64_t ns,
                              double L, double epsilon, void* stream) {
    if (!ctx) return fs::fail(FS_E_INVALID, "null context");
    if (nt < 0 || ns < 0 || nt + ns > INT32_MAX / 2)
        return fs::fail(FS_E_INVALID, "mixed direct sum: sizes nt=%lld ns=%lld out of range",
                        (long long)nt, (long long)ns);
    if (!(L > 0)) return fs::fail(FS_E_INVALID, "mixed direct sum: L must be > 0");
    if (!(epsilon > 0)) return fs::fail(FS_E_INVALID, "mixed direct sum: eps must be > 0");
    if (nt == 0) return FS_OK;
    if (!d_es || !d_targets || (ns > 0 && (!d_sources || !d_q_ws)))
        return fs::fail(FS_E_INVALID, "mixed direct sum: null pointer");
    // small calls: the FP64 sum is both exact and faster than the mixed
    // form's fixed costs
    if ((double)nt * (double)ns < (double)ctx->mixed_min_pairs)
        return fs_emq_periodic_f64_dev(ctx, d_es, d_targets, nt, d_sources, d_q_ws, ns, L,
                                       epsilon, stream);
    hipStream_t st = stream ? (hipStream_t)stream : ctx->stream;
    fs::FieldBatch& fb = fs::api_batch(ctx);
    const bool same = (d_targets == d_sources) && (nt == ns);
    std::vector<fs::BatchSeg> segs;
    if (same) {
        fs::BatchSeg a;
        a.off = 0;
        a.n = (int)nt;
        a.src = 0;
        a.eval = true;
        a.L = L;
        a.eps = epsilon;
        segs.push_back(a);
        FS_TRY(fb.plan(segs, (int)nt, fs::kPrecisionMixed, ctx->num_cus, ctx->tuning, st));
        FS_TRY(fs::shape(ctx, fb, nt, ns, true, d_sources, d_q_ws, d_es, st));
        fb.govern(ctx->mixed_target > 0, ctx->mixed_target);
        return fb.run(d_sources, d_q_ws, d_es, st);
    }
    // sources then targets in one array; the targets' segment is sourced
    // from the first
    const int n = (int)(ns + nt);
    fs::BatchSeg a, b;
    a.off = 0;
    a.n = (int)ns;
    a.src = 0;
    a.eval = false;
    a.L = L;
    a.eps = epsilon;
    b.off = (int)ns;
    b.n = (int)nt;
    b.src = 0;
    b.eval = true;
    b.L = L;
    b.eps = epsilon;
    segs.push_back(a);
    segs.push_back(b);
    const size_t bx = fs::align_up(sizeof(double) * n, 256);
    FS_TRY(ctx->aux.ensure(2 * bx, st));
    double* x = ctx->aux.as<double>(0);
    double* e = ctx->aux.as<double>(bx);
    if (ns > 0)
        FS_HIP_CHECK(hipMemcpyAsync(x, d_sources, sizeof(double) * ns, hipMemcpyDeviceToDevice, st));
    FS_HIP_CHECK(hipMemcpyAsync(x + ns, d_targets, sizeof(double) * nt, hipMemcpyDeviceToDevice, st));
    FS_TRY(fb.plan(segs, n, fs::kPrecisionMixed, ctx->num_cus, ctx->tuning, st));
    FS_TRY(fs::shape(ctx, fb, nt, ns, false, x, d_q_ws, e, st));
    fb.govern(ctx->mixed_target > 0, ctx->mixed_target);
    FS_TRY(fb.run(x, d_q_ws, e, st));
    FS_HIP_CHECK(hipMemcpyAsync(d_es, e + ns, sizeof(double) * nt, hipMemcpyDeviceToDevice, st));
    return FS_OK;
}

int fs_emq_set_mixed_target(fs_ctx* ctx, double rel_err) {
    if (!ctx) return fs::fail(FS_E_INVALID, "null context");
    if (!(rel_err >= 0) || rel_err > 1e-3)
        return fs::fail(FS_E_INVALID, "mixed target must be in [0, 1e-3], got %g", rel_err);
    ctx->mixed_target = rel_err;
    return FS_OK;
}

int fs_emq_set_mixed_min_pairs(fs_ctx* ctx, int64_t min_pairs) {
    if (!ctx) return fs::fail(FS_E_INVALID, "null context");
    if (min_pairs < 0)
        return fs::fail(FS_E_INVALID, "mixed min_pairs must be >= 0, got %lld",
                        (long long)min_pairs);
    ctx->mixed_min_pairs = min_pairs;
    return FS_OK;
}

int fs_emq_mixed_shape(fs_ctx* ctx, int* targets_per_thread, int* wgs_per_cu) {
    if (!ctx) return fs::fail(FS_E_INVALID, "null context");
    fs::FieldBatch& fb = fs::api_batch(ctx);
    if (targets_per_thread) *targets_per_thread = fb.tpt();
    if (wgs_per_cu) *wgs_per_cu = fb.wgs_per_cu();
    return FS_OK;
}

int fs_emq_periodic_mixed(fs_ctx* ctx, double* es, const double* targets, int64_t nt,
                          const double* sources, const double* q_ws, int64_t ns, double L,
                          double epsilon) {
    if (!ctx) return fs::fail(FS_E_INVALID, "null context");
    if (nt < 0 || ns < 0) return fs::fail(FS_E_INVALID, "negative size");
    if (nt == 0) return FS_OK;
    if (!es || !targets || (ns > 0 && (!sources || !q_ws)))
        return fs::fail(FS_E_INVALID, "null pointer");
    FS_HIP_CHECK(hipSetDevice(ctx->device));
    // targets and sources the same positions (one array or two copies): one
    // device copy and one shared sort, as in fs_emq_periodic_f64
    const bool same = nt == ns && (targets == sources ||
                                   std::memcmp(targets, sources, sizeof(double) * nt) == 0);
    const size_t bt = fs::align_up(sizeof(double) * nt, 256);
    const size_t bs = fs::align_up(sizeof(double) * ns, 256);
    FS_TRY(ctx->io.ensure(2 * bt + 2 * bs, ctx->stream));
    double* d_t = ctx->io.as<double>(0);
    double* d_e = ctx->io.as<double>(bt);
    double* d_s = same ? d_t : ctx->io.as<double>(2 * bt);
    double* d_w = ctx->io.as<double>(2 * bt + bs);
    FS_HIP_CHECK(hipMemcpyAsync(d_t, targets, sizeof(double) * nt, hipMemcpyHostToDevice,
                                ctx->stream));
    if (ns > 0) {
        if (!same)
            FS_HIP_CHECK(hipMemcpyAsync(d_s, sources, sizeof(double) * ns,
                                        hipMemcpyHostToDevice, ctx->stream));
        FS_HIP_CHECK(hipMemcpyAsync(d_w, q_ws, sizeof(double) * ns, hipMemcpyHostToDevice,
                                    ctx->stream));
    }
    FS_TRY(fs_emq_periodic_mixed_dev(ctx, d_e, d_t, nt, d_s, d_w, ns, L, epsilon, ctx->stream));
    FS_HIP_CHECK(hipMemcpyAsync(es, d_e, sizeof(double) * nt, hipMemcpyDeviceToHost,
                                ctx->stream));
    FS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return FS_OK;
}

}  // extern "C"
