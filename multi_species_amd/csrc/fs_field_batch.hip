// Batched periodic regularized direct sum (gfx950): many independent decks,
// each its own set of targets and sources, in one persistent launch.
//
// The field of each deck is the reference's E_MQ_DirectSum
// (src/FieldStructure.cpp:106-156) over that deck's points. The work of all
// decks is cut into items of (deck, block of 256*TPT targets, chunk of
// sources); a persistent kernel sized to the device pulls items from an
// atomic counter, largest first, so decks of very different sizes (2k to
// 500k points) share the chip without a launch per deck. Every item writes
// its own partial sum; a finalize kernel adds the partials of each target in
// chunk order. The chunking of a deck depends only on that deck's size and
// the device (not on the rest of the batch or on TPT), so a deck's field is
// bitwise the same alone and in any mix.
//
// Two precisions:
//  * FP64: per pair and summation order exactly those of fs_field.hip
//    (emq_direct_f64_kernel): d = x_i - x_j, z = d/L - rint, refined rsqrt,
//    TwoSum over 256-source runs, halved at the end.
//  * mixed: the decomposition of fs_field_mixed.hip's comment (FP64 step and
//    neutralizing part from sorted prefix sums, FP64 near band, FP32
//    far-field series, compensated FP64 accumulation), with positions sorted
//    per deck by the graph-safe radix sort of fs_radix.hip.
//
// run() launches kernels only (no allocation, memset or synchronization), so
// the caller can capture it in a HIP graph.
#include <algorithm>
#include <cmath>
#include <numeric>

#include "fs_device.h"
#include "fs_field_batch.h"

namespace fs {
namespace {

using dev::cyc_meets;
using dev::rsq_f64;
using dev::two_sum_acc;

constexpr int kBlock = 256;
constexpr int kSub = 256;          // FP64: sources per TwoSum run
constexpr int kTile = 32;          // mixed: sources per tile
constexpr int kScanItems = 2048;   // mixed: sorted positions per scan block
constexpr uint64_t kHalf = 0x8000000000000000ull;
constexpr uint64_t kAntiMargin = uint64_t(1) << 40;
constexpr uint64_t kLow32 = 0xffffffffull;

__device__ __forceinline__ int pull_item(int* counter) {
    __shared__ int s_item;
    if (threadIdx.x == 0) s_item = atomicAdd(counter, 1);
    __syncthreads();
    const int it = __builtin_amdgcn_readfirstlane(s_item);
    __syncthreads();
    return it;
}

// ---- FP64 ------------------------------------------------------------------

template <int TPT>
__global__ __launch_bounds__(kBlock) void batch_f64_kernel(const WorkItem* __restrict__ items,
                                                           int nitems,
                                                           const SegDev* __restrict__ segs,
                                                           const double* __restrict__ x,
                                                           const double* __restrict__ w,
                                                           double* __restrict__ part,
                                                           int* __restrict__ counter) {
    for (;;) {
        const int it = pull_item(counter);
        if (it >= nitems) break;
        const WorkItem W = items[it];
        const SegDev S = segs[W.seg];
        const double invL = S.invL, delta2 = S.delta2;
        const double C = rsq_f64(__builtin_fma(0.5, 0.5, delta2));
        double xi[TPT], hi[TPT], lo[TPT];
#pragma unroll
        for (int t = 0; t < TPT; ++t) {
            const int i = W.tb + (int)threadIdx.x + t * kBlock;
            xi[t] = (i < W.te) ? x[i] : 0.0;
            hi[t] = 0.0;
            lo[t] = 0.0;
        }
        for (int jb = W.sb; jb < W.se; jb += kSub) {
            const int je = min(W.se, jb + kSub);
            double acc[TPT];
#pragma unroll
            for (int t = 0; t < TPT; ++t) acc[t] = 0.0;
#pragma unroll 4
            for (int j = jb; j < je; ++j) {
                const double xj = x[j];
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
            const int i = W.tb + (int)threadIdx.x + t * kBlock;
            if (i < W.te)
                part[S.part_base + (int64_t)W.split * S.n + (i - S.off)] = 0.5 * (hi[t] + lo[t]);
        }
    }
}

__global__ __launch_bounds__(kBlock) void finalize_f64_kernel(
    int n_total, const int32_t* __restrict__ seg_of, const SegDev* __restrict__ segs,
    const double* __restrict__ part, double* __restrict__ es, int* __restrict__ counter) {
    const int g = blockIdx.x * kBlock + threadIdx.x;
    if (g == 0) *counter = 0;  // the pair kernel has finished: ready for the next run
    if (g >= n_total) return;
    const SegDev& S = segs[seg_of[g]];
    if (!S.eval) return;
    if (S.src_n == 0) {
        es[g] = 0.0;
        return;
    }
    double s = 0.0, lo = 0.0;
    for (int k = 0; k < S.splits; ++k)
        two_sum_acc(s, lo, part[S.part_base + (int64_t)k * S.n + (g - S.off)]);
    es[g] = s + lo;
}

// ---- mixed -----------------------------------------------------------------

__global__ __launch_bounds__(kBlock) void prep_kernel(int n_total, const int32_t* __restrict__ seg_of,
                                                      const SegDev* __restrict__ segs,
                                                      const double* __restrict__ x,
                                                      uint64_t* __restrict__ key) {
    // t must be rounded before floor(t) is taken from it: a fused
    // x*invL - floor(t) would return a negative residue when x*invL rounds up
    // to an integer
#pragma clang fp contract(off)
    const int g = blockIdx.x * kBlock + threadIdx.x;
    if (g >= n_total) return;
    const double t = x[g] * segs[seg_of[g]].invL;
    double a = t - __builtin_floor(t);  // exact, in [0, 1)
    if (!(a >= 0.0 && a < 1.0)) a = 0.0;  // NaN or infinite positions
    key[g] = (uint64_t)__builtin_ldexp(a, 64);
}

// Sorted weights, FP32 far-field weights, high words and per-block sums.
// Coincident sources (equal keys: the species of a deck share the x lattice
// of the remesh grid) have the same far-field kernel value, so within each
// 32-source tile (the unit that takes either the FP64 or the FP32 path) their
// FP32 far-field weight is their FP64 sum, held by the first of them (the
// others get 0): the FP32 terms then carry the net charge, and their rounding
// errors scale with it instead of with each species' charge.
__global__ __launch_bounds__(kBlock) void gather_kernel(
    const ScanBlock* __restrict__ blocks, const SegDev* __restrict__ segs,
    const uint64_t* __restrict__ us, const int32_t* __restrict__ perm,
    const double* __restrict__ w, uint32_t* __restrict__ hs, double* __restrict__ ws,
    float* __restrict__ wf, double* __restrict__ bsum, double* __restrict__ bsa) {
    __shared__ double sw[kBlock], swa[kBlock];
    const ScanBlock B = blocks[blockIdx.x];
    const SegDev& S = segs[B.seg];
    const double wscale = S.wscale;
    const int s_lo = S.off, s_hi = S.off + S.n;  // the tiles never cross these
    constexpr int kPer = kScanItems / kBlock;
    const int base = B.b + (int)threadIdx.x * kPer;
    double s = 0.0, sa = 0.0;
    for (int k = 0; k < kPer; ++k) {
        const int i = base + k;
        if (i < B.e) {
            const uint64_t u = us[i];
            const uint32_t top = (uint32_t)(u >> 32);
            const double wi = w[perm[i]];
            hs[i] = top;
            ws[i] = wi;
            s += wi;
            sa += wi * __builtin_ldexp((double)u, -64);
            // tiles start at multiples of kTile from the segment start
            const int t0 = s_lo + ((i - s_lo) & ~(kTile - 1));
            const int t1 = min(t0 + kTile, s_hi);
            bool first = true;
            for (int j = t0; j < i; ++j)
                if (us[j] == u) {
                    first = false;
                    break;
                }
            double net = 0.0;
            if (first)
                for (int j = i; j < t1; ++j)
                    if (us[j] == u) net += w[perm[j]];
            wf[i] = (float)(net * wscale);
        }
    }
    sw[threadIdx.x] = s;
    swa[threadIdx.x] = sa;
    __syncthreads();
    for (int h = kBlock / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            sw[threadIdx.x] += sw[threadIdx.x + h];
            swa[threadIdx.x] += swa[threadIdx.x + h];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        bsum[blockIdx.x] = sw[0];
        bsa[blockIdx.x] = swa[0];
    }
}

// Exclusive prefix sums of each source segment's sorted weights into
// pw[pw_base + 0 .. n], and totals[seg] = sum w a. Fixed order throughout.
__global__ __launch_bounds__(kBlock) void scan_kernel(
    const ScanBlock* __restrict__ blocks, const SegDev* __restrict__ segs,
    const double* __restrict__ ws, const double* __restrict__ bsum,
    const double* __restrict__ bsa, double* __restrict__ pw, double* __restrict__ totals) {
    __shared__ double red[kBlock], red2[kBlock];
    const ScanBlock B = blocks[blockIdx.x];
    const SegDev& S = segs[B.seg];
    const int me = (int)blockIdx.x - B.first;
    double o = 0.0, oa = 0.0;
    for (int k = threadIdx.x; k < me; k += kBlock) o += bsum[B.first + k];
    if (me == 0)
        for (int k = threadIdx.x; k < B.count; k += kBlock) oa += bsa[B.first + k];
    red[threadIdx.x] = o;
    red2[threadIdx.x] = oa;
    __syncthreads();
    for (int h = kBlock / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            red[threadIdx.x] += red[threadIdx.x + h];
            red2[threadIdx.x] += red2[threadIdx.x + h];
        }
        __syncthreads();
    }
    const double boff = red[0];
    if (me == 0 && threadIdx.x == 0) {
        totals[B.seg] = red2[0];
        if (S.n == 0) pw[S.pw_base] = 0.0;
    }
    __syncthreads();
    constexpr int kPer = kScanItems / kBlock;
    const int base = B.b + (int)threadIdx.x * kPer;
    double v[kPer];
    double s = 0.0;
    for (int k = 0; k < kPer; ++k) {
        const int i = base + k;
        v[k] = (i < B.e) ? ws[i] : 0.0;
        s += v[k];
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = 1; h < kBlock; h <<= 1) {
        const double add = ((int)threadIdx.x >= h) ? red[threadIdx.x - h] : 0.0;
        __syncthreads();
        red[threadIdx.x] += add;
        __syncthreads();
    }
    double run = boff + (red[threadIdx.x] - s);
    const int last = S.off + S.n - 1;
    for (int k = 0; k < kPer; ++k) {
        const int i = base + k;
        if (i < B.e) pw[S.pw_base + (i - S.off)] = run;
        run += v[k];
        if (i == last) pw[S.pw_base + S.n] = run;
    }
}

template <int TPT>
__global__ __launch_bounds__(kBlock) void batch_mixed_kernel(
    const WorkItem* __restrict__ items, int nitems, const SegDev* __restrict__ segs,
    const uint64_t* __restrict__ us, const uint32_t* __restrict__ hs,
    const double* __restrict__ ws, const float* __restrict__ wf, double* __restrict__ part,
    int* __restrict__ counter) {
    for (;;) {
        const int it = pull_item(counter);
        if (it >= nitems) break;
        const WorkItem W = items[it];
        const SegDev S = segs[W.seg];
        // target positions: sorted by their top words only, so widen the
        // block's range by one unit of 2^32 on each side
        const uint64_t uf = us[W.tb] & ~kLow32, ul = us[W.te - 1] | kLow32;
        const uint64_t tspan = ul - uf;
        uint64_t ui[TPT];
        uint32_t hi[TPT];
        double sh[TPT], sl[TPT];
#pragma unroll
        for (int t = 0; t < TPT; ++t) {
            const int i = min(W.tb + (int)threadIdx.x + t * kBlock, W.te - 1);
            ui[t] = us[i];
            hi[t] = (uint32_t)(ui[t] >> 32);
            sh[t] = 0.0;
            sl[t] = 0.0;
        }
        for (int jb = W.sb; jb < W.se; jb += kTile) {
            const int je = min(W.se, jb + kTile);
            const uint64_t vf = us[jb] & ~kLow32, vl = us[je - 1] | kLow32;
            const uint64_t vspan = vl - vf;
            const bool wide = tspan >= kHalf || vspan >= kHalf || tspan + vspan >= kHalf;
            const uint64_t len = tspan + vspan;
            const uint64_t lo = uf - vl;
            const bool band = S.all_band || wide || cyc_meets(lo + S.band, len, 2 * S.band) ||
                              cyc_meets(lo - kHalf + kAntiMargin, len, 2 * kAntiMargin);
            if (band) {
                double acc[TPT];
#pragma unroll
                for (int t = 0; t < TPT; ++t) acc[t] = 0.0;
#pragma unroll 2
                for (int j = jb; j < je; ++j) {
                    const uint64_t uj = us[j];
                    const double wj = ws[j];
#pragma unroll
                    for (int t = 0; t < TPT; ++t) {
                        const uint64_t d = ui[t] - uj;
                        const double zs = __builtin_fma((double)(int32_t)(uint32_t)(d >> 32),
                                                        4294967296.0, (double)(uint32_t)d);
                        const double y = rsq_f64(__builtin_fma(zs, zs, S.d2s));
                        const double s = __builtin_copysign(1.0, zs);
                        acc[t] = __builtin_fma(wj, __builtin_fma(zs, y, -s), acc[t]);
                    }
                }
#pragma unroll
                for (int t = 0; t < TPT; ++t) two_sum_acc(sh[t], sl[t], acc[t]);
            } else {
                float acc[TPT];
#pragma unroll
                for (int t = 0; t < TPT; ++t) acc[t] = 0.0f;
#pragma unroll 8
                for (int j = jb; j < je; ++j) {
                    const uint32_t hj = hs[j];
                    const float wj = wf[j];
#pragma unroll
                    for (int t = 0; t < TPT; ++t) {
                        const float zs = (float)(int32_t)(hi[t] - hj);
                        const float r = __builtin_amdgcn_rcpf(zs);
                        const float sq = r * __builtin_fabsf(r);  // s(z) / zs^2
                        const float q = __builtin_fabsf(sq);
                        const float p = __builtin_fmaf(q, __builtin_fmaf(q, S.b2, S.b1), 1.0f);
                        acc[t] = __builtin_fmaf(wj * sq, p, acc[t]);
                    }
                }
#pragma unroll
                for (int t = 0; t < TPT; ++t) two_sum_acc(sh[t], sl[t], (double)acc[t]);
            }
        }
#pragma unroll
        for (int t = 0; t < TPT; ++t) {
            const int i = W.tb + (int)threadIdx.x + t * kBlock;
            if (i < W.te) part[S.part_base + (int64_t)W.split * S.n + (i - S.off)] = sh[t] + sl[t];
        }
    }
}

// Sum of the sorted weights of [off, off+n) whose key is <= key. The keys
// are sorted by their top words; within a run of equal top words the full
// keys are compared one by one.
__device__ double w_le(const uint64_t* __restrict__ us, const double* __restrict__ ws,
                       const double* __restrict__ pw, int off, int n, int pwb, uint64_t key) {
    const uint32_t kt = (uint32_t)(key >> 32);
    int lo = 0, hi = n;  // first index with top >= kt
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if ((uint32_t)(us[off + mid] >> 32) < kt)
            lo = mid + 1;
        else
            hi = mid;
    }
    double s = pw[pwb + lo];
    for (int j = lo; j < n && (uint32_t)(us[off + j] >> 32) == kt; ++j)
        if (us[off + j] <= key) s += ws[off + j];
    return s;
}

// E = (S + G)/2 at every sorted position of the evaluated segments,
// scattered to the original index. With P(<= u) the weight of the sources at
// or left of u, W the total and A = sum w a:
//   sum_j w_j s_ij = 2 (P(<= u_i) - P(<= u_i - 1/2) - P(<= u_i + 1/2)) + W
//   sum_j w_j z_ij = a_i W - A - P(<= u_i - 1/2) + (W - P(<= u_i + 1/2)),
// where P(<= u_i - 1/2) = 0 for u_i < 1/2 and P(<= u_i + 1/2) = W for u_i >= 1/2.
__global__ __launch_bounds__(kBlock) void finalize_mixed_kernel(
    int n_total, const int32_t* __restrict__ seg_of, const SegDev* __restrict__ segs,
    const uint64_t* __restrict__ us, const int32_t* __restrict__ perm,
    const double* __restrict__ ws, const double* __restrict__ pw,
    const double* __restrict__ totals, const double* __restrict__ part,
    double* __restrict__ es, int* __restrict__ counter) {
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k == 0) *counter = 0;
    if (k >= n_total) return;
    const int d = seg_of[k];
    const SegDev& S = segs[d];
    if (!S.eval) return;
    if (S.src_n == 0) {
        es[perm[k]] = 0.0;
        return;
    }
    double g = 0.0, gl = 0.0;
    for (int s = 0; s < S.splits; ++s)
        two_sum_acc(g, gl, part[S.part_base + (int64_t)s * S.n + (k - S.off)]);
    const SegDev& R = segs[S.src];
    const uint64_t u = us[k];
    const double W = pw[R.pw_base + R.n], A = totals[S.src];
    const double P0 = w_le(us, ws, pw, R.off, R.n, R.pw_base, u);
    const double Pm = (u >= kHalf) ? w_le(us, ws, pw, R.off, R.n, R.pw_base, u - kHalf) : 0.0;
    const double Pp = (u < kHalf) ? w_le(us, ws, pw, R.off, R.n, R.pw_base, u + kHalf) : W;
    const double sum_s = 2.0 * ((P0 - Pm) - Pp) + W;
    const double ai = __builtin_ldexp((double)u, -64);
    const double sum_z = ((ai * W - A) - Pm) + (W - Pp);
    es[perm[k]] = 0.5 * ((sum_s - S.C * sum_z) + (g + gl));
}

__global__ void zero_int_kernel(int* p) { *p = 0; }

inline unsigned nblocks(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// Source chunking of one deck: the FP64 rule of fs_field.hip (plan_direct)
// and its mixed counterpart. Depends only on the deck's sizes and the device.
void chunking(int precision, int64_t nt, int64_t ns, int num_cus, const FieldTuning& tun,
              int& splits, int& chunk) {
    const int64_t want_waves = int64_t(16) * num_cus;
    const int unit = precision == kPrecisionF64 ? kSub : kTile;
    const int tpt_auto = (nt >= want_waves * 64 * 2) ? 2 : 1;
    const int64_t waves = (nt + 64 * tpt_auto - 1) / (64 * tpt_auto);
    int64_t sp = tun.source_splits > 0 ? tun.source_splits
                                       : (want_waves + waves - 1) / std::max<int64_t>(waves, 1);
    const int64_t max_splits =
        std::max<int64_t>(1, ns / (precision == kPrecisionF64 ? kSub : 8 * kTile));
    sp = std::min(std::max<int64_t>(sp, 1), max_splits);
    sp = std::min<int64_t>(sp, 65535);
    int64_t ch = (ns + sp - 1) / sp;
    ch = (ch + unit - 1) / unit * unit;
    ch = std::max<int64_t>(ch, unit);
    splits = (int)std::max<int64_t>(1, (ns + ch - 1) / ch);
    chunk = (int)ch;
}

}  // namespace

void FieldBatch::release() {
    tables_.release();
    work_.release();
    segs_.clear();
    items_.clear();
    scan_.clear();
}

int FieldBatch::plan(const std::vector<BatchSeg>& segs, int n_total, int precision, int num_cus,
                     const FieldTuning& tun, hipStream_t stream) {
    if (precision != kPrecisionF64 && precision != kPrecisionMixed)
        return fail(FS_E_INVALID, "field batch: unknown precision %d", precision);
    if (segs.empty() || n_total < 0) return fail(FS_E_INVALID, "field batch: no segments");
    precision_ = precision;
    n_total_ = n_total;
    nseg_ = (int)segs.size();
    tpt_ = (tun.targets_per_thread == 1 || tun.targets_per_thread == 2 ||
            tun.targets_per_thread == 4)
               ? tun.targets_per_thread
               : 2;
    segs_.assign(nseg_, SegDev{});
    items_.clear();
    scan_.clear();
    pairs_ = 0;
    std::vector<int32_t> seg_of(std::max(n_total, 1), 0);
    int64_t part_total = 0;
    for (int d = 0; d < nseg_; ++d) {
        const BatchSeg& b = segs[d];
        if (b.off < 0 || b.n < 0 || b.off + b.n > n_total || b.src < 0 || b.src >= nseg_)
            return fail(FS_E_INVALID, "field batch: segment %d out of range", d);
        if (!(b.L > 0)) return fail(FS_E_INVALID, "field batch: segment %d has L <= 0", d);
        if (precision == kPrecisionMixed && b.eval && !(b.eps > 0))
            return fail(FS_E_INVALID, "field batch: mixed precision needs eps > 0");
        for (int g = b.off; g < b.off + b.n; ++g) seg_of[g] = d;
        SegDev& S = segs_[d];
        S.off = b.off;
        S.n = b.n;
        S.src = b.src;
        S.src_off = segs[b.src].off;
        S.src_n = segs[b.src].n;
        S.eval = b.eval ? 1 : 0;
        S.invL = 1.0 / b.L;
        const double delta = b.eps / b.L;
        S.delta2 = delta * delta;
        S.C = 2.0 / std::sqrt(1.0 + 4.0 * S.delta2);
        S.pw_base = b.off + d;
        S.splits = 1;
        S.chunk = std::max(1, S.src_n);
        if (b.eval && b.n > 0) {
            chunking(precision, b.n, S.src_n, num_cus, tun, S.splits, S.chunk);
            S.part_base = part_total;
            part_total += (int64_t)S.splits * b.n;
            pairs_ += (int64_t)b.n * S.src_n;
        }
        if (precision == kPrecisionMixed) {
            const double two64 = 18446744073709551616.0;
            const double d2 = S.delta2;
            const double B = std::max({std::ldexp(1.0, -6), 16.0 * delta, 4000.0 * d2});
            S.all_band = (B >= 0.25) ? 1 : 0;
            S.band = S.all_band ? (uint64_t(1) << 62) : (uint64_t)std::ldexp(B, 64);
            S.d2s = d2 * two64 * two64;
            const double kq2 = d2 * two64;
            S.b1 = S.all_band ? 0.0f : (float)(-0.75 * kq2);
            S.b2 = S.all_band ? 0.0f : (float)(0.625 * kq2 * kq2);
        }
    }
    if (precision == kPrecisionMixed) {
        // the far-field weight scale belongs to the source segment (set by
        // the segment that evaluates against it)
        for (int d = 0; d < nseg_; ++d)
            if (segs[d].eval) segs_[segs[d].src].wscale = -0.5 * segs_[d].delta2 * 18446744073709551616.0;
        // scan blocks of every segment that serves as a source
        std::vector<char> is_src(nseg_, 0);
        for (int d = 0; d < nseg_; ++d)
            if (segs[d].eval) is_src[segs[d].src] = 1;
        for (int d = 0; d < nseg_; ++d) {
            if (!is_src[d]) continue;
            const SegDev& S = segs_[d];
            const int first = (int)scan_.size();
            const int cnt = std::max(1, (S.n + kScanItems - 1) / kScanItems);
            for (int c = 0; c < cnt; ++c)
                scan_.push_back(ScanBlock{d, S.off + c * kScanItems,
                                          std::min(S.off + S.n, S.off + (c + 1) * kScanItems),
                                          first, cnt});
        }
    }
    // work items, largest first
    const int tblock = kBlock * tpt_;
    for (int d = 0; d < nseg_; ++d) {
        const SegDev& S = segs_[d];
        if (!S.eval || S.n == 0 || S.src_n == 0) continue;
        for (int tb = S.off; tb < S.off + S.n; tb += tblock)
            for (int s = 0; s < S.splits; ++s) {
                const int sb = S.src_off + s * S.chunk;
                const int se = std::min(S.src_off + S.src_n, sb + S.chunk);
                items_.push_back(WorkItem{d, tb, std::min(S.off + S.n, tb + tblock), sb, se, s});
            }
    }
    std::stable_sort(items_.begin(), items_.end(), [](const WorkItem& a, const WorkItem& b) {
        return (int64_t)(a.te - a.tb) * (a.se - a.sb) > (int64_t)(b.te - b.tb) * (b.se - b.sb);
    });
    // device tables
    const size_t b_segs = align_up(sizeof(SegDev) * nseg_, 256);
    const size_t b_items = align_up(sizeof(WorkItem) * std::max<size_t>(items_.size(), 1), 256);
    const size_t b_scan = align_up(sizeof(ScanBlock) * std::max<size_t>(scan_.size(), 1), 256);
    const size_t b_segof = align_up(sizeof(int32_t) * seg_of.size(), 256);
    FS_TRY(tables_.ensure(b_segs + b_items + b_scan + b_segof + 256, stream));
    d_segs_ = tables_.as<SegDev>(0);
    d_items_ = tables_.as<WorkItem>(b_segs);
    d_scan_ = tables_.as<ScanBlock>(b_segs + b_items);
    d_seg_of_ = tables_.as<int32_t>(b_segs + b_items + b_scan);
    d_counter_ = tables_.as<int>(b_segs + b_items + b_scan + b_segof);
    FS_HIP_CHECK(hipMemcpyAsync(d_segs_, segs_.data(), sizeof(SegDev) * nseg_,
                                hipMemcpyHostToDevice, stream));
    if (!items_.empty())
        FS_HIP_CHECK(hipMemcpyAsync(d_items_, items_.data(), sizeof(WorkItem) * items_.size(),
                                    hipMemcpyHostToDevice, stream));
    if (!scan_.empty())
        FS_HIP_CHECK(hipMemcpyAsync(d_scan_, scan_.data(), sizeof(ScanBlock) * scan_.size(),
                                    hipMemcpyHostToDevice, stream));
    FS_HIP_CHECK(hipMemcpyAsync(d_seg_of_, seg_of.data(), sizeof(int32_t) * seg_of.size(),
                                hipMemcpyHostToDevice, stream));
    zero_int_kernel<<<1, 1, 0, stream>>>(d_counter_);
    FS_HIP_CHECK(hipGetLastError());
    // work space
    size_t o = 0;
    auto take = [&](size_t b) {
        const size_t at = o;
        o += align_up(std::max<size_t>(b, 1), 256);
        return at;
    };
    const size_t n = (size_t)std::max(n_total, 1);
    o_part_ = take(8 * (size_t)std::max<int64_t>(part_total, 1));
    if (precision == kPrecisionMixed) {
        o_key_ = take(8 * n);
        o_skey_ = take(8 * n);
        o_perm_ = take(4 * n);
        o_sort_ = take(radix_sort_bytes(n_total));
        o_hs_ = take(4 * n);
        o_ws_ = take(8 * n);
        o_wf_ = take(4 * n);
        o_pw_ = take(8 * (n + nseg_));
        o_bsum_ = take(8 * std::max<size_t>(scan_.size(), 1));
        o_bsa_ = take(8 * std::max<size_t>(scan_.size(), 1));
        o_tot_ = take(8 * (size_t)nseg_);
    }
    FS_TRY(work_.ensure(o, stream));
    // persistent grid: every item can start at once up to 8 workgroups per CU
    grid_ = (int)std::min<size_t>(items_.size(), (size_t)8 * num_cus);
    // the host tables must outlive the copies
    FS_HIP_CHECK(hipStreamSynchronize(stream));
    return FS_OK;
}

int FieldBatch::run(const double* x, const double* w, double* es, hipStream_t stream) {
    if (n_total_ == 0) return FS_OK;
    char* wk = static_cast<char*>(work_.ptr);
    double* part = reinterpret_cast<double*>(wk + o_part_);
    const int nit = (int)items_.size();
    if (precision_ == kPrecisionF64) {
        if (nit > 0) {
            switch (tpt_) {
                case 1:
                    batch_f64_kernel<1><<<grid_, kBlock, 0, stream>>>(d_items_, nit, d_segs_, x, w,
                                                                       part, d_counter_);
                    break;
                case 4:
                    batch_f64_kernel<4><<<grid_, kBlock, 0, stream>>>(d_items_, nit, d_segs_, x, w,
                                                                       part, d_counter_);
                    break;
                default:
                    batch_f64_kernel<2><<<grid_, kBlock, 0, stream>>>(d_items_, nit, d_segs_, x, w,
                                                                       part, d_counter_);
                    break;
            }
            FS_HIP_CHECK(hipGetLastError());
        }
        finalize_f64_kernel<<<nblocks(n_total_), kBlock, 0, stream>>>(n_total_, d_seg_of_, d_segs_,
                                                                      part, es, d_counter_);
        FS_HIP_CHECK(hipGetLastError());
        return FS_OK;
    }
    uint64_t* key = reinterpret_cast<uint64_t*>(wk + o_key_);
    uint64_t* skey = reinterpret_cast<uint64_t*>(wk + o_skey_);
    int32_t* perm = reinterpret_cast<int32_t*>(wk + o_perm_);
    uint32_t* hs = reinterpret_cast<uint32_t*>(wk + o_hs_);
    double* ws = reinterpret_cast<double*>(wk + o_ws_);
    float* wf = reinterpret_cast<float*>(wk + o_wf_);
    double* pw = reinterpret_cast<double*>(wk + o_pw_);
    double* bsum = reinterpret_cast<double*>(wk + o_bsum_);
    double* bsa = reinterpret_cast<double*>(wk + o_bsa_);
    double* tot = reinterpret_cast<double*>(wk + o_tot_);
    prep_kernel<<<nblocks(n_total_), kBlock, 0, stream>>>(n_total_, d_seg_of_, d_segs_, x, key);
    FS_HIP_CHECK(hipGetLastError());
    FS_TRY(radix_sort_seg(key, n_total_, d_seg_of_, nseg_, skey, perm, wk + o_sort_, stream));
    if (!scan_.empty()) {
        gather_kernel<<<(unsigned)scan_.size(), kBlock, 0, stream>>>(d_scan_, d_segs_, skey, perm,
                                                                     w, hs, ws, wf, bsum, bsa);
        FS_HIP_CHECK(hipGetLastError());
        scan_kernel<<<(unsigned)scan_.size(), kBlock, 0, stream>>>(d_scan_, d_segs_, ws, bsum, bsa,
                                                                   pw, tot);
        FS_HIP_CHECK(hipGetLastError());
    }
    if (nit > 0) {
        switch (tpt_) {
            case 1:
                batch_mixed_kernel<1><<<grid_, kBlock, 0, stream>>>(d_items_, nit, d_segs_, skey, hs,
                                                                     ws, wf, part, d_counter_);
                break;
            case 4:
                batch_mixed_kernel<4><<<grid_, kBlock, 0, stream>>>(d_items_, nit, d_segs_, skey, hs,
                                                                     ws, wf, part, d_counter_);
                break;
            default:
                batch_mixed_kernel<2><<<grid_, kBlock, 0, stream>>>(d_items_, nit, d_segs_, skey, hs,
                                                                     ws, wf, part, d_counter_);
                break;
        }
        FS_HIP_CHECK(hipGetLastError());
    }
    finalize_mixed_kernel<<<nblocks(n_total_), kBlock, 0, stream>>>(
        n_total_, d_seg_of_, d_segs_, skey, perm, ws, pw, tot, part, es, d_counter_);
    FS_HIP_CHECK(hipGetLastError());
    return FS_OK;
}

}  // namespace fs
