// FP64 periodic direct sum with the targets being the sources, each pair
// evaluated once. Shared by the single-call entry point (fs_field.hip) and
// the batched step loop (fs_field_batch.hip). Internal to libfarrsight_hip.
//
// K(d) = z (1/sqrt(z^2 + delta^2) - C) is odd in d (z = d/L - rint(d/L), and
// rint rounds half to even, which is symmetric), so the pair (i, j) gives
// w_j K to target i and -w_i K to target j: half the kernel evaluations of
// the full sum (reference src/FieldStructure.cpp:106-156 evaluates every
// ordered pair). Points are cut into row blocks of W * 64 * T points (W
// wavefronts per workgroup: 4 at one target per lane, else 8).
// A workgroup takes one row block's targets (T per lane) and a chunk of the
// source blocks at cyclic block distances d = 1..H_b from it (H = (nb - 1)/2,
// plus the distance nb/2 for the first half of the blocks when nb is even:
// every unordered pair of blocks exactly once); a block's pairs with itself
// are one more item, summed in full without the symmetry. Within a 64-source
// tile each lane holds one source (x_j, w_j and its accumulator) and the three
// move one lane on per step (DPP wavefront rotation), so after 64 steps every
// lane has met every source of the tile and each source's accumulator is back
// in its own lane; the workgroup's wavefronts add their accumulators of a
// tile through LDS in a fixed order. Targets keep their sums in registers
// (plain FP64 over a tile, TwoSum across tiles). Every item writes its row
// partials and its sources' column partials to their own slots, and the
// finalize adds each point's partials in a fixed order, so results are
// deterministic and independent of scheduling.
#pragma once

#include <algorithm>
#include <cstdint>

#include "fs_device.h"
#include "fs_internal.h"

namespace fs {
namespace sym {

// Wavefronts per workgroup: 4 for one target per lane (the decks below
// 65536 points, where more, smaller row blocks keep the chip busy), else 8.
template <int T> constexpr int waves() { return T == 1 ? 4 : 8; }
template <int T> constexpr int threads() { return 64 * waves<T>(); }
constexpr int kMaxWaves = 8;
// Points from which the pair-once form is the default (below it the row
// blocks are too few to fill the chip and the full sum is faster).
constexpr int64_t kMinPoints = 16384;

struct Plan {
    int n;       // points
    int t;       // targets per lane (1, 2, 4)
    int rb;      // points per row block, threads<t>() * t
    int nb;      // row blocks
    int h;       // (nb - 1) / 2
    int even;    // nb even: blocks b < nb/2 also take the distance nb/2
    int hmax;    // largest distance
    int chb;     // source blocks per item
    int nch;     // items per row block
    int nsym;    // nb * nch symmetric items, then nb self items
    int nitems;
};

__device__ __host__ __forceinline__ int hb(const Plan& P, int b) {
    return P.h + ((P.even && b < P.nb / 2) ? 1 : 0);
}

inline Plan plan(int64_t n, int num_cus, const FieldTuning& tun) {
    Plan P{};
    P.n = (int)n;
    if (tun.targets_per_thread == 1 || tun.targets_per_thread == 2 ||
        tun.targets_per_thread == 4)
        P.t = tun.targets_per_thread;
    else
        P.t = n >= (int64_t(1) << 18) ? 4 : n >= (int64_t(1) << 16) ? 2 : 1;
    P.rb = (P.t == 1 ? threads<1>() : P.t == 2 ? threads<2>() : threads<4>()) * P.t;
    P.nb = (int)((n + P.rb - 1) / P.rb);
    P.h = (P.nb - 1) / 2;
    P.even = (P.nb % 2 == 0) ? 1 : 0;
    P.hmax = P.h + P.even;
    // about 32 items per CU (two resident workgroups per CU, 16 rounds) so
    // the last round's imbalance stays small
    const int64_t pairs = (int64_t)P.nb * P.h + (P.even ? P.nb / 2 : 0);
    const int64_t want = int64_t(32) * num_cus;
    int64_t chb = tun.source_splits > 0 ? (P.hmax + tun.source_splits - 1) / tun.source_splits
                                        : (pairs + want - 1) / want;
    chb = std::max<int64_t>(1, std::min<int64_t>(chb, std::max(P.hmax, 1)));
    P.chb = (int)chb;
    P.nch = P.hmax > 0 ? (P.hmax + P.chb - 1) / P.chb : 0;
    P.nsym = P.nb * P.nch;
    P.nitems = P.nsym + P.nb;
    return P;
}

// Partial slots (doubles): one row partial per point of a row block per item,
// one column partial per point of a source block per item and distance.
inline int64_t rowpart_elems(const Plan& P) { return (int64_t)P.nitems * P.rb; }
inline int64_t colpart_elems(const Plan& P) { return (int64_t)P.nsym * P.chb * P.rb; }

// Kernel evaluations of one pass: block pairs at distance >= 1 once, each
// block with itself in full.
inline int64_t evals(const Plan& P) {
    const int64_t bp = (int64_t)P.nb * P.h + (P.even ? P.nb / 2 : 0) + P.nb;
    return bp * P.rb * P.rb;
}

// one lane on across the wavefront (DPP wave_ror:1), 64-bit
__device__ __forceinline__ double wave_rotate(double v) {
    const long long u = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_mov_dpp((int)u, 0x13C, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(u >> 32), 0x13C, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned)lo);
}

// One item (row block and source chunk, or a block with itself) of the plan
// P over the points x[0..n), w[0..n). red: the workgroup's LDS,
// [2][kMaxWaves][64] doubles. Called by all threads<T>() threads of a workgroup.
// kOpen: the open-boundary kernel z / sqrt(z^2 + delta2) with z = x_i - x_j
// (E_MQ_DirectSum_openbcs, src/FieldStructure.cpp:166-191; also odd in z;
// invL unused) instead of the periodic one.
template <int T, bool kOpen = false>
__device__ __forceinline__ void item(const double* __restrict__ x, const double* __restrict__ w,
                                     const Plan& P, int id, double invL, double delta2,
                                     double* __restrict__ rowpart, double* __restrict__ colpart,
                                     double (*red)[kMaxWaves][64]) {
    constexpr int kWaves = waves<T>();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int rb = threads<T>() * T;
    int b, d0, d1;
    const bool self = id >= P.nsym;
    if (!self) {
        b = id / P.nch;
        const int c = id - b * P.nch;
        d0 = 1 + c * P.chb;
        d1 = min(hb(P, b), (c + 1) * P.chb) + 1;
    } else {
        b = id - P.nsym;
        d0 = 0;
        d1 = 1;
    }
    const int n = P.n;
    const int p0 = b * rb + wave * 64 * T + lane;
    double xi[T], nwi[T], acc[T], hi[T], lo[T];
#pragma unroll
    for (int t = 0; t < T; ++t) {
        const int p = p0 + t * 64;
        xi[t] = p < n ? x[p] : 0.0;
        nwi[t] = p < n ? -w[p] : 0.0;
        acc[t] = 0.0;
        hi[t] = 0.0;
        lo[t] = 0.0;
    }
    const double C = kOpen ? 0.0 : dev::rsq_f64(__builtin_fma(0.5, 0.5, delta2));
    int tile = 0;
    for (int d = d0; d < d1; ++d) {
        const int B = b + d < P.nb ? b + d : b + d - P.nb;
        for (int jt = 0; jt < rb / 64; ++jt, ++tile) {
            const int q = B * rb + jt * 64 + lane;
            double xj = q < n ? x[q] : 0.0;
            double wj = q < n ? w[q] : 0.0;
            double accj = 0.0;
#pragma unroll 2
            for (int s = 0; s < 64; ++s) {
#pragma unroll
                for (int t = 0; t < T; ++t) {
                    double k;
                    if constexpr (kOpen) {
                        const double z = xi[t] - xj;
                        k = z * dev::rsq_f64(__builtin_fma(z, z, delta2));
                    } else {
                        const double u = (xi[t] - xj) * invL;
                        const double z = u - __builtin_rint(u);
                        const double y = dev::rsq_f64(__builtin_fma(z, z, delta2));
                        k = z * (y - C);
                    }
                    acc[t] = __builtin_fma(wj, k, acc[t]);
                    accj = __builtin_fma(nwi[t], k, accj);
                }
                xj = wave_rotate(xj);
                wj = wave_rotate(wj);
                accj = wave_rotate(accj);
            }
#pragma unroll
            for (int t = 0; t < T; ++t) {
                dev::two_sum_acc(hi[t], lo[t], acc[t]);
                acc[t] = 0.0;
            }
            if (!self) {
                red[tile & 1][wave][lane] = accj;
                __syncthreads();
                if (wave == (tile & (kWaves - 1))) {
                    double sum = red[tile & 1][0][lane];
#pragma unroll
                    for (int v = 1; v < kWaves; ++v) sum += red[tile & 1][v][lane];
                    colpart[((int64_t)id * P.chb + (d - d0)) * rb + jt * 64 + lane] = sum;
                }
            }
        }
    }
#pragma unroll
    for (int t = 0; t < T; ++t)
        rowpart[(int64_t)id * rb + wave * 64 * T + t * 64 + lane] = hi[t] + lo[t];
}

// E at point p: its self item's row partial, its row block's items in chunk
// order, then the column partials of the items that reached its block, by
// increasing distance; TwoSum throughout, halved at the end.
__device__ __forceinline__ double point(const Plan& P, int p, const double* __restrict__ rowpart,
                                        const double* __restrict__ colpart) {
    const int b = p / P.rb, loc = p - b * P.rb;
    double s = 0.0, lo = 0.0;
    dev::two_sum_acc(s, lo, rowpart[(int64_t)(P.nsym + b) * P.rb + loc]);
    for (int c = 0; c < P.nch; ++c)
        dev::two_sum_acc(s, lo, rowpart[(int64_t)(b * P.nch + c) * P.rb + loc]);
    for (int d = 1; d <= P.hmax; ++d) {
        const int b2 = b - d >= 0 ? b - d : b - d + P.nb;
        if (d > hb(P, b2)) continue;
        const int id2 = b2 * P.nch + (d - 1) / P.chb;
        dev::two_sum_acc(s, lo, colpart[((int64_t)id2 * P.chb + (d - 1) % P.chb) * P.rb + loc]);
    }
    return 0.5 * (s + lo);
}

}  // namespace sym
}  // namespace fs
