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
// tile each lane holds one source's accumulator, which moves one lane on per
// step (DPP wavefront rotation), and at step s reads source (lane - s) mod 64
// from the wavefront's LDS copy of the tile (written twice over, so the
// source sits at slot lane + 64 - s): after 64 steps every lane has met every
// source of the tile and each source's accumulator is back in its own lane;
// the workgroup's wavefronts add their accumulators of a tile through LDS in
// a fixed order. Targets keep their sums in registers
// (plain FP64 over a tile, TwoSum across tiles). Every item writes its row
// partials and its sources' column partials to their own slots, and the
// finalize adds each point's partials in a fixed order, so results are
// deterministic and independent of scheduling.
//
// The column partials take 8 B per point and source block of a chunk, about
// n^2 / (2 * rb) doubles for a whole evaluation (2 GiB at n = 2^20, 32 GiB at
// 2^22). To bound them, the chunks of every row block are evaluated in passes
// of as many chunks as a byte budget holds; after each pass a fold kernel
// adds the pass's partials to a per-point (s, lo) TwoSum state, chunk by
// chunk, so the sequence of additions -- and the result -- does not depend on
// how many passes the budget allows.
#pragma once

#include <algorithm>
#include <cstdint>

#include "fs_device.h"
#include "fs_internal.h"

namespace fs {
namespace sym {

// Wavefronts per workgroup: 4 for one target per lane (the decks below
// 65536 points, where more, smaller row blocks keep the chip busy) and for
// eight (the same 2048-point row blocks as four), else 8.
template <int T> constexpr int waves() { return T == 1 || T == 8 ? 4 : 8; }
template <int T> constexpr int threads() { return 64 * waves<T>(); }
inline int threads_of(int t) { return t == 1 || t == 8 ? threads<1>() : threads<2>(); }
constexpr int kMaxWaves = 8;
// Points from which the pair-once form is the default (below it the row
// blocks are too few to fill the chip and the full sum is faster).
constexpr int64_t kMinPoints = 16384;

// Reference CU count for the item sizes, so a plan (and with it the
// summation order) is a function of the point count and the tuning alone.
constexpr int kPlanCUs = 256;
// Source blocks per item at most: keeps one chunk's column partials at
// 8 * 16 B per point, so the passes below can bound the scratch at any size.
constexpr int kMaxChunkBlocks = 16;
// Default bound on the partials (row, column, carried state) of one
// pair-once evaluation.
constexpr int64_t kDefaultBudget = int64_t(4) << 30;
// Batched decks run their pair-once evaluation in one pass; a deck whose
// single-pass partials would exceed this takes the full sum (about 740k
// points and up).
constexpr int64_t kBatchDeckLimit = int64_t(1) << 30;

struct Plan {
    int n;       // points
    int t;       // targets per lane (1, 2, 4)
    int rb;      // points per row block, threads<t>() * t
    int nb;      // row blocks
    int h;       // (nb - 1) / 2
    int even;    // nb even: blocks b < nb/2 also take the distance nb/2
    int hmax;    // largest distance
    int chb;     // source blocks (distances) per chunk
    int nch;     // chunks per row block
    int nsym;    // nb * nch symmetric items, then nb self items
    int nitems;  // items of the whole evaluation
    int cpp;     // chunks per pass
    int npass;   // passes (> 1 when the partials of all chunks exceed the budget)
};

// Pass k of a plan: the chunks [c0, c0 + cnt) of every row block, and in
// pass 0 also every block with itself. Item ids of a pass: b * cnt + (c - c0)
// for the chunk items, then nb * cnt + b for the self items.
struct Pass {
    int c0, cnt, self, nitems;
};

__device__ __host__ __forceinline__ int hb(const Plan& P, int b) {
    return P.h + ((P.even && b < P.nb / 2) ? 1 : 0);
}

inline Pass pass(const Plan& P, int k) {
    Pass S{};
    S.c0 = k * P.cpp;
    S.cnt = std::max(0, std::min(P.cpp, P.nch - S.c0));
    S.self = k == 0 ? 1 : 0;
    S.nitems = P.nb * S.cnt + (S.self ? P.nb : 0);
    return S;
}

// Bytes of each partial array (per pass; reused by the next pass).
inline int64_t rowpart_elems(const Plan& P) { return ((int64_t)P.nb * P.cpp + P.nb) * P.rb; }
inline int64_t colpart_elems(const Plan& P) { return (int64_t)P.nb * P.cpp * P.chb * P.rb; }
// (s, lo) carried per point from one pass to the next
inline int64_t state_elems(const Plan& P) { return P.npass > 1 ? 2 * (int64_t)P.n : 0; }
inline size_t al256(int64_t elems) { return ((size_t)elems * sizeof(double) + 255) / 256 * 256; }
inline size_t scratch_bytes(const Plan& P) {
    return al256(rowpart_elems(P)) + al256(colpart_elems(P)) + al256(state_elems(P));
}

// budget <= 0: kDefaultBudget. The chunking depends on n and the tuning only;
// the budget decides how many chunks share a pass, which changes no result.
// wide: eight targets per lane by default from 2^18 points (the same
// 2048-point row blocks as four; the standalone sums take it).
inline Plan plan(int64_t n, const FieldTuning& tun, int64_t budget = 0, bool wide = false) {
    Plan P{};
    P.n = (int)n;
    if (tun.targets_per_thread == 1 || tun.targets_per_thread == 2 ||
        tun.targets_per_thread == 4 || tun.targets_per_thread == 8)
        P.t = tun.targets_per_thread;
    else
        P.t = n >= (int64_t(1) << 18) ? (wide ? 8 : 4) : n >= (int64_t(1) << 16) ? 2 : 1;
    P.rb = threads_of(P.t) * P.t;
    P.nb = (int)((n + P.rb - 1) / P.rb);
    P.h = (P.nb - 1) / 2;
    P.even = (P.nb % 2 == 0) ? 1 : 0;
    P.hmax = P.h + P.even;
    // about 32 items per CU (two resident workgroups per CU, 16 rounds) so
    // the last round's imbalance stays small
    const int64_t pairs = (int64_t)P.nb * P.h + (P.even ? P.nb / 2 : 0);
    const int64_t want = int64_t(32) * kPlanCUs;
    int64_t chb = tun.source_splits > 0 ? (P.hmax + tun.source_splits - 1) / tun.source_splits
                                        : (pairs + want - 1) / want;
    chb = std::min<int64_t>(chb, kMaxChunkBlocks);
    chb = std::max<int64_t>(1, std::min<int64_t>(chb, std::max(P.hmax, 1)));
    P.chb = (int)chb;
    P.nch = P.hmax > 0 ? (P.hmax + P.chb - 1) / P.chb : 0;
    P.nsym = P.nb * P.nch;
    P.nitems = P.nsym + P.nb;
    // passes: as many chunks per pass as the budget holds (at least one)
    const int64_t lim = budget > 0 ? budget : kDefaultBudget;
    const int64_t blk = (int64_t)P.nb * P.rb * (int64_t)sizeof(double);
    const int64_t per_chunk = blk * (1 + P.chb);
    if (P.nch == 0 || blk + per_chunk * P.nch <= lim) {
        P.cpp = std::max(P.nch, 1);
    } else {
        const int64_t avail = lim - blk - 16 * (int64_t)P.n;
        P.cpp = (int)std::max<int64_t>(1, std::min<int64_t>(P.nch, avail / per_chunk));
    }
    P.npass = P.nch > 0 ? (P.nch + P.cpp - 1) / P.cpp : 1;
    return P;
}

// Partials of a batched deck of n points evaluated in one pass (0 where the
// deck takes the full sum instead).
inline size_t batch_bytes(int64_t n, const FieldTuning& tun) {
    if (n < 1 || n > INT32_MAX || !use_sym(n, n, true, tun)) return 0;
    const size_t b = scratch_bytes(plan(n, tun, INT64_MAX));
    return b <= (size_t)kBatchDeckLimit ? b : 0;
}

// Kernel evaluations of a whole evaluation: block pairs at distance >= 1
// once, each block with itself in full.
inline int64_t evals(const Plan& P) {
    const int64_t bp = (int64_t)P.nb * P.h + (P.even ? P.nb / 2 : 0) + P.nb;
    return bp * P.rb * P.rb;
}

// x W - M with the product rounded on its own (not fused), as the tiles' sums
// round theirs: a lone point's linear part is then exactly zero.
__device__ __forceinline__ double lin_part(double x, double W, double M) {
#pragma clang fp contract(off)
    return x * W - M;
}

// one lane on across the wavefront (DPP wave_ror:1), 64-bit
__device__ __forceinline__ double wave_rotate(double v) {
    const long long u = __builtin_bit_cast(long long, v);
    const int lo = __builtin_amdgcn_mov_dpp((int)u, 0x13C, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(u >> 32), 0x13C, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned)lo);
}

// One item (row block and source chunk, or a block with itself) of pass S of
// the plan P over the points x[0..n), w[0..n). id: the item's index within
// the pass. red, stage: the workgroup's LDS, [2][kMaxWaves][64] doubles and
// [kMaxWaves][128] (x, w) pairs. tsum: the tiles' sums (kSorted). Called by
// all threads<T>() threads of a workgroup.
// kOpen: the open-boundary kernel z / sqrt(z^2 + delta2) with z = x_i - x_j
// (E_MQ_DirectSum_openbcs, src/FieldStructure.cpp:166-191; also odd in z;
// invL unused) instead of the periodic one.
// kSorted: x is sorted ascending. A row block and a 64-source tile then span
// the displacements [x_first - x'_last, x_last - x'_first]; when that range
// sits inside one period's window (|d| < L/2, or d - L, or d + L, with a
// margin), every pair of the tile has the same wrap, and the tile computes
// z = (x_i - x_j) / L with the tile's sources moved by one period beforehand
// (forward where d > L/2, back where d < -L/2): no rint per pair. Other tiles (those
// straddling the half-period) take the rint form. For |d| < L/2 both forms
// round alike; K is odd and vanishes at z = +-1/2, so a tile whose range
// touches the half-period within the margin loses nothing to either choice.
// Such a whole-wrap tile is evaluated in x units, K = d / sqrt(d^2 + eps^2) -
// C_x d with d = x_i - x_j (shifted source), eps = delta L, C_x = C / L (the
// same K: z / sqrt(z^2 + delta^2) = d / sqrt(d^2 + eps^2)), and its linear part
// for all 64 sources at once: sum_j w_j C_x d_ij = C_x (x_i W - M) from the
// tile's sums W = sum w_j, M = sum w_j x_j (tsum, per 64-point tile of the
// sorted points), and for the sources C_x (M_r - x_j W_r) from the row block's
// sums.
template <int T, bool kOpen = false, bool kSorted = false>
__device__ __forceinline__ void item(const double* __restrict__ x, const double* __restrict__ w,
                                     const Plan& P, const Pass& S, int id, double L, double invL,
                                     double delta2, double* __restrict__ rowpart,
                                     double* __restrict__ colpart,
                                     double (*red)[kMaxWaves][64], double2 (*stage)[128],
                                     const double2* __restrict__ tsum = nullptr) {
    constexpr int kWaves = waves<T>();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int rb = threads<T>() * T;
    int b, d0, d1;
    const bool self = id >= P.nb * S.cnt;
    if (!self) {
        b = id / S.cnt;
        const int c = S.c0 + (id - b * S.cnt);
        d0 = 1 + c * P.chb;
        d1 = min(hb(P, b), (c + 1) * P.chb) + 1;
    } else {
        b = id - P.nb * S.cnt;
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
    // the row block's first and last points, and (whole-wrap tiles) eps^2,
    // C_x and the row block's sums W_r, M_r (kSorted)
    double r0 = 0.0, r1 = 0.0, eps2 = 0.0, Cx = 0.0, Wr = 0.0, Mr = 0.0;
    if constexpr (kSorted) {
        const int pa = b * rb;
        r0 = x[pa];
        r1 = x[min(pa + rb, n) - 1];
        if constexpr (!kOpen) {
            eps2 = delta2 * (L * L);
            Cx = dev::rsq_f64(__builtin_fma(0.5 * L, 0.5 * L, eps2));
            for (int k = 0; k < rb / 64 && pa + 64 * k < n; ++k) {
                const double2 s = tsum[pa / 64 + k];
                Wr += s.x;
                Mr += s.y;
            }
        }
    }
    int tile = 0;
    for (int d = d0; d < d1; ++d) {
        const int B = b + d < P.nb ? b + d : b + d - P.nb;
        for (int jt = 0; jt < rb / 64; ++jt, ++tile) {
            const int q = B * rb + jt * 64 + lane;
            const double xj = q < n ? x[q] : 0.0;
            const double wj = q < n ? w[q] : 0.0;
            double accj = 0.0;
            double src_lin = 0.0;  // the sources' linear part (whole-wrap tiles)
            // the tile's wrap: 0 (none), +1 (targets back one period), -1
            // (sources back one period), 2 (per pair)
            int wrap = 2;
            if constexpr (kSorted && !kOpen) {
                const int q0 = B * rb + jt * 64;
                if (q0 < n) {
                    const double lo = (r0 - x[min(q0 + 63, n - 1)]) * invL;
                    const double hi = (r1 - x[q0]) * invL;
                    constexpr double m = 1e-9;
                    if (hi < 0.5 - m && lo > -0.5 + m) wrap = 0;
                    else if (lo > 0.5 + m && hi < 1.5 - m) wrap = 1;
                    else if (hi < -0.5 - m && lo > -1.5 + m) wrap = -1;
                }
            }
            // the tile's sources (one period on or back for a whole-wrap tile)
            // in the wavefront's LDS copy
            const double sh = wrap == 1 ? L : (wrap == -1 ? -L : 0.0);
            const double xs = xj + sh;
            stage[wave][lane] = make_double2(xs, wj);
            stage[wave][lane + 64] = make_double2(xs, wj);
            __syncthreads();
            if (wrap != 2) {
                if constexpr (kSorted && !kOpen) {
#pragma unroll 1
                    for (int s0 = 0; s0 < 64; s0 += 4) {
                        double2 src[4];
#pragma unroll
                        for (int u = 0; u < 4; ++u) src[u] = stage[wave][lane + 64 - (s0 + u)];
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            const double2 sj = src[u];
#pragma unroll
                            for (int t = 0; t < T; ++t) {
                                const double dd = xi[t] - sj.x;
                                const double k = dd * dev::rsq_f64(__builtin_fma(dd, dd, eps2));
                                acc[t] = __builtin_fma(sj.y, k, acc[t]);
                                accj = __builtin_fma(nwi[t], k, accj);
                            }
                            accj = wave_rotate(accj);
                        }
                    }
                    // the linear part -C_x d of the tile's pairs
                    const double2 ts = tsum[(B * rb) / 64 + jt];
                    const double Mt = __builtin_fma(sh, ts.x, ts.y);
#pragma unroll
                    for (int t = 0; t < T; ++t)
                        acc[t] = __builtin_fma(-Cx, lin_part(xi[t], ts.x, Mt), acc[t]);
                    src_lin = -Cx * lin_part(xs, Wr, Mr);
                }
            } else {
#pragma unroll 1
                for (int s0 = 0; s0 < 64; s0 += 4) {
                    double2 src[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) src[u] = stage[wave][lane + 64 - (s0 + u)];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const double2 sj = src[u];
#pragma unroll
                        for (int t = 0; t < T; ++t) {
                            double k;
                            if constexpr (kOpen) {
                                const double z = xi[t] - sj.x;
                                k = z * dev::rsq_f64(__builtin_fma(z, z, delta2));
                            } else {
                                const double v = (xi[t] - sj.x) * invL;
                                const double z = v - __builtin_rint(v);
                                const double y = dev::rsq_f64(__builtin_fma(z, z, delta2));
                                k = z * (y - C);
                            }
                            acc[t] = __builtin_fma(sj.y, k, acc[t]);
                            accj = __builtin_fma(nwi[t], k, accj);
                        }
                        accj = wave_rotate(accj);
                    }
                }
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
                    colpart[((int64_t)id * P.chb + (d - d0)) * rb + jt * 64 + lane] =
                        sum + src_lin;
                }
            }
        }
    }
#pragma unroll
    for (int t = 0; t < T; ++t)
        rowpart[(int64_t)id * rb + wave * 64 * T + t * 64 + lane] = hi[t] + lo[t];
}

// Adds point p's partials of pass S to (s, lo) in a fixed order: its self
// item's row partial (pass 0), then chunk by chunk its row block's row
// partial of the chunk and the column partials of the items of that chunk
// that reached its block, by increasing distance. TwoSum throughout. The
// sequence of additions is the same however the chunks are cut into passes,
// so every budget gives the same result bit for bit.
__device__ __forceinline__ void fold(const Plan& P, const Pass& S, int p,
                                     const double* __restrict__ rowpart,
                                     const double* __restrict__ colpart, double& s, double& lo) {
    const int b = p / P.rb, loc = p - b * P.rb;
    if (S.self) dev::two_sum_acc(s, lo, rowpart[((int64_t)P.nb * S.cnt + b) * P.rb + loc]);
    for (int k = 0; k < S.cnt; ++k) {
        const int c = S.c0 + k;
        dev::two_sum_acc(s, lo, rowpart[((int64_t)b * S.cnt + k) * P.rb + loc]);
        const int dl = min((c + 1) * P.chb, P.hmax);
        for (int d = c * P.chb + 1; d <= dl; ++d) {
            const int b2 = b - d >= 0 ? b - d : b - d + P.nb;
            if (d > hb(P, b2)) continue;
            dev::two_sum_acc(
                s, lo,
                colpart[(((int64_t)b2 * S.cnt + k) * P.chb + (d - 1 - c * P.chb)) * P.rb + loc]);
        }
    }
}

// Pass S at point p: (s, lo) from the carried state (zero in pass 0), the
// pass's partials added, written back to the state unless this is the last
// pass. Returns (s + lo) / 2, which is E at p after the last pass.
__device__ __forceinline__ double pass_point(const Plan& P, const Pass& S, int p, bool last,
                                             const double* __restrict__ rowpart,
                                             const double* __restrict__ colpart,
                                             double* __restrict__ state) {
    double s = 0.0, lo = 0.0;
    if (!S.self) {
        s = state[2 * (int64_t)p];
        lo = state[2 * (int64_t)p + 1];
    }
    fold(P, S, p, rowpart, colpart, s, lo);
    if (!last) {
        state[2 * (int64_t)p] = s;
        state[2 * (int64_t)p + 1] = lo;
    }
    return 0.5 * (s + lo);
}

template <int T, bool kOpen, bool kSorted>
__global__ __launch_bounds__(threads<T>(), T == 8 ? 4 : 1) void pass_kernel(
    const double* __restrict__ x, const double* __restrict__ w, Plan P, Pass S, double L,
    double invL, double delta2, double* __restrict__ rowpart, double* __restrict__ colpart,
    const double2* __restrict__ tsum) {
    __shared__ double red[2][kMaxWaves][64];
    __shared__ double2 stage[kMaxWaves][128];
    item<T, kOpen, kSorted>(x, w, P, S, blockIdx.x, L, invL, delta2, rowpart, colpart, red, stage,
                            tsum);
}

// kOpen: E = (s + lo) / 2 - x_p (the open-boundary form's uniform background).
// perm (sorted points): E of sorted point p goes to es[perm[p]].
template <bool kOpen>
__global__ __launch_bounds__(256) void fold_kernel(Plan P, Pass S, int last,
                                                   const double* __restrict__ rowpart,
                                                   const double* __restrict__ colpart,
                                                   double* __restrict__ state,
                                                   const double* __restrict__ x,
                                                   const int32_t* __restrict__ perm,
                                                   double* __restrict__ es) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= P.n) return;
    const double e = pass_point(P, S, p, last != 0, rowpart, colpart, state);
    if (last) es[perm ? perm[p] : p] = kOpen ? e - x[p] : e;
}

template <bool kOpen, bool kSorted>
inline hipError_t launch_passes(const double* x, const double* w, const Plan& P, double L,
                                double invL, double delta2, void* scratch,
                                const int32_t* perm, const double2* tsum, double* es,
                                hipStream_t stream) {
    double* rowpart = static_cast<double*>(scratch);
    double* colpart = reinterpret_cast<double*>(static_cast<char*>(scratch) +
                                                al256(rowpart_elems(P)));
    double* state = reinterpret_cast<double*>(static_cast<char*>(scratch) +
                                              al256(rowpart_elems(P)) +
                                              al256(colpart_elems(P)));
    for (int k = 0; k < P.npass; ++k) {
        const Pass S = pass(P, k);
        switch (P.t) {
            case 8:
                pass_kernel<8, kOpen, kSorted><<<S.nitems, threads<8>(), 0, stream>>>(
                    x, w, P, S, L, invL, delta2, rowpart, colpart, tsum);
                break;
            case 4:
                pass_kernel<4, kOpen, kSorted><<<S.nitems, threads<4>(), 0, stream>>>(
                    x, w, P, S, L, invL, delta2, rowpart, colpart, tsum);
                break;
            case 2:
                pass_kernel<2, kOpen, kSorted><<<S.nitems, threads<2>(), 0, stream>>>(
                    x, w, P, S, L, invL, delta2, rowpart, colpart, tsum);
                break;
            default:
                pass_kernel<1, kOpen, kSorted><<<S.nitems, threads<1>(), 0, stream>>>(
                    x, w, P, S, L, invL, delta2, rowpart, colpart, tsum);
                break;
        }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        fold_kernel<kOpen><<<(P.n + 255) / 256, 256, 0, stream>>>(
            P, S, k == P.npass - 1 ? 1 : 0, rowpart, colpart, state, x, perm, es);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// Every pass of plan P over x[0..n), w[0..n) on `stream`, E into es, in the
// points' own order. scratch: scratch_bytes(P) bytes.
template <bool kOpen>
inline hipError_t run(const double* x, const double* w, const Plan& P, double L, double invL,
                      double delta2, void* scratch, double* es, hipStream_t stream) {
    return launch_passes<kOpen, false>(x, w, P, L, invL, delta2, scratch, nullptr, nullptr, es,
                                       stream);
}

// ---- the periodic sum on points sorted by x -------------------------------

// 64-point tiles of n sorted points.
__host__ __device__ inline int64_t tiles(int64_t n) { return (n + 63) / 64; }

// Work space of the sort in front of the passes: keys, sorted keys, the
// permutation, sorted x and w, the tiles' sums, and the radix sort's own.
inline size_t sort_bytes(int64_t n) {
    return 4 * al256(n) + (((size_t)n * 4 + 255) / 256 * 256) + al256(2 * tiles(n)) +
           radix_sort_bytes((int)n);
}

// Order-preserving image of a double (+0 and -0 one value).
__device__ __forceinline__ uint64_t order_key(double v) {
    if (v == 0.0) v = 0.0;
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

namespace {  // one copy per translation unit that launches them

__global__ __launch_bounds__(256) void key_kernel(const double* __restrict__ x, int n,
                                                  uint64_t* __restrict__ key) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) key[i] = order_key(x[i]);
}

__global__ __launch_bounds__(256) void gather_kernel(const double* __restrict__ x,
                                                     const double* __restrict__ w,
                                                     const int32_t* __restrict__ perm, int n,
                                                     double* __restrict__ xs,
                                                     double* __restrict__ ws) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    const int i = perm[k];
    xs[k] = x[i];
    ws[k] = w[i];
}

// (sum w, sum w x) of each 64-point tile of the sorted points, in point order.
__global__ __launch_bounds__(256) void tile_sums_kernel(const double* __restrict__ xs,
                                                        const double* __restrict__ ws, int n,
                                                        double2* __restrict__ tsum) {
#pragma clang fp contract(off)
    const int k = blockIdx.x * 256 + threadIdx.x;
    if ((int64_t)k >= tiles(n)) return;
    double W = 0.0, M = 0.0;
    const int e = min(64 * k + 64, n);
    for (int p = 64 * k; p < e; ++p) {
        W += ws[p];
        M += ws[p] * xs[p];
    }
    tsum[k] = make_double2(W, M);
}

}  // namespace

// The periodic pair-once sum with the points sorted by x first (stable: equal
// x keep their order), so that nearly every tile pair wraps as a whole (see
// item); E is written back in the points' own order. scratch: scratch_bytes(P)
// + sort_bytes(n) bytes.
inline int run_sorted(const double* x, const double* w, const Plan& P, double L, double invL,
                      double delta2, void* scratch, double* es, hipStream_t stream) {
    const int n = P.n;
    char* c = static_cast<char*>(scratch) + scratch_bytes(P);
    uint64_t* key = reinterpret_cast<uint64_t*>(c);
    c += al256(n);
    uint64_t* skey = reinterpret_cast<uint64_t*>(c);
    c += al256(n);
    double* xs = reinterpret_cast<double*>(c);
    c += al256(n);
    double* ws = reinterpret_cast<double*>(c);
    c += al256(n);
    int32_t* perm = reinterpret_cast<int32_t*>(c);
    c += ((size_t)n * 4 + 255) / 256 * 256;
    double2* tsum = reinterpret_cast<double2*>(c);
    c += al256(2 * tiles(n));
    const unsigned g = (unsigned)((n + 255) / 256);
    key_kernel<<<g, 256, 0, stream>>>(x, n, key);
    if (hipGetLastError() != hipSuccess) return fail(FS_E_HIP, "pair-once sort: key launch failed");
    FS_TRY(radix_sort_full(key, n, skey, perm, c, stream));
    gather_kernel<<<g, 256, 0, stream>>>(x, w, perm, n, xs, ws);
    if (hipGetLastError() != hipSuccess)
        return fail(FS_E_HIP, "pair-once sort: gather launch failed");
    tile_sums_kernel<<<(unsigned)((tiles(n) + 255) / 256), 256, 0, stream>>>(xs, ws, n, tsum);
    if (hipGetLastError() != hipSuccess)
        return fail(FS_E_HIP, "pair-once sort: tile sums launch failed");
    const hipError_t e = launch_passes<false, true>(xs, ws, P, L, invL, delta2, scratch, perm,
                                                    tsum, es, stream);
    if (e != hipSuccess) return fail(FS_E_HIP, "pair-once sum: %s", hipGetErrorString(e));
    return FS_OK;
}

}  // namespace sym
}  // namespace fs
