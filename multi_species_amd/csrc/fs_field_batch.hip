// Batched periodic regularized direct sum (gfx950): many independent decks,
// each its own set of targets and sources, in one persistent launch.
//
// The field of each deck is the reference's E_MQ_DirectSum
// (src/FieldStructure.cpp:106-156) over that deck's points. The work of all
// decks is cut into items of (deck, block of 256*TPT targets, chunk of
// sources; TPT is 4 in mixed precision); a persistent kernel sized to the device pulls items from an
// atomic counter, largest first, so decks of very different sizes (2k to
// 500k points) share the chip without a launch per deck. Every item writes
// its own partial sum; a finalize kernel adds the partials of each target in
// chunk order. The chunking of a deck depends only on that deck's size and
// the device (not on the rest of the batch or on TPT), so a deck's field is
// bitwise the same alone and in any mix.
//
// Two precisions:
//  * FP64: per pair exactly fs_field.hip's emq_direct_f64_kernel: d = x_i - x_j,
//    z = d/L - rint, refined rsqrt, TwoSum over 256-source runs of each chunk,
//    halved at the end (the chunking, and so the last bits of the sums, can
//    differ from the single-call kernel's).
//  * mixed: the decomposition of fs_field_mixed.hip's comment (FP64 step and
//    neutralizing part from sorted prefix sums; group and tile expansions,
//    FP32 beyond the band and FP64 within it; pairs where no expansion
//    applies; compensated FP64 accumulation; the accuracy governor), with
//    positions sorted per deck by the graph-safe radix sort of fs_radix.hip.
//
// run() launches kernels only (no allocation, memset or synchronization), so
// the caller can capture it in a HIP graph.
#include <algorithm>
#include <cmath>
#include <cstring>
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
constexpr int kSortShift = 40;  // keys are sorted by bits 40..63 (fs_radix.hip)
constexpr uint64_t kLowSort = (uint64_t(1) << kSortShift) - 1;
constexpr int kGroupSrc = 4 * kTile;   // mixed: sources per group
// mixed: far-field units above the group (512 and 2048 sorted sources), each
// with its own span and FP32 expansion; a source chunk is a whole number of
// the largest, so every unit lies inside one chunk
constexpr int kLevels = 2;
__device__ __host__ constexpr int unit_src(int level) { return kGroupSrc << (2 * (level + 1)); }
// Unit levels used for a source segment of n points (from its size alone,
// so a deck's results do not depend on its batch): small decks keep finer
// source chunks (more work items) instead of the larger units. The mixed
// chunk unit is the largest unit in use, so no unit straddles two chunks.
inline int unit_levels(int64_t n) { return n >= (int64_t(1) << 16) ? 2 : n >= (int64_t(1) << 13) ? 1 : 0; }
inline int chunk_unit_src(int64_t n) { return kGroupSrc << (2 * unit_levels(n)); }
// mixed: targets per thread, fixed, since each wavefront chooses each source
// range's path for its 64 * kMixedTpt consecutive sorted targets: a tuned
// value would make results depend on timing
constexpr int kMixedTpt = 4;
constexpr int kXShift = 4;             // tiles expand when their sources lie within band/16
constexpr int kMoments = 8;            // moments of a tile's expansion
constexpr int kCoef = kMoments + 4;    // its polynomial coefficients (TileRec::c)
constexpr int kMoments64 = 13;         // the probe's FP64 reference: moments,
constexpr int kTerms64 = 7;            // terms of the series in d^2/z^2,
constexpr int kCoef64 = 2 * kTerms64 + kMoments64 - 2;  // and coefficients (TileRec64::c)
constexpr int kNearTerms = 12;         // near-band FP64 expansion: Legendre terms
constexpr int kLadder = kGovVals - 1;  // governor: candidate bands
constexpr int kProbeMax = 128;         // governor: probed targets per segment
constexpr double kGovTarget = 4e-9;    // governor: default error target (relative to max |E|)

__device__ __forceinline__ int pull_item(int* counter) {
    __shared__ int s_item;
    if (threadIdx.x == 0) s_item = atomicAdd(counter, 1);
    __syncthreads();
    const int it = __builtin_amdgcn_readfirstlane(s_item);
    __syncthreads();
    return it;
}

// ---- FP64 ------------------------------------------------------------------

// One (deck, target block, source split) item of the FP64 full sum: the
// periodic kernel z (1/sqrt(z^2 + delta^2) - C), z wrapped, or with kOpen the
// open-boundary kernel z / sqrt(z^2 + eps^2), z = x_i - x_j; runs of kSub
// sources summed in FP64, the runs added with TwoSum.
template <int TPT, bool kOpen>
__device__ __forceinline__ void f64_item(const WorkItem& W, const SegDev& S,
                                         const double* __restrict__ x,
                                         const double* __restrict__ w, double* __restrict__ part) {
    const double invL = S.invL, delta2 = S.delta2;
    const double C = kOpen ? 0.0 : rsq_f64(__builtin_fma(0.5, 0.5, delta2));
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
                if constexpr (kOpen) {
                    const double z = xi[t] - xj;
                    const double y = rsq_f64(__builtin_fma(z, z, delta2));
                    acc[t] = __builtin_fma(wj * z, y, acc[t]);
                } else {
                    const double s = (xi[t] - xj) * invL;
                    const double z = s - __builtin_rint(s);
                    const double y = rsq_f64(__builtin_fma(z, z, delta2));
                    acc[t] = __builtin_fma(wj * z, y - C, acc[t]);
                }
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

template <int TPT>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8))) void batch_f64_kernel(const WorkItem* __restrict__ items,
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
        if (S.open) f64_item<TPT, true>(W, S, x, w, part);
        else f64_item<TPT, false>(W, S, x, w, part);
    }
}

// Pair-once segments (fs_sym.h) of one targets-per-lane value: workgroup
// blockIdx.x belongs to the last table entry whose item_base is not above it.
template <int T>
__global__ __launch_bounds__(sym::threads<T>()) void batch_sym_kernel(
    const SymSeg* __restrict__ tab, int first, int count, const double* __restrict__ x,
    const double* __restrict__ w, double* __restrict__ sympart) {
    __shared__ double red[2][sym::kMaxWaves][64];
    __shared__ double2 stage[sym::kMaxWaves][128];
    int lo = first, hi = first + count - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tab[mid].item_base <= (int)blockIdx.x) lo = mid;
        else hi = mid - 1;
    }
    const SymSeg E = tab[lo];
    if (E.open)
        sym::item<T, true>(x + E.off, w + E.off, E.P, E.S, (int)blockIdx.x - E.item_base, 0.0,
                           E.invL, E.delta2, sympart + E.rbase, sympart + E.cbase, red, stage);
    else
        sym::item<T, false>(x + E.off, w + E.off, E.P, E.S, (int)blockIdx.x - E.item_base, 0.0,
                            E.invL, E.delta2, sympart + E.rbase, sympart + E.cbase, red, stage);
}

// A point's partials in a fixed order; open-boundary segments subtract the
// point's own x (E_MQ_DirectSum_openbcs's uniform background).
__global__ __launch_bounds__(kBlock) void finalize_f64_kernel(
    int n_total, const int32_t* __restrict__ seg_of, const SegDev* __restrict__ segs,
    const double* __restrict__ part, const SymSeg* __restrict__ symsegs,
    const double* __restrict__ sympart, const double* __restrict__ x, double* __restrict__ es,
    int* __restrict__ counter) {
    const int g = blockIdx.x * kBlock + threadIdx.x;
    if (g == 0) *counter = 0;  // the pair kernel has finished: ready for the next run
    if (g >= n_total) return;
    const SegDev& S = segs[seg_of[g]];
    if (!S.eval) return;
    if (S.src_n == 0) {
        es[g] = S.open ? -x[g] : 0.0;
        return;
    }
    if (S.sym >= 0) {
        const SymSeg& E = symsegs[S.sym];
        double s = 0.0, lo = 0.0;
        sym::fold(E.P, E.S, g - S.off, sympart + E.rbase, sympart + E.cbase, s, lo);
        const double e = 0.5 * (s + lo);
        es[g] = S.open ? e - x[g] : e;
        return;
    }
    double s = 0.0, lo = 0.0;
    for (int k = 0; k < S.splits; ++k)
        two_sum_acc(s, lo, part[S.part_base + (int64_t)k * S.n + (g - S.off)]);
    es[g] = S.open ? (s + lo) - x[g] : s + lo;
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

// Per sorted source: FP64 position a = U 2^-64 and a - 1 (both exact; the
// near band reads the second across the periodic seam), FP64 weight, and the
// far-field record. Coincident sources (equal keys: the species of a deck
// share the x lattice of the remesh grid) have the same kernel value, so
// within each 32-source tile (the unit that takes one path) their far-field
// weight is their FP64 sum, held by the first of them (the others get 0):
// the FP32 terms then carry the net charge, and their rounding errors scale
// with it instead of with each species' charge. One thread per position
// (weights and records only for source segments).
__global__ __launch_bounds__(kBlock) void gather_kernel(
    int n_total, const int32_t* __restrict__ seg_of, const SegDev* __restrict__ segs,
    const uint64_t* __restrict__ us, const int32_t* __restrict__ perm,
    const double* __restrict__ w, double* __restrict__ as, double* __restrict__ asm1,
    double* __restrict__ ws, FarSrc* __restrict__ far) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n_total) return;
    const SegDev& S = segs[seg_of[i]];
    const uint64_t u = us[i];
    const double a = __builtin_ldexp((double)u, -64);
    as[i] = a;  // targets read it too
    asm1[i] = a - 1.0;
    if (!S.is_src) return;
    ws[i] = w[perm[i]];
    // tiles start at multiples of kTile from the segment start
    const int t0 = S.off + ((i - S.off) & ~(kTile - 1));
    const int t1 = min(t0 + kTile, S.off + S.n);
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
    // g w = -(w delta^2/2) s/z^2 (1 - 3q/4 + 5q^2/8), z = zs 2^-32, q = delta^2/z^2:
    // with w'' = -w delta^2/2 the FP32 term is sq (w'' + q (w'' b1 + q w'' b2)),
    // sq = s/zs^2 (zs units), q = |sq|, and the tile sum is scaled by 2^64.
    const double w2 = -0.5 * net * S.delta2;
    FarSrc f;
    f.h = (uint32_t)(u >> 32);
    f.w = (float)w2;
    f.wb1 = (float)(w2 * S.b1z);
    f.wb2 = (float)(w2 * S.b2z);
    far[i] = f;
}

// C(2i+k-1, k), i = 1..7, k = 0..12: the coefficients of (Z + e)^-2i in (-e/Z)^k.
__device__ constexpr double kBinom[7][13] = {
    {1.0, 2.0, 3.0, 4.0, 5.0, 6.0, 7.0, 8.0, 9.0, 10.0, 11.0, 12.0, 13.0},
    {1.0, 4.0, 10.0, 20.0, 35.0, 56.0, 84.0, 120.0, 165.0, 220.0, 286.0, 364.0, 455.0},
    {1.0, 6.0, 21.0, 56.0, 126.0, 252.0, 462.0, 792.0, 1287.0, 2002.0, 3003.0, 4368.0, 6188.0},
    {1.0, 8.0, 36.0, 120.0, 330.0, 792.0, 1716.0, 3432.0, 6435.0, 11440.0, 19448.0, 31824.0, 50388.0},
    {1.0, 10.0, 55.0, 220.0, 715.0, 2002.0, 5005.0, 11440.0, 24310.0, 48620.0, 92378.0, 167960.0, 293930.0},
    {1.0, 12.0, 78.0, 364.0, 1365.0, 4368.0, 12376.0, 31824.0, 75582.0, 167960.0, 352716.0, 705432.0, 1352078.0},
    {1.0, 14.0, 105.0, 560.0, 2380.0, 8568.0, 27132.0, 77520.0, 203490.0, 497420.0, 1144066.0, 2496144.0, 5200300.0},
};

// Far-field expansion of each 32-source tile. The regularization remainder
// beyond the band is G(z) = s(z) (a1 z^-2 + a2 z^-4 + a3 z^-6) with
// a1 = -d^2/2, a2 = 3d^4/8, a3 = -5d^6/16 (the binomial series of
// z/sqrt(z^2+d^2) - s(z) in d^2/z^2). With z_j = Z + e_j, e_j = c - x_j for
// the tile centre c and Z = t - c, (Z + e)^-n = Z^-n sum_k C(n+k-1, k)(-e/Z)^k,
// so the tile sums to s(Z) sum_n rho^n sum_{2i+k=n} a_i H^-2i C(2i+k-1, k)
// (-1)^k M_k with rho = H/Z and the moments M_k = sum_j w_j (e_j/H)^k,
// k < kMoments. The centre is on the 2^-32 grid, so Z is the difference of
// high words, as in the pair path. Coefficients are summed in FP64 and
// rounded once to FP32.
// The centre of a range of keys (its midpoint, on the 2^-32 grid).
__device__ __forceinline__ uint32_t range_centre(uint64_t umin, uint64_t umax) {
    const uint64_t mid = umin + ((umax - umin) >> 1);
    return (uint32_t)(mid >> 32);
}

// Adds source (key u, weight w) to the moments M_k = sum w (e/H)^k about the
// centre uc and to the largest distance emax.
__device__ __forceinline__ void add_moments(double* M, uint64_t& emax, uint64_t uc, uint64_t u,
                                            double w, int log2h) {
    const int64_t d = (int64_t)(uc - u);
    const uint64_t ad = (uint64_t)(d < 0 ? -d : d);
    emax = ad > emax ? ad : emax;
    const double eh = __builtin_ldexp((double)d, -64 - log2h);
    double pw = w;
#pragma unroll
    for (int k = 0; k < kMoments64; ++k) {
        M[k] += pw;
        pw *= eh;
    }
}

__device__ void expand_coeffs(const double* M, uint32_t hc, uint64_t emax, const SegDev& S,
                              TileRec& T, TileMom* N, TileRec64* T64);

// The sources of [b, e) in order through src(j) -> (key, weight).
template <class Src>
__device__ __forceinline__ void expand_range(Src src, int b, int e, const SegDev& S, TileRec& T,
                                             TileMom& N, TileRec64* T64) {
    uint64_t umin = src.u(b), umax = src.u(b);
    for (int j = b + 1; j < e; ++j) {
        const uint64_t u = src.u(j);
        umin = u < umin ? u : umin;
        umax = u > umax ? u : umax;
    }
    const uint32_t hc = range_centre(umin, umax);
    const uint64_t uc = (uint64_t)hc << 32;
    uint64_t emax = 0;
    double M[kMoments64];
#pragma unroll
    for (int k = 0; k < kMoments64; ++k) M[k] = 0.0;
    for (int j = b; j < e; ++j) add_moments(M, emax, uc, src.u(j), src.w(j), S.log2h);
    expand_coeffs(M, hc, emax, S, T, &N, T64);
}

// Sources read from the sorted arrays.
struct GlobalSrc {
    const uint64_t* __restrict__ us;
    const double* __restrict__ ws;
    __device__ uint64_t u(int j) const { return us[j]; }
    __device__ double w(int j) const { return ws[j]; }
};

// The expansion kernels give each thread (or each four lanes) whole tiles
// of consecutive sources, which a wavefront's loads would touch 32 sources
// apart; a workgroup of them first copies its tiles' sources [b0, b0 + len)
// (len <= kStageSrc, coalesced) into LDS, one pad slot after every 32 so
// that the lanes' tiles fall in different banks, and reads them from there
// in the same order as from memory.
constexpr int kStageThreads = 64;                    // tiles per workgroup
constexpr int kStageSrc = kStageThreads * kTile;     // 2048 sources
constexpr int kStageSlots = kStageSrc + kStageSrc / kTile;
__device__ __forceinline__ int stage_slot(int j) { return j + (j >> 5); }
struct StagedSrc {
    const uint64_t* su;
    const double* sw;
    int b0;
    __device__ uint64_t u(int j) const { return su[stage_slot(j - b0)]; }
    __device__ double w(int j) const { return sw[stage_slot(j - b0)]; }
};

// Copies the sources [b0, e1) of a workgroup's tiles into LDS when they fit;
// returns whether they did (the same for every thread of the workgroup).
__device__ __forceinline__ bool stage_sources(const uint64_t* __restrict__ us,
                                              const double* __restrict__ ws, int b0, int e1,
                                              uint64_t* su, double* sw) {
    const bool fits = e1 > b0 && e1 - b0 <= kStageSrc;
    if (fits)
        for (int j = threadIdx.x; j < e1 - b0; j += kStageThreads) {
            su[stage_slot(j)] = us[b0 + j];
            sw[stage_slot(j)] = ws[b0 + j];
        }
    __syncthreads();
    return fits;
}

// Coefficients of the FP32 far expansion (and, when asked, the FP64 near
// moments and the probe's FP64 reference) from the moments about the centre.
__device__ void expand_coeffs(const double* M, uint32_t hc, uint64_t emax, const SegDev& S,
                              TileRec& T, TileMom* N, TileRec64* T64) {
    // a_i (d/H)^2i: the binomial series (1+q)^-1/2 - 1 = sum_i C(-1/2, i) q^i
    const double dh2 = __builtin_ldexp(S.delta2, -2 * S.log2h);  // (d/H)^2
    double a[kTerms64];
    {
        double c = 1.0, p = 1.0;
        for (int ii = 1; ii <= kTerms64; ++ii) {
            c = -c * (2.0 * ii - 1.0) / (2.0 * ii);  // C(-1/2, i), exact (dyadic)
            p *= dh2;
            a[ii - 1] = c * p;
        }
    }
    T.hc = hc;
    T.pad = 0;
    T.emax = emax;
    if (T64) {
        T64->m0 = M[0];
        T64->m1 = M[1];
    }
    // coefficient of rho^(m+2): n = m + 2 = 2i + k
    for (int m = 0; m < (T64 ? kCoef64 : kCoef); ++m) {
        double c32 = 0.0, c64 = 0.0;
        for (int ii = 1; ii <= (T64 ? kTerms64 : 3); ++ii) {
            const int k = m + 2 - 2 * ii;
            if (k < 0 || k >= kMoments64) continue;
            const double t = a[ii - 1] * kBinom[ii - 1][k] * ((k & 1) ? -M[k] : M[k]);
            c64 += t;
            if (ii <= 3 && k < kMoments) c32 += t;
        }
        if (m < kCoef) T.c[m] = (float)c32;
        if (T64) T64->c[m] = c64;
    }
    if (!N) return;
    // the near expansion's moments in units of the period, scaled by
    // gamma_k = C(2k, k) / 2^k (exact)
    double gk = 1.0;
    for (int k = 0; k < kNearTerms; ++k) {
        N->a[k] = gk * __builtin_ldexp(M[k], k * S.log2h);
        N->b[k] = gk * __builtin_ldexp(M[k + 1], (k + 1) * S.log2h);
        gk = gk * (2.0 * k + 1.0) / (k + 1.0);
    }
}

// The sources [b, e) of tile t (tile_seg: the segment of each tile).
__device__ __forceinline__ int2 tile_range(int t, const int32_t* __restrict__ tile_seg,
                                           const SegDev* __restrict__ segs) {
    const SegDev& S = segs[tile_seg[t]];
    const int b = S.off + (t - S.tile_base) * kTile;
    return make_int2(b, min(b + kTile, S.off + S.n));
}

// One thread per tile, kStageThreads tiles per workgroup, the tiles' sources
// staged in LDS (from memory where a workgroup's tiles span more than
// kStageSrc sources).
__global__ __launch_bounds__(kStageThreads) void tile_kernel(int ntiles, const int32_t* __restrict__ tile_seg,
                                                      const SegDev* __restrict__ segs,
                                                      const uint64_t* __restrict__ us,
                                                      const double* __restrict__ ws,
                                                      TileRec* __restrict__ tiles,
                                                      TileRec64* __restrict__ tiles64,
                                                      TileMom* __restrict__ moms) {
    __shared__ uint64_t su[kStageSlots];
    __shared__ double sw[kStageSlots];
    const int t0 = blockIdx.x * kStageThreads;
    const int t1 = min(t0 + kStageThreads, ntiles) - 1;
    const int b0 = tile_range(t0, tile_seg, segs).x, e1 = tile_range(t1, tile_seg, segs).y;
    const bool staged = stage_sources(us, ws, b0, e1, su, sw);
    const int t = t0 + (int)threadIdx.x;
    if (t >= ntiles) return;
    const SegDev& S = segs[tile_seg[t]];
    const int2 r = tile_range(t, tile_seg, segs);
    TileRec T;
    TileMom N;
    TileRec64 T64;
    if (staged && r.x >= b0 && r.y <= e1)
        expand_range(StagedSrc{su, sw, b0}, r.x, r.y, S, T, N, &T64);
    else
        expand_range(GlobalSrc{us, ws}, r.x, r.y, S, T, N, &T64);
    tiles[t] = T;
    tiles64[t] = T64;
    moms[t] = N;
}

// The sources [b, e) of group g (grp_seg: the segment of each group).
__device__ __forceinline__ int2 group_range(int g, const int32_t* __restrict__ grp_seg,
                                            const SegDev* __restrict__ segs) {
    const SegDev& S = segs[grp_seg[g]];
    const int b = S.off + (g - S.group_base) * kGroupSrc;
    return make_int2(b, min(b + kGroupSrc, S.off + S.n));
}

// Span and largest tile emax of group g of 4 tiles (after tile_kernel), and
// the group's own expansions: four adjacent lanes, one tile each (sub),
// combined by butterfly shuffles in a fixed order.
template <class Src>
__device__ __forceinline__ void group_one(int g, int sub, const SegDev& S, Src src,
                                          const uint64_t* __restrict__ us,
                                          const TileRec* __restrict__ tiles,
                                          GroupRec* __restrict__ groups,
                                          TileRec* __restrict__ gexp,
                                          TileRec64* __restrict__ gexp64,
                                          TileMom* __restrict__ gmoms) {
    const int i = S.off + (g - S.group_base) * kGroupSrc;
    const int ge = min(i + kGroupSrc, S.off + S.n);
    const int jb = i + sub * kTile, je = min(jb + kTile, ge);
    uint64_t umin = ~uint64_t(0), umax = 0, temax = 0;
    for (int j = jb; j < je; ++j) {
        const uint64_t u = src.u(j);
        umin = u < umin ? u : umin;
        umax = u > umax ? u : umax;
    }
    if (jb < je) temax = tiles[S.tile_base + (jb - S.off) / kTile].emax;
#pragma unroll
    for (int m = 1; m < 4; m <<= 1) {
        const uint64_t a = __shfl_xor(umin, m, 4), b = __shfl_xor(umax, m, 4);
        const uint64_t c = __shfl_xor(temax, m, 4);
        umin = a < umin ? a : umin;
        umax = b > umax ? b : umax;
        temax = c > temax ? c : temax;
    }
    const uint32_t hc = range_centre(umin, umax);
    const uint64_t uc = (uint64_t)hc << 32;
    double M[kMoments64];
#pragma unroll
    for (int k = 0; k < kMoments64; ++k) M[k] = 0.0;
    uint64_t emax = 0;
    for (int j = jb; j < je; ++j) add_moments(M, emax, uc, src.u(j), src.w(j), S.log2h);
#pragma unroll
    for (int m = 1; m < 4; m <<= 1) {
#pragma unroll
        for (int k = 0; k < kMoments64; ++k) M[k] += __shfl_xor(M[k], m, 4);
        const uint64_t e = __shfl_xor(emax, m, 4);
        emax = e > emax ? e : emax;
    }
    if (sub != 0) return;
    GroupRec G;
    G.vf = us[i] & ~kLowSort;
    G.vl = us[ge - 1] | kLowSort;
    G.emax = temax;
    G.pad = 0;
    groups[g] = G;
    // the group's own expansions about its centre (FP32 far, FP64 near)
    TileRec T;
    TileMom N;
    TileRec64 T64;
    expand_coeffs(M, hc, emax, S, T, &N, &T64);
    gexp[g] = T;
    gexp64[g] = T64;
    gmoms[g] = N;
}

// Four adjacent lanes per group, kStageThreads / 4 groups per workgroup, the
// groups' sources staged in LDS as in tile_kernel.
__global__ __launch_bounds__(kStageThreads) void group_kernel(int ngroups, const int32_t* __restrict__ grp_seg,
                                                       const SegDev* __restrict__ segs,
                                                       const uint64_t* __restrict__ us,
                                                       const double* __restrict__ ws,
                                                       const TileRec* __restrict__ tiles,
                                                       GroupRec* __restrict__ groups,
                                                       TileRec* __restrict__ gexp,
                                                       TileRec64* __restrict__ gexp64,
                                                       TileMom* __restrict__ gmoms) {
    __shared__ uint64_t su[kStageSlots];
    __shared__ double sw[kStageSlots];
    constexpr int kGroups = kStageThreads / 4;
    const int g0 = blockIdx.x * kGroups;
    const int g1 = min(g0 + kGroups, ngroups) - 1;
    const int b0 = group_range(g0, grp_seg, segs).x, e1 = group_range(g1, grp_seg, segs).y;
    const bool staged = stage_sources(us, ws, b0, e1, su, sw);
    const int g = g0 + ((int)threadIdx.x >> 2), sub = (int)threadIdx.x & 3;
    if (g >= ngroups) return;  // all four lanes of a group return together
    const SegDev& S = segs[grp_seg[g]];
    const int2 r = group_range(g, grp_seg, segs);
    // the four lanes of a group take the same path
    if (staged && r.x >= b0 && r.y <= e1)
        group_one(g, sub, S, StagedSrc{su, sw, b0}, us, tiles, groups, gexp, gexp64, gmoms);
    else
        group_one(g, sub, S, GlobalSrc{us, ws}, us, tiles, groups, gexp, gexp64, gmoms);
}

// Span and FP32 expansion of each far-field unit of one level: one
// wavefront per unit (unit_seg: the segment of each unit record), the keys'
// extremes, then the 13 moments and the largest distance about the centre,
// summed per lane over a fixed stride and reduced by butterfly shuffles in a
// fixed order. The record's emax is the unit's own.
__global__ __launch_bounds__(kBlock) void unit_kernel(int level, int lstart, int nunits,
                                                      int ngroups,
                                                      const int32_t* __restrict__ unit_seg,
                                                      const SegDev* __restrict__ segs,
                                                      const uint64_t* __restrict__ us,
                                                      const double* __restrict__ ws,
                                                      GroupRec* __restrict__ groups,
                                                      TileRec* __restrict__ gexp,
                                                      TileMom* __restrict__ gmoms) {
    const int tid = blockIdx.x * kBlock + threadIdx.x;
    const int wu = tid >> 6, lane = tid & 63;
    if (wu >= nunits) return;  // whole wavefronts return together
    const int ua = lstart + wu;  // record index (after the groups)
    const int usrc = unit_src(level);
    const SegDev& S = segs[unit_seg[ua - ngroups]];
    const int b = S.off + (ua - S.unit_base[level]) * usrc;
    const int e = min(b + usrc, S.off + S.n);
    uint64_t umin = ~uint64_t(0), umax = 0;
    for (int j = b + lane; j < e; j += 64) {
        const uint64_t u = us[j];
        umin = u < umin ? u : umin;
        umax = u > umax ? u : umax;
    }
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) {
        const uint64_t a0 = __shfl_xor(umin, m), a1 = __shfl_xor(umax, m);
        umin = a0 < umin ? a0 : umin;
        umax = a1 > umax ? a1 : umax;
    }
    const uint32_t hc = range_centre(umin, umax);
    const uint64_t uc = (uint64_t)hc << 32;
    double M[kMoments64];
#pragma unroll
    for (int k = 0; k < kMoments64; ++k) M[k] = 0.0;
    uint64_t emax = 0;
    for (int j = b + lane; j < e; j += 64) add_moments(M, emax, uc, us[j], ws[j], S.log2h);
#pragma unroll
    for (int m = 1; m < 64; m <<= 1) {
#pragma unroll
        for (int k = 0; k < kMoments64; ++k) M[k] += __shfl_xor(M[k], m);
        const uint64_t a1 = __shfl_xor(emax, m);
        emax = a1 > emax ? a1 : emax;
    }
    if (lane != 0) return;
    TileRec T;
    TileMom N;
    expand_coeffs(M, hc, emax, S, T, &N, nullptr);
    GroupRec G;
    G.vf = us[b] & ~kLowSort;
    G.vl = us[e - 1] | kLowSort;
    G.emax = T.emax;
    G.pad = 0;
    groups[ua] = G;
    gexp[ua] = T;
    gmoms[ua] = N;
}

// The FP32 value of one expanded tile at a target (high word ht): the pair
// kernel and the probe evaluate it with the same operations.
__device__ __forceinline__ float tile_value(const TileRec& T, uint32_t ht, int hexp) {
    const float zs = (float)(int32_t)(ht - T.hc);
    const float rho = __builtin_amdgcn_rcpf(__builtin_ldexpf(zs, hexp));
    const float sq = rho * __builtin_fabsf(rho);
    float p = T.c[kCoef - 1];
#pragma unroll
    for (int m = kCoef - 2; m >= 0; --m) p = __builtin_fmaf(p, rho, T.c[m]);
    return sq * p;
}

// A zero the compiler cannot see through, held in a VGPR: added to a uniform
// index it makes a load a vector load, so the moments of a near expansion
// sit in VGPRs (an FP64 FMA may take two of them; in SGPRs they would need a
// move each and crowd out the kernel's other scalar state).
__device__ __forceinline__ int vgpr_zero() {
    int z;
    asm volatile("v_mov_b32 %0, 0" : "=v"(z));
    return z;
}

// The near expansion of a range within emax of its centre, whose
// displacements to every target lie on one side, at least gap away (both in
// units of 2^-64 of the period): every target is at r = sqrt(Z^2 + d^2) >=
// r0 = sqrt(gap^2 + d^2) from the centre, and K Legendre terms leave a
// truncation below (emax/r)^K of sum |w|. The number of terms that keeps it
// at most 2^-48 (16^-12): 12 when emax <= r0/16, 10 within r0/32, 8 within
// r0/64, 6 within r0/256; 0 = not compact enough for the expansion.
__device__ __forceinline__ int near_terms(uint64_t emax, uint64_t gap, double d2) {
    const double e = __builtin_ldexp((double)emax, -64);
    const double g = __builtin_ldexp((double)gap, -64);
    const double r2 = __builtin_fma(g, g, d2), e2 = e * e;
    return e2 * 65536.0 <= r2 ? 6 : e2 * 4096.0 <= r2 ? 8 : e2 * 1024.0 <= r2 ? 10
                                                          : e2 * 256.0 <= r2 ? 12 : 0;
}

// sum_j w_j g(z_j) of a compact range (scaled moments N about the centre uc)
// on the side s of a target with key ut, in FP64: sum_j w_j z_j/sqrt(z_j^2 +
// d^2) from the Legendre generating function, minus s sum_j w_j. With
// r = sqrt(Z^2 + d^2), x = -Z/r and V_k = P_k(x) r^-(k+1) / gamma_k, the sum is
// sum_k V_k (Z a_k + b_k), and V obeys V_k+1 = A V_k - c_k B V_k-1 with
// A = -Z/r^2, B = 1/r^2, c_k = k^2/(4k^2 - 1): three operations per term
// for the recurrence, two for the sum. K terms (6, 8, 10 or 12, uniform over
// the wavefront; the moments are indexed with constants only, so they stay in
// registers).
template <int k>
__device__ __forceinline__ void near_term(const TileMom& N, double Z, double A, double B, double& vm,
                                          double& v, double& acc) {
    constexpr double ck = (double)(k * k) / (double)(4 * k * k - 1);
    const double vn = __builtin_fma(-ck, B * vm, A * v);
    vm = v;
    v = vn;
    acc = __builtin_fma(v, __builtin_fma(Z, N.a[k + 1], N.b[k + 1]), acc);
}

__device__ __forceinline__ double near_value(const TileMom& N, uint64_t uc, uint64_t ut, double d2,
                                             double s, int K) {
    const double Z = __builtin_ldexp((double)(int64_t)(ut - uc), -64);
    const double ri = rsq_f64(__builtin_fma(Z, Z, d2));
    const double B = ri * ri;
    const double A = -Z * B;
    double vm = ri, v = A * ri;
    double acc = vm * __builtin_fma(Z, N.a[0], N.b[0]);
    acc = __builtin_fma(v, __builtin_fma(Z, N.a[1], N.b[1]), acc);
    near_term<1>(N, Z, A, B, vm, v, acc);
    near_term<2>(N, Z, A, B, vm, v, acc);
    near_term<3>(N, Z, A, B, vm, v, acc);
    near_term<4>(N, Z, A, B, vm, v, acc);
    if (K > 6) {
        near_term<5>(N, Z, A, B, vm, v, acc);
        near_term<6>(N, Z, A, B, vm, v, acc);
    }
    if (K > 8) {
        near_term<7>(N, Z, A, B, vm, v, acc);
        near_term<8>(N, Z, A, B, vm, v, acc);
    }
    if (K > 10) {
        near_term<9>(N, Z, A, B, vm, v, acc);
        near_term<10>(N, Z, A, B, vm, v, acc);
    }
    return acc - s * N.a[0];
}

// Per scan block of a source segment: sum of w and of w a.
__global__ __launch_bounds__(kBlock) void blocksum_kernel(const ScanBlock* __restrict__ blocks,
                                                          const double* __restrict__ ws,
                                                          const double* __restrict__ as,
                                                          double* __restrict__ bsum,
                                                          double* __restrict__ bsa) {
    __shared__ double sw[kBlock], swa[kBlock];
    const ScanBlock B = blocks[blockIdx.x];
    constexpr int kPer = kScanItems / kBlock;
    double s = 0.0, sa = 0.0;
    for (int k = 0; k < kPer; ++k) {
        const int i = B.b + k * kBlock + (int)threadIdx.x;
        if (i < B.e) {
            const double wi = ws[i];
            s += wi;
            sa += wi * as[i];
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

// Pair kernel. Per (a wavefront's 256 consecutive sorted targets, range of
// sources: far-field unit, group or 32-source tile), one path, the choice
// uniform over the wavefront; for a tile one of:
//  * near tile (some displacement within B of 0): FP64 from the FP64
//    positions, z = a_i - a_j with the periodic seam handled per tile (the
//    targets' a - 1 and the sources' a - 1 are exact where they are used);
//  * antipodal tile (some displacement within 2^-24 of 1/2) or a tile whose
//    displacements span half a period: FP64 from the exact 64-bit difference;
//  * far tile: the FP32 series from the high words.
template <int TPT>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) void batch_mixed_kernel(
    const WorkItem* __restrict__ items, int nitems, const SegDev* __restrict__ segs,
    const uint64_t* __restrict__ us, const double* __restrict__ as,
    const double* __restrict__ asm1, const double* __restrict__ ws,
    const FarSrc* __restrict__ far, const TileRec* __restrict__ tiles,
    const GroupRec* __restrict__ groups, const TileMom* __restrict__ moms,
    const TileRec* __restrict__ gexp, const TileMom* __restrict__ gmoms,
    double* __restrict__ part, int* __restrict__ counter) {
    for (;;) {
        const int it = pull_item(counter);
        if (it >= nitems) break;
        const WorkItem W = items[it];
        const SegDev S = segs[W.seg];
        const double d2 = S.delta2;
        const int hexp = -32 - S.log2h;
        // each wavefront takes 64 * TPT consecutive sorted targets of the item
        // and makes its own near/far choices (uniform over the wavefront: no
        // divergence, and scalar as long as the wavefront index is known to
        // be uniform): the narrower the target range, the fewer sources need
        // pairs or the smaller ranges
        const int lane = (int)threadIdx.x & 63;
        const int wtb = W.tb + __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6) * 64 * TPT;
        const int wte = min(W.te, wtb + 64 * TPT);
        if (wtb >= wte) continue;
        // positions are sorted by their top bits only: widen the range by one
        // unit of the sort on each side
        const uint64_t uf = us[wtb] & ~kLowSort, ul = us[wte - 1] | kLowSort;
        const uint64_t tspan = ul - uf;
        uint64_t ui[TPT];
        uint32_t hi[TPT];
        double ai[TPT], sh[TPT], sl[TPT], gx[TPT];
#pragma unroll
        for (int t = 0; t < TPT; ++t) {
            const int i = min(wtb + lane + t * 64, wte - 1);
            ui[t] = us[i];
            hi[t] = (uint32_t)(ui[t] >> 32);
            ai[t] = as[i];
            sh[t] = 0.0;
            sl[t] = 0.0;
            gx[t] = 0.0;
        }
        int nk = 0;  // terms of a near expansion
        for (int gb = W.sb; gb < W.se;) {
          // the largest unit (2048 or 512 sources) starting here that is far
          // from all the wavefront's targets and compact: one FP32 expansion
          // for all its sources; else the largest on one side of them and
          // compact for its distance: one FP64 near expansion
          if (tspan < kHalf) {
            int step = 0;
#pragma unroll
            for (int L = kLevels - 1; L >= 0; --L) {
              const int rel = gb - S.src_off;
              if (L >= S.levels || step != 0 || (rel & (unit_src(L) - 1)) != 0) continue;
              const int uix = S.src_unit_base[L] + rel / unit_src(L);
              const GroupRec U = groups[uix];
              const uint64_t uspan = U.vl - U.vf;
              const uint64_t ulen = tspan + uspan, ulo = uf - U.vl;
              if (uspan >= kHalf || ulen >= kHalf ||
                  cyc_meets(ulo - kHalf + kAntiMargin, ulen, 2 * kAntiMargin))
                continue;
              if (!S.all_band && U.emax <= S.xband && !cyc_meets(ulo + S.band, ulen, 2 * S.band)) {
                const TileRec T = gexp[uix];
#pragma unroll
                for (int t = 0; t < TPT; ++t) gx[t] += (double)tile_value(T, hi[t], hexp);
                step = unit_src(L);
              } else if (!cyc_meets(ulo, ulen, 1) &&
                         (nk = near_terms(U.emax, ulo < kHalf ? ulo : 0 - (ulo + ulen), d2)) > 0) {
                const uint64_t uc = (uint64_t)gexp[uix].hc << 32;
                const TileMom N = gmoms[uix + vgpr_zero()];
                const double s = (ulo < kHalf) ? 1.0 : -1.0;
#pragma unroll
                for (int t = 0; t < TPT; ++t)
                    two_sum_acc(sh[t], sl[t], near_value(N, uc, ui[t], d2, s, nk));
                step = unit_src(L);
              }
            }
            if (step != 0) {
              gb = min(W.se, gb + step);
              continue;
            }
          }
          const int gcur = gb;
          const int ge = min(W.se, gb + kGroupSrc);
          gb = ge;
          if (tspan < kHalf) {
            const int gix = S.src_group_base + (gcur - S.src_off) / kGroupSrc;
            const GroupRec G = groups[gix];
            const uint64_t gspan = G.vl - G.vf;
            const uint64_t glen = tspan + gspan, glo = uf - G.vl;
            if (gspan < kHalf && glen < kHalf &&
                !cyc_meets(glo - kHalf + kAntiMargin, glen, 2 * kAntiMargin)) {
                const uint64_t gemax = gexp[gix].emax;
                if (!S.all_band && !cyc_meets(glo + S.band, glen, 2 * S.band)) {
                    // a group far from all the wavefront's targets: its own FP32
                    // expansion when compact, else its tiles' when they are
                    if (gemax <= S.xband) {
                        const TileRec T = gexp[gix];
#pragma unroll
                        for (int t = 0; t < TPT; ++t) gx[t] += (double)tile_value(T, hi[t], hexp);
                        continue;
                    }
                    if (G.emax <= S.xband) {
                        for (int jb = gcur; jb < ge; jb += kTile) {
                            const TileRec T = tiles[S.src_tile_base + (jb - S.src_off) / kTile];
#pragma unroll
                            for (int t = 0; t < TPT; ++t) gx[t] += (double)tile_value(T, hi[t], hexp);
                        }
                        continue;
                    }
                }
                if (!cyc_meets(glo, glen, 1) &&
                    (nk = near_terms(gemax, glo < kHalf ? glo : 0 - (glo + glen), d2)) > 0) {
                    // a group on one side of all the wavefront's targets and
                    // compact for its distance (within reach of the band, or far
                    // but too spread for the FP32 expansion): its FP64 expansion
                    const uint64_t uc = (uint64_t)gexp[gix].hc << 32;
                    const TileMom N = gmoms[gix + vgpr_zero()];
                    const double s = (glo < kHalf) ? 1.0 : -1.0;
#pragma unroll
                    for (int t = 0; t < TPT; ++t)
                        two_sum_acc(sh[t], sl[t], near_value(N, uc, ui[t], d2, s, nk));
                    continue;
                }
            }
          }
          for (int jb = gcur; jb < ge; jb += kTile) {
            const int je = min(W.se, jb + kTile);
            const uint64_t vf = us[jb] & ~kLowSort, vl = us[je - 1] | kLowSort;
            const uint64_t vspan = vl - vf;
            const bool wide = tspan >= kHalf || vspan >= kHalf || tspan + vspan >= kHalf;
            const uint64_t len = tspan + vspan;
            const uint64_t lo = uf - vl;
            const bool near = S.all_band || cyc_meets(lo + S.band, len, 2 * S.band);
            const bool anti = cyc_meets(lo - kHalf + kAntiMargin, len, 2 * kAntiMargin);
            if (wide || anti) {
                double acc[TPT];
#pragma unroll
                for (int t = 0; t < TPT; ++t) acc[t] = 0.0;
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
            } else if (near && !cyc_meets(lo, len, 1) &&
                       (nk = near_terms(tiles[S.src_tile_base + (jb - S.src_off) / kTile].emax,
                                        lo < kHalf ? lo : 0 - (lo + len), d2)) > 0) {
                // near tile on one side of all the wavefront's targets, compact for
                // its distance: its FP64 expansion about the centre (Legendre
                // generating function; truncation below 16^-12 of sum |w|), and
                // s = the side
                const int tix = S.src_tile_base + (jb - S.src_off) / kTile;
                const uint64_t uc = (uint64_t)tiles[tix].hc << 32;
                const TileMom N = moms[tix + vgpr_zero()];
                const double s = (lo < kHalf) ? 1.0 : -1.0;
#pragma unroll
                for (int t = 0; t < TPT; ++t) two_sum_acc(sh[t], sl[t], near_value(N, uc, ui[t], d2, s, nk));
            } else if (near) {
                // no displacement of this tile reaches 1/2, so each pair's
                // wrap is the tile's: targets above the sources by more than
                // half a period are seen one period down, and the reverse
                const bool t_down = uf > vl && uf - vl > kHalf;
                const bool s_down = vf > ul && vf - ul > kHalf;
                const double* __restrict__ src = s_down ? asm1 : as;
                double acc[TPT], at[TPT];
#pragma unroll
                for (int t = 0; t < TPT; ++t) {
                    acc[t] = 0.0;
                    at[t] = t_down ? ai[t] - 1.0 : ai[t];
                }
#pragma unroll 2
                for (int j = jb; j < je; ++j) {
                    const double aj = src[j];
                    const double wj = ws[j];
#pragma unroll
                    for (int t = 0; t < TPT; ++t) {
                        const double z = at[t] - aj;
                        const double y = rsq_f64(__builtin_fma(z, z, d2));
                        const double s = __builtin_copysign(1.0, z);
                        acc[t] = __builtin_fma(wj, __builtin_fma(z, y, -s), acc[t]);
                    }
                }
#pragma unroll
                for (int t = 0; t < TPT; ++t) two_sum_acc(sh[t], sl[t], acc[t]);
            } else if (tiles[S.src_tile_base + (jb - S.src_off) / kTile].emax <= S.xband) {
                // compact far tile: its expansion about the centre
                const TileRec T = tiles[S.src_tile_base + (jb - S.src_off) / kTile];
#pragma unroll
                for (int t = 0; t < TPT; ++t) gx[t] += (double)tile_value(T, hi[t], hexp);
            } else if ((nk = near_terms(tiles[S.src_tile_base + (jb - S.src_off) / kTile].emax,
                                        lo < kHalf ? lo : 0 - (lo + len), d2)) > 0) {
                // far tile too spread for the FP32 expansion but compact for
                // its distance: its FP64 expansion (a far tile is on one side)
                const int tix = S.src_tile_base + (jb - S.src_off) / kTile;
                const uint64_t uc = (uint64_t)tiles[tix].hc << 32;
                const TileMom N = moms[tix + vgpr_zero()];
                const double s = (lo < kHalf) ? 1.0 : -1.0;
#pragma unroll
                for (int t = 0; t < TPT; ++t) two_sum_acc(sh[t], sl[t], near_value(N, uc, ui[t], d2, s, nk));
            } else {
                float acc[TPT];
#pragma unroll
                for (int t = 0; t < TPT; ++t) acc[t] = 0.0f;
                if (je - jb == kTile) {
                    // whole tile: records in groups of 4 (one scalar load each;
                    // a second buffer in flight would crowd the SGPRs)
                    FarSrc cur[4];
#pragma unroll
                    for (int g = 0; g < kTile / 4; ++g) {
#pragma unroll
                        for (int k = 0; k < 4; ++k) cur[k] = far[jb + 4 * g + k];
#pragma unroll
                        for (int k = 0; k < 4; ++k) {
                            const FarSrc f = cur[k];
#pragma unroll
                            for (int t = 0; t < TPT; ++t) {
                                const float zs = (float)(int32_t)(hi[t] - f.h);
                                const float r = __builtin_amdgcn_rcpf(zs);
                                const float sq = r * __builtin_fabsf(r);
                                const float q = __builtin_fabsf(sq);
                                const float p = __builtin_fmaf(q, __builtin_fmaf(q, f.wb2, f.wb1), f.w);
                                acc[t] = __builtin_fmaf(sq, p, acc[t]);
                            }
                        }
                    }
                } else {
                    for (int j = jb; j < je; ++j) {
                        const FarSrc f = far[j];
#pragma unroll
                        for (int t = 0; t < TPT; ++t) {
                            const float zs = (float)(int32_t)(hi[t] - f.h);
                            const float r = __builtin_amdgcn_rcpf(zs);
                            const float sq = r * __builtin_fabsf(r);  // s(z) / zs^2
                            const float q = __builtin_fabsf(sq);
                            const float p = __builtin_fmaf(q, __builtin_fmaf(q, f.wb2, f.wb1), f.w);
                            acc[t] = __builtin_fmaf(sq, p, acc[t]);
                        }
                    }
                }
#pragma unroll
                for (int t = 0; t < TPT; ++t)
                    two_sum_acc(sh[t], sl[t], __builtin_ldexp((double)acc[t], 64));
            }
          }
        }
#pragma unroll
        for (int t = 0; t < TPT; ++t) {
            two_sum_acc(sh[t], sl[t], gx[t]);
            const int i = wtb + lane + t * 64;
            if (i < wte) part[S.part_base + (int64_t)W.split * S.n + (i - S.off)] = sh[t] + sl[t];
        }
    }
}

// ---- accuracy governor ----------------------------------------------------
// Before the pair kernel of a governed run, the FP32 far-field error of each
// evaluated segment is measured on the run's own inputs: a probe takes up to
// 128 evenly spaced targets of the segment and sums, for each target,
//  * the exact FP64 field (every source, the reference kernel on the exact
//    64-bit displacement), whose largest magnitude over the probed targets
//    stands for max|E|;
//  * for each of the ladder bands B_q = B0 2^((q-2)/2) (q = 0..7, at least
//    2^-6), over the source tiles that are far from the target at B_q, the
//    difference between the tile's FP32 value, computed exactly as the pair
//    kernel computes it (tile expansion or FP32 pair sum), and the exact FP64
//    regularization remainder of the same pairs.
// The governor then gives each segment the narrowest band whose largest
// error (over the probed targets) is within target * max|E|, or every pair in
// FP64 when none is. A tile that is far for the pair kernel's whole target
// block is far for each of its targets, so the probe counts every FP32 tile
// of the kernel (and a few more). The band is a function of the run's inputs
// alone: evaluating the same points again gives the same field, bit for bit.

__device__ __forceinline__ uint64_t ladder_band(const SegDev& S, int q) {
    double f = __builtin_ldexp(1.0, (q >> 1) - 1);  // 2^((q - 2)/2)
    if (q & 1) f *= 1.4142135623730951;
    double b = __builtin_ldexp((double)S.band0, -64) * f;
    b = b > 0.015625 ? b : 0.015625;
    return b >= 0.25 ? 0 : (uint64_t)__builtin_ldexp(b, 64);  // 0: all pairs FP64
}

// The FP64 reference remainder of an expanded range at rho = H/Z (coefficients
// read through the pointer, a few at a time, so they do not all sit in
// registers at once).
__device__ __forceinline__ double ref_value(const TileRec64* __restrict__ T, double rho) {
    double p = T->c[kCoef64 - 1];
#pragma unroll 4
    for (int c = kCoef64 - 2; c >= 0; --c) p = __builtin_fma(p, rho, T->c[c]);
    return rho * __builtin_fabs(rho) * p;
}

__device__ __forceinline__ int probe_samples(int n) {
    const int s = n / 64;
    return min(n, min(kProbeMax, max(16, s)));
}

// grid (kProbeMax, nseg): one workgroup per probed target. gov (zeroed
// before) receives per segment max|E| and the error at each ladder band.
__global__ __launch_bounds__(kBlock) void probe_kernel(const SegDev* __restrict__ segs,
                                                       const uint64_t* __restrict__ us,
                                                       const double* __restrict__ ws,
                                                       const FarSrc* __restrict__ far,
                                                       const TileRec* __restrict__ tiles,
                                                       const TileRec64* __restrict__ tiles64,
                                                       const GroupRec* __restrict__ groups,
                                                       const TileRec* __restrict__ gexp,
                                                       const TileRec64* __restrict__ gexp64,
                                                       uint64_t* __restrict__ gov) {
    __shared__ double red[kGovVals][kBlock];
    const int d = blockIdx.y;
    const SegDev S = segs[d];
    if (!S.eval || S.all_band0 || S.n == 0 || S.src_n == 0) return;
    const int ns = probe_samples(S.n);
    if ((int)blockIdx.x >= ns) return;
    const int k = S.off + (int)(((int64_t)(2 * blockIdx.x + 1) * S.n) / (2 * ns));
    const uint64_t ut = us[k];
    const uint32_t ht = (uint32_t)(ut >> 32);
    const uint64_t uf = ut & ~kLowSort, ul = ut | kLowSort;
    const uint64_t tspan = ul - uf;
    uint64_t band[kLadder];
#pragma unroll
    for (int q = 0; q < kLadder; ++q) band[q] = ladder_band(S, q);
    double err[kLadder];
#pragma unroll
    for (int q = 0; q < kLadder; ++q) err[q] = 0.0;
    double efield = 0.0;
    const SegDev R = segs[S.src];
    const double d2 = S.delta2;
    const int ntile = (R.n + kTile - 1) / kTile;
    const int ngroup = (R.n + kGroupSrc - 1) / kGroupSrc;
    const double H = __builtin_ldexp(1.0, S.log2h);
    const int hexp = -32 - S.log2h;
    for (int g = threadIdx.x; g < ngroup; g += kBlock) {
        // the group as one: far at the bands q < kfg, its own expansion
        // compact at q >= qg (the pair kernel takes it when the group is far
        // from all the targets of a wavefront)
        const GroupRec G = groups[R.group_base + g];
        const TileRec GE = gexp[R.group_base + g];
        const uint64_t gspan = G.vl - G.vf;
        const uint64_t glen = tspan + gspan, glo = uf - G.vl;
        int kfg = 0;
        if (gspan < kHalf && glen < kHalf && !cyc_meets(glo - kHalf + kAntiMargin, glen, 2 * kAntiMargin))
            while (kfg < kLadder && band[kfg] != 0 && !cyc_meets(glo + band[kfg], glen, 2 * band[kfg]))
                ++kfg;
        int qg = 0;
        while (qg < kfg && GE.emax > (band[qg] >> kXShift)) ++qg;
        // the far-field units that contain the group: far at the bands
        // q < kfu, compact at q >= qu; the pair kernel takes the largest
        // such unit, and the unit's value is counted once (with its first
        // group), against the exact remainders of all its groups
        int kfu[kLevels], qu[kLevels];
        double vu[kLevels];
#pragma unroll
        for (int L = 0; L < kLevels; ++L) {
            const int gpu = unit_src(L) / kGroupSrc;
            const int uix = R.unit_base[L] + g / gpu;
            const GroupRec U = groups[uix];
            const uint64_t uspan = U.vl - U.vf;
            const uint64_t ulen = tspan + uspan, ulo = uf - U.vl;
            kfu[L] = 0;
            if (L < S.levels && tspan < kHalf && uspan < kHalf && ulen < kHalf &&
                !cyc_meets(ulo - kHalf + kAntiMargin, ulen, 2 * kAntiMargin))
                while (kfu[L] < kLadder && band[kfu[L]] != 0 &&
                       !cyc_meets(ulo + band[kfu[L]], ulen, 2 * band[kfu[L]]))
                    ++kfu[L];
            qu[L] = 0;
            while (qu[L] < kfu[L] && U.emax > (band[qu[L]] >> kXShift)) ++qu[L];
            vu[L] = (g % gpu == 0 && qu[L] < kfu[L]) ? (double)tile_value(gexp[uix], ht, hexp)
                                                      : 0.0;
        }
        double et[kLadder];
#pragma unroll
        for (int q = 0; q < kLadder; ++q) et[q] = 0.0;
        double exg = 0.0;
        // a group far from the target at band 0 and compact there: at every
        // band where it is far the pair kernel takes its own expansion (or a
        // unit), so its exact field and remainder come from its FP64
        // reference expansion, and its tiles count only at the bands where
        // they are far and the group is not
        const bool gfast = qg == 0 && kfg > 0;
        if (gfast) {
            const TileRec64* T64 = gexp64 + R.group_base + g;
            const double Z = __builtin_ldexp((double)(int64_t)(ut - ((uint64_t)GE.hc << 32)), -64);
            exg = ref_value(T64, H / Z);
            efield += 0.5 * (__builtin_copysign(1.0, Z) * T64->m0 -
                             S.C * (Z * T64->m0 + H * T64->m1) + exg);
        }
        for (int m = g * (kGroupSrc / kTile); m < min(ntile, (g + 1) * (kGroupSrc / kTile)); ++m) {
            const int jb = R.off + m * kTile, je = min(jb + kTile, R.off + R.n);
            const uint64_t vf = us[jb] & ~kLowSort, vl = us[je - 1] | kLowSort;
            const uint64_t vspan = vl - vf;
            const bool wide = tspan >= kHalf || vspan >= kHalf || tspan + vspan >= kHalf;
            const uint64_t len = tspan + vspan;
            const uint64_t lo = uf - vl;
            const bool anti = !wide && cyc_meets(lo - kHalf + kAntiMargin, len, 2 * kAntiMargin);
            int kf = 0;
            if (!wide && !anti)
                while (kf < kLadder && band[kf] != 0 && !cyc_meets(lo + band[kf], len, 2 * band[kf]))
                    ++kf;
            if (gfast && kf <= kfg) continue;
            const TileRec T = tiles[R.tile_base + m];
            // the tile expands at the bands q >= qx and takes the pair path below
            int qx = 0;
            while (qx < kf && T.emax > (band[qx] >> kXShift)) ++qx;
            // the exact field of the tile's sources, K(z) = (z/sqrt(z^2+d^2) - C z)/2
            // = ((s - C z) + g)/2, and its regularization remainder g: for a
            // compact far tile from the FP64 reference expansion (7 series terms,
            // 13 moments: truncation below 1e-12 of the remainder at |z| >= 8d
            // and e/|Z| <= 1/15, so the FP32 expansion's own truncations show in
            // the measured error) and moments, otherwise pair by pair (for
            // |z| >= 2^-6 the cancellation in z/sqrt(z^2+d^2) - s costs
            // ~1e-16 d^-2 z^2 relative, far below the FP32 errors being measured)
            double ek, ex;
            if (kf > 0 && qx == 0) {
                const TileRec64* T64 = tiles64 + R.tile_base + m;
                const double Z = __builtin_ldexp((double)(int64_t)(ut - ((uint64_t)T.hc << 32)), -64);
                ex = ref_value(T64, H / Z);
                ek = __builtin_copysign(1.0, Z) * T64->m0 - S.C * (Z * T64->m0 + H * T64->m1) + ex;
            } else {
                ek = 0.0;
                ex = 0.0;
                for (int j = jb; j < je; ++j) {
                    const double z = __builtin_ldexp((double)(int64_t)(ut - us[j]), -64);
                    const double tz = z * rsq_f64(__builtin_fma(z, z, d2));
                    ek += ws[j] * (tz - S.C * z);
                    ex += ws[j] * (tz - __builtin_copysign(1.0, z));
                }
            }
            if (!gfast) {
                efield += 0.5 * ek;
                exg += ex;
            }
            if (kf == 0) continue;
            float acc = 0.0f;
            if (qx > 0)
                for (int j = jb; j < je; ++j) {
                    const FarSrc f = far[j];
                    const float zs = (float)(int32_t)(ht - f.h);
                    const float r = __builtin_amdgcn_rcpf(zs);
                    const float sq = r * __builtin_fabsf(r);
                    const float qq = __builtin_fabsf(sq);
                    const float p = __builtin_fmaf(qq, __builtin_fmaf(qq, f.wb2, f.wb1), f.w);
                    acc = __builtin_fmaf(sq, p, acc);
                }
            const double ep = 0.5 * (__builtin_ldexp((double)acc, 64) - ex);
            const double ee = (qx < kf) ? 0.5 * ((double)tile_value(T, ht, hexp) - ex) : 0.0;
            for (int q = gfast ? kfg : 0; q < kf; ++q) et[q] += (q < qx) ? ep : ee;
        }
        const double eg = (qg < kfg) ? 0.5 * ((double)tile_value(GE, ht, hexp) - exg) : 0.0;
        for (int q = 0; q < kLadder; ++q) {
            double v = (q >= qg && q < kfg) ? eg : et[q];
#pragma unroll
            for (int L = 0; L < kLevels; ++L)
                if (q >= qu[L] && q < kfu[L]) v = 0.5 * (vu[L] - exg);
            err[q] += v;
        }
    }
#pragma unroll
    for (int q = 0; q < kLadder; ++q) red[1 + q][threadIdx.x] = err[q];
    red[0][threadIdx.x] = efield;
    __syncthreads();
    for (int h = kBlock / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h)
#pragma unroll
            for (int q = 0; q < kGovVals; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + h];
        __syncthreads();
    }
    if ((int)threadIdx.x < kGovVals) {
        const double e = __builtin_fabs(red[threadIdx.x][0]);
        atomicMax((unsigned long long*)&gov[(int64_t)d * kGovVals + threadIdx.x],
                  (unsigned long long)__double_as_longlong(e));
    }
}

// Band of every evaluated segment from the probe's measurements (above).
__global__ void governor_kernel(int nseg, SegDev* __restrict__ segs,
                                const uint64_t* __restrict__ gov, double target) {
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= nseg) return;
    SegDev& S = segs[d];
    if (!S.eval || S.all_band0) return;
    const uint64_t* g = gov + (int64_t)d * kGovVals;
    const double em = __longlong_as_double((long long)g[0]);
    uint64_t band = 0;
    for (int q = 0; q < kLadder; ++q) {
        const uint64_t b = ladder_band(S, q);
        if (b == 0) break;
        if (__longlong_as_double((long long)g[1 + q]) <= target * em) {
            band = b;
            break;
        }
    }
    S.all_band = band == 0 ? 1 : 0;
    S.band = band == 0 ? (uint64_t(1) << 62) : band;
    S.xband = S.band >> kXShift;
}

// Sum of the sorted weights of [off, off+n) whose key is <= key. The keys
// are sorted by their top bits; within a run of equal top bits the full keys
// are compared one by one.
__device__ double w_le(const uint64_t* __restrict__ us, const double* __restrict__ ws,
                       const double* __restrict__ pw, int off, int n, int pwb, uint64_t key) {
    const uint64_t kt = key >> kSortShift;
    int lo = 0, hi = n;  // first index with top >= kt
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if ((us[off + mid] >> kSortShift) < kt)
            lo = mid + 1;
        else
            hi = mid;
    }
    double s = pw[pwb + lo];
    for (int j = lo; j < n && (us[off + j] >> kSortShift) == kt; ++j)
        if (us[off + j] <= key) s += ws[off + j];
    return s;
}

// E at sorted position k (segment d), written to es at the original index;
// returns it (0 for points of segments that are not evaluated).
__device__ __forceinline__ double field_at(int k, int d, const SegDev* __restrict__ segs,
                                           const uint64_t* __restrict__ us,
                                           const int32_t* __restrict__ perm,
                                           const double* __restrict__ ws,
                                           const double* __restrict__ pw,
                                           const double* __restrict__ totals,
                                           const double* __restrict__ part,
                                           double* __restrict__ es) {
    const SegDev& S = segs[d];
    if (!S.eval) return 0.0;
    if (S.src_n == 0) {
        es[perm[k]] = 0.0;
        return 0.0;
    }
    double g = 0.0, gl = 0.0;
    for (int s = 0; s < S.splits; ++s)
        two_sum_acc(g, gl, part[S.part_base + (int64_t)s * S.n + (k - S.off)]);
    const SegDev& R = segs[S.src];
    const uint64_t u = us[k];
    const double W = pw[R.pw_base + R.n], A = totals[S.src];
    double P0;
    if (S.src == d) {
        // u is this segment's own sorted key at k: the run of its top bits
        // starts at or just before k (the search w_le makes, found locally)
        const uint64_t kt = u >> kSortShift;
        int lo = k - S.off;
        while (lo > 0 && (us[S.off + lo - 1] >> kSortShift) == kt) --lo;
        P0 = pw[R.pw_base + lo];
        for (int j = lo; j < R.n && (us[R.off + j] >> kSortShift) == kt; ++j)
            if (us[R.off + j] <= u) P0 += ws[R.off + j];
    } else {
        P0 = w_le(us, ws, pw, R.off, R.n, R.pw_base, u);
    }
    const double Pm = (u >= kHalf) ? w_le(us, ws, pw, R.off, R.n, R.pw_base, u - kHalf) : 0.0;
    const double Pp = (u < kHalf) ? w_le(us, ws, pw, R.off, R.n, R.pw_base, u + kHalf) : W;
    const double sum_s = 2.0 * ((P0 - Pm) - Pp) + W;
    const double ai = __builtin_ldexp((double)u, -64);
    const double sum_z = ((ai * W - A) - Pm) + (W - Pp);
    const double e = 0.5 * ((sum_s - S.C * sum_z) + (g + gl));
    es[perm[k]] = e;
    return e;
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
    (void)field_at(k, seg_of[k], segs, us, perm, ws, pw, totals, part, es);
}

__global__ void zero_int_kernel(int* p) { *p = 0; }

// The segment of every point for segments in ascending order: the last one
// starting at or before it, 0 where none holds it (as the host fill leaves).
__global__ __launch_bounds__(kBlock) void seg_of_kernel(const SegDev* __restrict__ segs, int nseg,
                                                        int n, int32_t* __restrict__ seg_of) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    int lo = 0, hi = nseg - 1;
    while (lo < hi) {
        const int m = (lo + hi + 1) >> 1;
        if (segs[m].off <= i) lo = m;
        else hi = m - 1;
    }
    seg_of[i] = (segs[lo].off <= i && i < segs[lo].off + segs[lo].n) ? lo : 0;
}

__global__ void zero_u64_kernel(uint64_t* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

inline unsigned nblocks(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// Source chunking of one deck: enough (target block, source chunk) items for
// about 16 wavefronts per SIMD when the deck runs alone (a persistent kernel
// with up to 8 resident workgroups per CU then always has work). Depends only
// on the deck's sizes and the device, never on the rest of the batch.
void chunking(int precision, int64_t nt, int64_t ns, int num_cus, const FieldTuning& tun,
              int& splits, int& chunk) {
    const int64_t want_waves = int64_t(64) * num_cus;
    const int unit = precision == kPrecisionF64 ? kSub : chunk_unit_src(ns);
    const int tpt_auto = (nt >= want_waves * 64 * 2) ? 2 : 1;
    const int64_t waves = (nt + 64 * tpt_auto - 1) / (64 * tpt_auto);
    int64_t sp = tun.source_splits > 0 ? tun.source_splits
                                       : (want_waves + waves - 1) / std::max<int64_t>(waves, 1);
    const int64_t max_splits =
        std::max<int64_t>(1, ns / unit);
    sp = std::min(std::max<int64_t>(sp, 1), max_splits);
    sp = std::min<int64_t>(sp, 65535);
    int64_t ch = (ns + sp - 1) / sp;
    ch = (ch + unit - 1) / unit * unit;
    ch = std::max<int64_t>(ch, unit);
    splits = (int)std::max<int64_t>(1, (ns + ch - 1) / ch);
    chunk = (int)ch;
}

}  // namespace

double min_band(double delta) {
    return std::max({std::ldexp(1.0, -6), 16.0 * delta, 4000.0 * delta * delta});
}

double governor_band(double delta, int q) {
    double f = std::ldexp(1.0, (q >> 1) - 1);
    if (q & 1) f *= 1.4142135623730951;
    return std::max(min_band(delta) * f, 0.015625);
}

void FieldBatch::release() {
    tables_.release();
    work_.release();
    items_buf_.release();
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
    governed_ = false;
    n_total_ = n_total;
    nseg_ = (int)segs.size();
    for (int L = 0; L < kLevels; ++L) lstart_[L] = nunits_[L] = 0;
    ntiles_ = ngroups_ = nrecs_ = 0;
    // mixed precision: a fixed number of targets per thread, so each
    // wavefront's targets are the same whatever the launch shape (its
    // near/far choices are made for them)
    tpt_ = precision == kPrecisionMixed ? kMixedTpt
           : (tun.targets_per_thread == 1 || tun.targets_per_thread == 2 ||
              tun.targets_per_thread == 4)
               ? tun.targets_per_thread
               : 2;
    segs_.assign(nseg_, SegDev{});
    items_.clear();
    scan_.clear();
    pairs_ = evals_ = 0;
    symsegs_.clear();
    // segments in ascending order of their points (the step loop's are):
    // each point's segment is found on the device; else written here
    bool ascending = true;
    for (int d = 0, end = 0; d < nseg_; ++d) {
        if (segs[d].off < end) ascending = false;
        end = std::max(end, segs[d].off + segs[d].n);
    }
    std::vector<int32_t> seg_of(ascending ? 0 : std::max(n_total, 1), 0);
    int64_t part_total = 0;
    for (int d = 0; d < nseg_; ++d) {
        const BatchSeg& b = segs[d];
        if (b.off < 0 || b.n < 0 || b.off + b.n > n_total || b.src < 0 || b.src >= nseg_)
            return fail(FS_E_INVALID, "field batch: segment %d out of range", d);
        if (!(b.L > 0)) return fail(FS_E_INVALID, "field batch: segment %d has L <= 0", d);
        if (precision == kPrecisionMixed && b.eval && !(b.eps > 0))
            return fail(FS_E_INVALID, "field batch: mixed precision needs eps > 0");
        if (precision == kPrecisionMixed && b.eval && b.open)
            return fail(FS_E_INVALID, "field batch: open boundaries run in FP64 only");
        if (!ascending)
            for (int g = b.off; g < b.off + b.n; ++g) seg_of[g] = d;
        SegDev& S = segs_[d];
        S.off = b.off;
        S.n = b.n;
        S.src = b.src;
        S.src_off = segs[b.src].off;
        S.src_n = segs[b.src].n;
        S.eval = b.eval ? 1 : 0;
        S.open = b.open ? 1 : 0;
        S.invL = 1.0 / b.L;
        const double delta = b.eps / b.L;
        S.delta2 = delta * delta;
        S.C = 2.0 / std::sqrt(1.0 + 4.0 * S.delta2);
        if (b.open) {  // the open kernel in absolute units
            S.invL = 1.0;
            S.delta2 = b.eps * b.eps;
            S.C = 0.0;
        }
        S.pw_base = b.off + d;
        S.splits = 1;
        S.chunk = std::max(1, S.src_n);
        S.sym = -1;
        bool pair_once = precision == kPrecisionF64 && b.eval && b.src == d && b.n > 0 &&
                         use_sym(b.n, b.n, true, tun);
        sym::Plan SP{};
        if (pair_once) {
            SP = sym::plan(b.n, tun, INT64_MAX);  // one pass
            pair_once = (int64_t)sym::scratch_bytes(SP) <= sym::kBatchDeckLimit;
        }
        if (pair_once) {
            SymSeg E{};
            E.P = SP;
            E.S = sym::pass(SP, 0);
            E.seg = d;
            E.off = b.off;
            E.invL = S.invL;
            E.delta2 = S.delta2;
            E.open = S.open;
            S.sym = (int)symsegs_.size();  // renumbered below, grouped by targets per lane
            symsegs_.push_back(E);
            pairs_ += (int64_t)b.n * b.n;
            evals_ += sym::evals(E.P);
        } else if (b.eval && b.n > 0) {
            chunking(precision, b.n, S.src_n, num_cus, tun, S.splits, S.chunk);
            S.part_base = part_total;
            part_total += (int64_t)S.splits * b.n;
            pairs_ += (int64_t)b.n * S.src_n;
            evals_ += (int64_t)b.n * S.src_n;
        }
        if (precision == kPrecisionMixed) {
            const double two64 = 18446744073709551616.0;
            const double d2 = S.delta2;
            const double B = min_band(delta);
            S.all_band = (B >= 0.25) ? 1 : 0;
            S.band = S.all_band ? (uint64_t(1) << 62) : (uint64_t)std::ldexp(B, 64);
            S.all_band0 = S.all_band;
            S.band0 = S.band;
            S.xband = S.band >> kXShift;
            S.levels = unit_levels(S.src_n);
            S.log2h = (int)std::floor(std::log2(B));
            S.d2s = d2 * two64 * two64;
            const double kq2 = d2 * two64;
            S.b1z = S.all_band ? 0.0 : -0.75 * kq2;
            S.b2z = S.all_band ? 0.0 : 0.625 * kq2 * kq2;
        }
    }
    // pair-once segments grouped by targets per lane (one launch each), with
    // their workgroups and partial slots
    int64_t sym_elems = 0;
    {
        std::vector<SymSeg> grouped;
        for (int k = 0; k < 4; ++k) {
            const int t = 1 << k;
            sym_first_[k] = (int)grouped.size();
            sym_items_[k] = 0;
            for (const SymSeg& E0 : symsegs_) {
                if (E0.P.t != t) continue;
                SymSeg E = E0;
                E.item_base = sym_items_[k];
                sym_items_[k] += E.S.nitems;
                E.rbase = sym_elems;
                sym_elems += sym::rowpart_elems(E.P);
                E.cbase = sym_elems;
                sym_elems += sym::colpart_elems(E.P);
                segs_[E.seg].sym = (int)grouped.size();
                grouped.push_back(E);
            }
            sym_count_[k] = (int)grouped.size() - sym_first_[k];
        }
        symsegs_ = grouped;
    }
    if (precision == kPrecisionMixed) {
        // scan blocks of every segment that serves as a source
        std::vector<char> is_src(nseg_, 0);
        for (int d = 0; d < nseg_; ++d)
            if (segs[d].eval) is_src[segs[d].src] = 1;
        ntiles_ = ngroups_ = nrecs_ = 0;
        for (int d = 0; d < nseg_; ++d) {
            segs_[d].is_src = is_src[d];
            segs_[d].tile_base = ntiles_;
            segs_[d].group_base = ngroups_;
            if (is_src[d]) {
                ntiles_ += (segs_[d].n + kTile - 1) / kTile;
                ngroups_ += (segs_[d].n + kGroupSrc - 1) / kGroupSrc;
            }
        }
        // far-field units: records after every group, level by level
        int nrec = ngroups_;
        for (int L = 0; L < kLevels; ++L) {
            lstart_[L] = nrec;
            for (int d = 0; d < nseg_; ++d) {
                segs_[d].unit_base[L] = nrec;
                if (is_src[d]) nrec += (segs_[d].n + unit_src(L) - 1) / unit_src(L);
            }
            nunits_[L] = nrec - lstart_[L];
        }
        nrecs_ = nrec;
        for (int d = 0; d < nseg_; ++d) {
            segs_[d].src_tile_base = segs_[segs_[d].src].tile_base;
            segs_[d].src_group_base = segs_[segs_[d].src].group_base;
            for (int L = 0; L < kLevels; ++L)
                segs_[d].src_unit_base[L] = segs_[segs_[d].src].unit_base[L];
        }
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
    // segment of each tile and of each group (mixed precision)
    std::vector<int32_t> tile_seg(std::max(ntiles_, 1), 0), grp_seg(std::max(ngroups_, 1), 0);
    std::vector<int32_t> unit_seg(std::max(nrecs_ - ngroups_, 1), 0);
    if (precision == kPrecisionMixed)
        for (int d = 0; d < nseg_; ++d) {
            if (!segs_[d].is_src) continue;
            const int nt = (segs_[d].n + kTile - 1) / kTile;
            const int ng = (segs_[d].n + kGroupSrc - 1) / kGroupSrc;
            for (int t = 0; t < nt; ++t) tile_seg[segs_[d].tile_base + t] = d;
            for (int g = 0; g < ng; ++g) grp_seg[segs_[d].group_base + g] = d;
            for (int L = 0; L < kLevels; ++L) {
                const int nu = (segs_[d].n + unit_src(L) - 1) / unit_src(L);
                for (int u = 0; u < nu; ++u) unit_seg[segs_[d].unit_base[L] - ngroups_ + u] = d;
            }
        }
    // device tables
    const size_t b_segs = align_up(sizeof(SegDev) * nseg_, 256);
    const size_t b_scan = align_up(sizeof(ScanBlock) * std::max<size_t>(scan_.size(), 1), 256);
    const size_t b_segof = align_up(sizeof(int32_t) * std::max(n_total, 1), 256);
    const size_t b_sym = align_up(sizeof(SymSeg) * std::max<size_t>(symsegs_.size(), 1), 256);
    const size_t b_tseg = align_up(sizeof(int32_t) * tile_seg.size(), 256);
    const size_t b_gseg = align_up(sizeof(int32_t) * grp_seg.size(), 256);
    const size_t b_useg = align_up(sizeof(int32_t) * unit_seg.size(), 256);
    FS_TRY(grow(tables_, b_segs + b_scan + b_segof + b_sym + b_tseg + b_gseg + b_useg + 256,
                stream));
    d_segs_ = tables_.as<SegDev>(0);
    d_scan_ = tables_.as<ScanBlock>(b_segs);
    d_seg_of_ = tables_.as<int32_t>(b_segs + b_scan);
    d_symsegs_ = tables_.as<SymSeg>(b_segs + b_scan + b_segof);
    d_tile_seg_ = tables_.as<int32_t>(b_segs + b_scan + b_segof + b_sym);
    d_grp_seg_ = tables_.as<int32_t>(b_segs + b_scan + b_segof + b_sym + b_tseg);
    d_unit_seg_ = tables_.as<int32_t>(b_segs + b_scan + b_segof + b_sym + b_tseg + b_gseg);
    d_counter_ = tables_.as<int>(b_segs + b_scan + b_segof + b_sym + b_tseg + b_gseg + b_useg);
    FS_HIP_CHECK(hipMemcpyAsync(d_unit_seg_, unit_seg.data(), sizeof(int32_t) * unit_seg.size(),
                                hipMemcpyHostToDevice, stream));
    FS_HIP_CHECK(hipMemcpyAsync(d_tile_seg_, tile_seg.data(), sizeof(int32_t) * tile_seg.size(),
                                hipMemcpyHostToDevice, stream));
    FS_HIP_CHECK(hipMemcpyAsync(d_grp_seg_, grp_seg.data(), sizeof(int32_t) * grp_seg.size(),
                                hipMemcpyHostToDevice, stream));
    if (!symsegs_.empty())
        FS_HIP_CHECK(hipMemcpyAsync(d_symsegs_, symsegs_.data(), sizeof(SymSeg) * symsegs_.size(),
                                    hipMemcpyHostToDevice, stream));
    FS_HIP_CHECK(hipMemcpyAsync(d_segs_, segs_.data(), sizeof(SegDev) * nseg_,
                                hipMemcpyHostToDevice, stream));
    if (!scan_.empty())
        FS_HIP_CHECK(hipMemcpyAsync(d_scan_, scan_.data(), sizeof(ScanBlock) * scan_.size(),
                                    hipMemcpyHostToDevice, stream));
    if (!ascending) {
        FS_HIP_CHECK(hipMemcpyAsync(d_seg_of_, seg_of.data(), sizeof(int32_t) * seg_of.size(),
                                    hipMemcpyHostToDevice, stream));
    } else if (n_total > 0) {
        seg_of_kernel<<<nblocks(n_total), kBlock, 0, stream>>>(d_segs_, nseg_, n_total, d_seg_of_);
        FS_HIP_CHECK(hipGetLastError());
    }
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
    o_sympart_ = take(8 * (size_t)std::max<int64_t>(sym_elems, 1));
    if (precision == kPrecisionMixed) {
        o_key_ = take(8 * n);
        o_skey_ = take(8 * n);
        o_perm_ = take(4 * n);
        o_sort_ = take(radix_sort_bytes(n_total));
        o_as_ = take(8 * n);
        o_asm1_ = take(8 * n);
        o_ws_ = take(8 * n);
        o_far_ = take(sizeof(FarSrc) * n);
        o_pw_ = take(8 * (n + nseg_));
        o_bsum_ = take(8 * std::max<size_t>(scan_.size(), 1));
        o_bsa_ = take(8 * std::max<size_t>(scan_.size(), 1));
        o_tot_ = take(8 * (size_t)nseg_);
        o_gov_ = take(8 * (size_t)nseg_ * kGovVals);
        o_tiles_ = take(sizeof(TileRec) * (size_t)std::max(ntiles_, 1));
        o_groups_ = take(sizeof(GroupRec) * (size_t)std::max(nrecs_, 1));
        o_tiles64_ = take(sizeof(TileRec64) * (size_t)std::max(ntiles_, 1));
        o_moms_ = take(sizeof(TileMom) * (size_t)std::max(ntiles_, 1));
        o_gexp_ = take(sizeof(TileRec) * (size_t)std::max(nrecs_, 1));
        o_gexp64_ = take(sizeof(TileRec64) * (size_t)std::max(ngroups_, 1));
        o_gmoms_ = take(sizeof(TileMom) * (size_t)std::max(nrecs_, 1));
    }
    FS_TRY(grow(work_, o, stream));
    if (precision == kPrecisionMixed) {
        zero_u64_kernel<<<nblocks(nseg_ * kGovVals), kBlock, 0, stream>>>(gov(), nseg_ * kGovVals);
        FS_HIP_CHECK(hipGetLastError());
    }
    num_cus_ = num_cus;
    // the host tables must outlive the copies
    FS_HIP_CHECK(hipStreamSynchronize(stream));
    return set_shape(tpt_, wgs_per_cu_, stream);
}

// Work items for blocks of 256*tpt targets, largest first, and the
// persistent grid (wgs_per_cu workgroups per CU, at most one per item).
void FieldBatch::build_items(int tpt, std::vector<WorkItem>& items) const {
    items.clear();
    const int tblock = kBlock * tpt;
    for (int d = 0; d < nseg_; ++d) {
        const SegDev& S = segs_[d];
        if (!S.eval || S.n == 0 || S.src_n == 0 || S.sym >= 0) continue;
        for (int tb = S.off; tb < S.off + S.n; tb += tblock)
            for (int s = 0; s < S.splits; ++s) {
                const int sb = S.src_off + s * S.chunk;
                const int se = std::min(S.src_off + S.src_n, sb + S.chunk);
                items.push_back(WorkItem{d, tb, std::min(S.off + S.n, tb + tblock), sb, se, s});
            }
    }
    std::stable_sort(items.begin(), items.end(), [](const WorkItem& a, const WorkItem& b) {
        return (int64_t)(a.te - a.tb) * (a.se - a.sb) > (int64_t)(b.te - b.tb) * (b.se - b.sb);
    });
}

int FieldBatch::upload_items(const std::vector<WorkItem>& items, hipStream_t stream) {
    FS_TRY(grow(items_buf_, sizeof(WorkItem) * std::max<size_t>(items.size(), 1), stream));
    d_items_ = items_buf_.as<WorkItem>();
    if (!items.empty())
        FS_HIP_CHECK(hipMemcpyAsync(d_items_, items.data(), sizeof(WorkItem) * items.size(),
                                    hipMemcpyHostToDevice, stream));
    FS_HIP_CHECK(hipStreamSynchronize(stream));
    return FS_OK;
}

int FieldBatch::set_shape(int tpt, int wgs_per_cu, hipStream_t stream) {
    if (tpt != 1 && tpt != 2 && tpt != 4) return fail(FS_E_INVALID, "field batch: tpt %d", tpt);
    tpt_ = precision_ == kPrecisionMixed ? kMixedTpt : tpt;
    wgs_per_cu_ = std::max(1, wgs_per_cu);
    build_items(tpt_, items_);
    FS_TRY(upload_items(items_, stream));
    grid_ = (int)std::min<size_t>(items_.size(), (size_t)wgs_per_cu_ * num_cus_);
    return FS_OK;
}

int FieldBatch::launch_pairs(int nit, int grid, hipStream_t stream) {
    if (nit <= 0 || grid <= 0) return FS_OK;
    char* wk = static_cast<char*>(work_.ptr);
    double* part = reinterpret_cast<double*>(wk + o_part_);
    if (precision_ == kPrecisionF64) {
        const double* x = x_;
        const double* w = w_;
        switch (tpt_) {
            case 1: batch_f64_kernel<1><<<grid, kBlock, 0, stream>>>(d_items_, nit, d_segs_, x, w, part, d_counter_); break;
            case 4: batch_f64_kernel<4><<<grid, kBlock, 0, stream>>>(d_items_, nit, d_segs_, x, w, part, d_counter_); break;
            default: batch_f64_kernel<2><<<grid, kBlock, 0, stream>>>(d_items_, nit, d_segs_, x, w, part, d_counter_); break;
        }
    } else {
        const uint64_t* skey = reinterpret_cast<const uint64_t*>(wk + o_skey_);
        const double* as = reinterpret_cast<const double*>(wk + o_as_);
        const double* asm1 = reinterpret_cast<const double*>(wk + o_asm1_);
        const double* ws = reinterpret_cast<const double*>(wk + o_ws_);
        const FarSrc* far = reinterpret_cast<const FarSrc*>(wk + o_far_);
        const TileRec* tiles = reinterpret_cast<const TileRec*>(wk + o_tiles_);
        const GroupRec* groups = reinterpret_cast<const GroupRec*>(wk + o_groups_);
        const TileMom* moms = reinterpret_cast<const TileMom*>(wk + o_moms_);
        const TileRec* gexp = reinterpret_cast<const TileRec*>(wk + o_gexp_);
        const TileMom* gmoms = reinterpret_cast<const TileMom*>(wk + o_gmoms_);
        batch_mixed_kernel<kMixedTpt><<<grid, kBlock, 0, stream>>>(
            d_items_, nit, d_segs_, skey, as, asm1, ws, far, tiles, groups, moms, gexp, gmoms, part,
            d_counter_);
    }
    FS_HIP_CHECK(hipGetLastError());
    return FS_OK;
}

// Times the pair kernel for every (tpt, workgroups per CU) on this batch's
// data and keeps the fastest. A trial runs all the launch's items with their
// source ranges shortened to about 2e10 pairs in total. The shape changes
// neither the items' summation order nor any result. es receives a full
// evaluation (it is scratch here) at the minimum band.
int FieldBatch::autotune(const double* x, const double* w, double* es, hipStream_t stream,
                         int* best_tpt, int* best_wgs) {
    const bool governed = governed_;
    governed_ = false;
    const int rc0 = run(x, w, es, stream);  // sorted arrays and plan tables in place
    governed_ = governed;
    FS_TRY(rc0);
    static const int tpts[3] = {1, 2, 4};
    static const int wgss[3] = {2, 4, 8};
    hipEvent_t e0, e1;
    FS_HIP_CHECK(hipEventCreate(&e0));
    FS_HIP_CHECK(hipEventCreate(&e1));
    double best_rate = -1.0;
    int bt = tpt_, bw = wgs_per_cu_;
    std::vector<WorkItem> trial;
    int rc = FS_OK;
    // mixed precision: tpt fixed (its near/far choices are made per item)
    const bool mixed = precision_ == kPrecisionMixed;
    const int a0 = mixed ? (kMixedTpt == 1 ? 0 : kMixedTpt == 2 ? 1 : 2) : 0;
    for (int a = a0; a < (mixed ? a0 + 1 : 3) && rc == FS_OK; ++a) {
        // every item of the real launch (so the grid fills as it will), each
        // cut to a source count that keeps the trial near 2e10 pairs
        build_items(tpts[a], trial);
        int64_t targets = 0;
        for (const WorkItem& it : trial) targets += it.te - it.tb;
        const int64_t cut = std::max<int64_t>(512, (int64_t)2e10 / std::max<int64_t>(targets, 1));
        // the windows are spread over each chunk (a window at the chunk start
        // would hold the near-band sources of most target blocks)
        int64_t pairs = 0;
        uint64_t hsh = 0x9e3779b97f4a7c15ull;
        for (WorkItem& it : trial) {
            const int64_t len = it.se - it.sb;
            if (len > cut) {
                hsh = hsh * 6364136223846793005ull + 1442695040888963407ull;
                const int64_t slots = (len - cut) / kTile + 1;
                it.sb += (int)((int64_t)((hsh >> 33) % (uint64_t)slots) * kTile);
                it.se = (int)(it.sb + cut);
            }
            pairs += (int64_t)(it.te - it.tb) * (it.se - it.sb);
        }
        tpt_ = tpts[a];
        if ((rc = upload_items(trial, stream)) != FS_OK) break;
        for (int b = 0; b < 3 && rc == FS_OK; ++b) {
            const int grid = (int)std::min<size_t>(trial.size(), (size_t)wgss[b] * num_cus_);
            float ms = 0.f;
            for (int rep = 0; rep < 2 && rc == FS_OK; ++rep) {
                zero_int_kernel<<<1, 1, 0, stream>>>(d_counter_);
                if (hipEventRecord(e0, stream) != hipSuccess) rc = fail(FS_E_HIP, "autotune: event");
                if (rc == FS_OK) rc = launch_pairs((int)trial.size(), grid, stream);
                if (rc == FS_OK && hipEventRecord(e1, stream) != hipSuccess) rc = fail(FS_E_HIP, "autotune: event");
                if (rc == FS_OK && hipEventSynchronize(e1) != hipSuccess) rc = fail(FS_E_HIP, "autotune: sync");
                if (rc == FS_OK) (void)hipEventElapsedTime(&ms, e0, e1);
            }
            const double rate = ms > 0 ? pairs / (double)ms : 0.0;
            if (rc == FS_OK && rate > best_rate) {
                best_rate = rate;
                bt = tpts[a];
                bw = wgss[b];
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    zero_int_kernel<<<1, 1, 0, stream>>>(d_counter_);
    FS_TRY(rc);
    FS_TRY(set_shape(bt, bw, stream));
    if (best_tpt) *best_tpt = bt;
    if (best_wgs) *best_wgs = bw;
    return FS_OK;
}

void FieldBatch::govern(bool on, double target) {
    governed_ = on && precision_ == kPrecisionMixed;
    gov_target_ = target > 0 ? target : kGovTarget;
}

uint64_t* FieldBatch::gov() const {
    if (precision_ != kPrecisionMixed || !work_.ptr) return nullptr;
    return reinterpret_cast<uint64_t*>(static_cast<char*>(work_.ptr) + o_gov_);
}

int FieldBatch::get_gov(std::vector<double>& out, hipStream_t stream) const {
    out.assign((size_t)nseg_ * kGovVals, 0.0);
    if (!gov() || nseg_ == 0) return FS_OK;
    FS_HIP_CHECK(hipMemcpyAsync(out.data(), gov(), sizeof(double) * out.size(),
                                hipMemcpyDeviceToHost, stream));
    FS_HIP_CHECK(hipStreamSynchronize(stream));
    return FS_OK;
}

int FieldBatch::run(const double* x, const double* w, double* es, hipStream_t stream) {
    if (n_total_ == 0) return FS_OK;
    char* wk = static_cast<char*>(work_.ptr);
    double* part = reinterpret_cast<double*>(wk + o_part_);
    const int nit = (int)items_.size();
    x_ = x;
    w_ = w;
    if (precision_ == kPrecisionF64) {
        FS_TRY(launch_pairs(nit, grid_, stream));
        double* sympart = reinterpret_cast<double*>(wk + o_sympart_);
        if (sym_items_[0] > 0)
            batch_sym_kernel<1><<<sym_items_[0], sym::threads<1>(), 0, stream>>>(
                d_symsegs_, sym_first_[0], sym_count_[0], x, w, sympart);
        if (sym_items_[1] > 0)
            batch_sym_kernel<2><<<sym_items_[1], sym::threads<2>(), 0, stream>>>(
                d_symsegs_, sym_first_[1], sym_count_[1], x, w, sympart);
        if (sym_items_[2] > 0)
            batch_sym_kernel<4><<<sym_items_[2], sym::threads<4>(), 0, stream>>>(
                d_symsegs_, sym_first_[2], sym_count_[2], x, w, sympart);
        if (sym_items_[3] > 0)
            batch_sym_kernel<8><<<sym_items_[3], sym::threads<8>(), 0, stream>>>(
                d_symsegs_, sym_first_[3], sym_count_[3], x, w, sympart);
        FS_HIP_CHECK(hipGetLastError());
        finalize_f64_kernel<<<nblocks(n_total_), kBlock, 0, stream>>>(
            n_total_, d_seg_of_, d_segs_, part, d_symsegs_, sympart, x, es, d_counter_);
        FS_HIP_CHECK(hipGetLastError());
        return FS_OK;
    }
    uint64_t* key = reinterpret_cast<uint64_t*>(wk + o_key_);
    uint64_t* skey = reinterpret_cast<uint64_t*>(wk + o_skey_);
    int32_t* perm = reinterpret_cast<int32_t*>(wk + o_perm_);
    double* as = reinterpret_cast<double*>(wk + o_as_);
    double* asm1 = reinterpret_cast<double*>(wk + o_asm1_);
    double* ws = reinterpret_cast<double*>(wk + o_ws_);
    FarSrc* far = reinterpret_cast<FarSrc*>(wk + o_far_);
    double* pw = reinterpret_cast<double*>(wk + o_pw_);
    double* bsum = reinterpret_cast<double*>(wk + o_bsum_);
    double* bsa = reinterpret_cast<double*>(wk + o_bsa_);
    double* tot = reinterpret_cast<double*>(wk + o_tot_);
    prep_kernel<<<nblocks(n_total_), kBlock, 0, stream>>>(n_total_, d_seg_of_, d_segs_, x, key);
    FS_HIP_CHECK(hipGetLastError());
    FS_TRY(radix_sort_seg(key, n_total_, d_seg_of_, nseg_, skey, perm, wk + o_sort_, stream));
    gather_kernel<<<nblocks(n_total_), kBlock, 0, stream>>>(n_total_, d_seg_of_, d_segs_, skey,
                                                            perm, w, as, asm1, ws, far);
    FS_HIP_CHECK(hipGetLastError());
    TileRec* tiles = reinterpret_cast<TileRec*>(wk + o_tiles_);
    if (ntiles_ > 0) {
        tile_kernel<<<(unsigned)((ntiles_ + kStageThreads - 1) / kStageThreads), kStageThreads, 0,
                      stream>>>(
            ntiles_, d_tile_seg_, d_segs_, skey, ws, tiles,
            reinterpret_cast<TileRec64*>(wk + o_tiles64_), reinterpret_cast<TileMom*>(wk + o_moms_));
        FS_HIP_CHECK(hipGetLastError());
    }
    if (ngroups_ > 0)
    group_kernel<<<(unsigned)((4 * (int64_t)ngroups_ + kStageThreads - 1) / kStageThreads),
                   kStageThreads, 0, stream>>>(
        ngroups_, d_grp_seg_, d_segs_, skey, ws, tiles, reinterpret_cast<GroupRec*>(wk + o_groups_),
        reinterpret_cast<TileRec*>(wk + o_gexp_), reinterpret_cast<TileRec64*>(wk + o_gexp64_),
        reinterpret_cast<TileMom*>(wk + o_gmoms_));
    FS_HIP_CHECK(hipGetLastError());
    for (int L = 0; L < kLevels; ++L) {
        if (nunits_[L] == 0) continue;
        unit_kernel<<<nblocks(64 * (int64_t)nunits_[L]), kBlock, 0, stream>>>(
            L, lstart_[L], nunits_[L], ngroups_, d_unit_seg_, d_segs_, skey, ws,
            reinterpret_cast<GroupRec*>(wk + o_groups_), reinterpret_cast<TileRec*>(wk + o_gexp_),
            reinterpret_cast<TileMom*>(wk + o_gmoms_));
        FS_HIP_CHECK(hipGetLastError());
    }
    if (!scan_.empty()) {
        blocksum_kernel<<<(unsigned)scan_.size(), kBlock, 0, stream>>>(d_scan_, ws, as, bsum, bsa);
        FS_HIP_CHECK(hipGetLastError());
        scan_kernel<<<(unsigned)scan_.size(), kBlock, 0, stream>>>(d_scan_, d_segs_, ws, bsum, bsa,
                                                                   pw, tot);
        FS_HIP_CHECK(hipGetLastError());
    }
    if (governed_ && probe_) {
        zero_u64_kernel<<<nblocks(nseg_ * kGovVals), kBlock, 0, stream>>>(gov(), nseg_ * kGovVals);
        FS_HIP_CHECK(hipGetLastError());
        probe_kernel<<<dim3(kProbeMax, nseg_), kBlock, 0, stream>>>(
            d_segs_, skey, ws, far, tiles, reinterpret_cast<TileRec64*>(wk + o_tiles64_),
            reinterpret_cast<GroupRec*>(wk + o_groups_), reinterpret_cast<TileRec*>(wk + o_gexp_),
            reinterpret_cast<TileRec64*>(wk + o_gexp64_), gov());
        FS_HIP_CHECK(hipGetLastError());
        governor_kernel<<<nblocks(nseg_), kBlock, 0, stream>>>(nseg_, d_segs_, gov(), gov_target_);
        FS_HIP_CHECK(hipGetLastError());
    }
    FS_TRY(launch_pairs(nit, grid_, stream));
    finalize_mixed_kernel<<<nblocks(n_total_), kBlock, 0, stream>>>(
        n_total_, d_seg_of_, d_segs_, skey, perm, ws, pw, tot, part, es, d_counter_);
    FS_HIP_CHECK(hipGetLastError());
    return FS_OK;
}

}  // namespace fs
