// Stable LSD radix sort of points by (segment, top 24 bits of a 64-bit key).
//
// The mixed-precision field sorts positions (64-bit fixed point) within each
// deck of a batch. Ordering by the top 24 bits (2^-24 of the period) is
// enough for its uses (range tests widen by one such unit; the prefix-sum cut
// points compare full keys within a run of equal top bits), and every pass is
// stable, so equal keys keep index order and the result is deterministic. A pass is
// per-block digit histograms, an exclusive scan over (digit, block) (three
// small kernels), and a stable scatter (wavefront ballots rank equal digits,
// per-wave counts in LDS order the waves). Nothing here uses memset or host
// synchronization, so the sort can be captured in a HIP graph.
#include "fs_internal.h"

namespace fs {
namespace {

constexpr int kBlock = 256;
constexpr int kRounds = 8;
constexpr int kItems = kBlock * kRounds;  // items per block and pass
constexpr int kScanThreads = 1024;

struct Digit {
    int shift;               // key bit of the digit (key passes)
    const int32_t* seg_of;   // segment of each original index (segment passes)
    int seg_shift;
    __device__ __forceinline__ int operator()(uint64_t k, int32_t id) const {
        if (seg_of) return (seg_of[id] >> seg_shift) & 255;
        return (int)((k >> shift) & 255);
    }
};

__global__ __launch_bounds__(kBlock) void radix_hist(const uint64_t* __restrict__ kin,
                                                     const int32_t* __restrict__ iin, int n,
                                                     Digit dg, int nb, int* __restrict__ hist) {
    __shared__ int cnt[256];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    for (int r = 0; r < kRounds; ++r) {
        const int i = blockIdx.x * kItems + r * kBlock + (int)threadIdx.x;
        if (i < n) atomicAdd(&cnt[dg(kin[i], iin ? iin[i] : i)], 1);
    }
    __syncthreads();
    hist[threadIdx.x * nb + blockIdx.x] = cnt[threadIdx.x];
}

// The exclusive scan of the (digit, block) counts: block sums of 4096
// counts, one workgroup's scan of those sums, then every block's own scan
// (all loads and stores coalesced).
constexpr int kScanItems = kBlock * 16;

__global__ __launch_bounds__(kBlock) void scan_reduce(const int* __restrict__ a, int m,
                                                      int* __restrict__ partial) {
    __shared__ int red[kBlock];
    const int b0 = blockIdx.x * kScanItems;
    int s = 0;
    for (int r = 0; r < kScanItems / kBlock; ++r) {
        const int i = b0 + r * kBlock + (int)threadIdx.x;
        if (i < m) s += a[i];
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = kBlock / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(kBlock) void scan_apply(int* __restrict__ a, int m,
                                                     const int* __restrict__ partial) {
    constexpr int kPer = kScanItems / kBlock;
    __shared__ int buf[kScanItems];
    __shared__ int tot[kBlock];
    const int b0 = blockIdx.x * kScanItems;
    for (int r = 0; r < kPer; ++r) {
        const int i = b0 + r * kBlock + (int)threadIdx.x;
        buf[r * kBlock + threadIdx.x] = (i < m) ? a[i] : 0;
    }
    __syncthreads();
    // each thread's kPer consecutive counts, then the threads' totals
    int s = 0;
    for (int k = 0; k < kPer; ++k) s += buf[threadIdx.x * kPer + k];
    tot[threadIdx.x] = s;
    __syncthreads();
    for (int h = 1; h < kBlock; h <<= 1) {
        const int add = ((int)threadIdx.x >= h) ? tot[threadIdx.x - h] : 0;
        __syncthreads();
        tot[threadIdx.x] += add;
        __syncthreads();
    }
    int run = partial[blockIdx.x] + tot[threadIdx.x] - s;
    for (int k = 0; k < kPer; ++k) {
        const int v = buf[threadIdx.x * kPer + k];
        buf[threadIdx.x * kPer + k] = run;
        run += v;
    }
    __syncthreads();
    for (int r = 0; r < kPer; ++r) {
        const int i = b0 + r * kBlock + (int)threadIdx.x;
        if (i < m) a[i] = buf[r * kBlock + threadIdx.x];
    }
}

// Exclusive scan of m ints in place, one workgroup (the block sums).
__global__ __launch_bounds__(kScanThreads) void radix_scan(int* __restrict__ a, int m) {
    __shared__ int part[kScanThreads];
    const int per = (m + kScanThreads - 1) / kScanThreads;
    const int b = threadIdx.x * per, e = min(m, b + per);
    int s = 0;
    for (int k = b; k < e; ++k) s += a[k];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int h = 1; h < kScanThreads; h <<= 1) {
        const int add = ((int)threadIdx.x >= h) ? part[threadIdx.x - h] : 0;
        __syncthreads();
        part[threadIdx.x] += add;
        __syncthreads();
    }
    int run = part[threadIdx.x] - s;
    for (int k = b; k < e; ++k) {
        const int v = a[k];
        a[k] = run;
        run += v;
    }
}

__global__ __launch_bounds__(kBlock) void radix_scatter(
    const uint64_t* __restrict__ kin, const int32_t* __restrict__ iin, int n, Digit dg, int nb,
    const int* __restrict__ offs, uint64_t* __restrict__ kout, int32_t* __restrict__ iout) {
    __shared__ int base[256], run[256];
    __shared__ int wc[kBlock / 64][256];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    base[t] = offs[t * nb + blockIdx.x];
    run[t] = 0;
    __syncthreads();
    for (int r = 0; r < kRounds; ++r) {
        const int i = blockIdx.x * kItems + r * kBlock + t;
        const bool valid = i < n;
        uint64_t k = 0;
        int32_t id = 0;
        int d = 0;
        if (valid) {
            k = kin[i];
            id = iin ? iin[i] : i;
            d = dg(k, id);
        }
        // lanes of this wavefront holding the same digit
        uint64_t m = __ballot(valid);
        for (int b = 0; b < 8; ++b) {
            const uint64_t bb = __ballot(valid && ((d >> b) & 1));
            m &= ((d >> b) & 1) ? bb : ~bb;
        }
        for (int v = 0; v < kBlock / 64; ++v) wc[v][t] = 0;
        __syncthreads();
        const int rank = __popcll(m & ((1ull << lane) - 1ull));
        if (valid && rank == 0) wc[w][d] = __popcll(m);
        __syncthreads();
        if (valid) {
            int before = 0;
            for (int v = 0; v < w; ++v) before += wc[v][d];
            const int pos = base[d] + run[d] + before + rank;
            kout[pos] = k;
            iout[pos] = id;
        }
        __syncthreads();
        int add = 0;
        for (int v = 0; v < kBlock / 64; ++v) add += wc[v][t];
        run[t] += add;
        __syncthreads();
    }
}

constexpr int kKeyPasses = 3;  // key bits 40..63
int passes_for(int key_passes, int nseg) {
    return key_passes + (nseg > 1 ? 1 : 0) + (nseg > 256 ? 1 : 0);
}

}  // namespace

size_t radix_sort_bytes(int n) {
    if (n <= 0) return 0;
    const size_t nb = (size_t)(n + kItems - 1) / kItems;
    const size_t ns = (256 * nb + kScanItems - 1) / kScanItems;
    return align_up(8 * (size_t)n, 256) + align_up(4 * (size_t)n, 256) +
           align_up(4 * 256 * nb, 256) + align_up(4 * ns, 256);
}

namespace {

// Sorts by (segment, top 8*key_passes key bits).
int radix_sort_bits(const uint64_t* kin, int n, const int32_t* seg_of, int nseg, int key_passes,
                    uint64_t* kout, int32_t* iout, void* tmp, hipStream_t stream) {
    if (n <= 0) return FS_OK;
    if (nseg < 1 || nseg > 65536) return fail(FS_E_INVALID, "radix sort: 1..65536 segments");
    const int nb = (n + kItems - 1) / kItems;
    char* p = static_cast<char*>(tmp);
    uint64_t* kb = reinterpret_cast<uint64_t*>(p);
    p += align_up(8 * (size_t)n, 256);
    int32_t* ib = reinterpret_cast<int32_t*>(p);
    p += align_up(4 * (size_t)n, 256);
    int* hist = reinterpret_cast<int*>(p);
    p += align_up(4 * 256 * (size_t)nb, 256);
    int* partial = reinterpret_cast<int*>(p);
    const int m = 256 * nb;
    const int ns = (m + kScanItems - 1) / kScanItems;
    const int passes = passes_for(key_passes, nseg);
    // ping-pong so that the last pass writes kout/iout
    const uint64_t* ksrc = kin;
    const int32_t* isrc = nullptr;  // identity on the first pass
    for (int q = 0; q < passes; ++q) {
        const bool to_out = ((passes - 1 - q) % 2) == 0;
        uint64_t* kdst = to_out ? kout : kb;
        int32_t* idst = to_out ? iout : ib;
        Digit dg{0, nullptr, 0};
        if (q < key_passes) {
            dg.shift = 64 - 8 * (key_passes - q);
        } else {
            dg.seg_of = seg_of;
            dg.seg_shift = 8 * (q - key_passes);
        }
        radix_hist<<<nb, kBlock, 0, stream>>>(ksrc, isrc, n, dg, nb, hist);
        FS_HIP_CHECK(hipGetLastError());
        scan_reduce<<<ns, kBlock, 0, stream>>>(hist, m, partial);
        FS_HIP_CHECK(hipGetLastError());
        radix_scan<<<1, kScanThreads, 0, stream>>>(partial, ns);
        FS_HIP_CHECK(hipGetLastError());
        scan_apply<<<ns, kBlock, 0, stream>>>(hist, m, partial);
        FS_HIP_CHECK(hipGetLastError());
        radix_scatter<<<nb, kBlock, 0, stream>>>(ksrc, isrc, n, dg, nb, hist, kdst, idst);
        FS_HIP_CHECK(hipGetLastError());
        ksrc = kdst;
        isrc = idst;
    }
    return FS_OK;
}

}  // namespace

int radix_sort_seg(const uint64_t* kin, int n, const int32_t* seg_of, int nseg, uint64_t* kout,
                   int32_t* iout, void* tmp, hipStream_t stream) {
    return radix_sort_bits(kin, n, seg_of, nseg, kKeyPasses, kout, iout, tmp, stream);
}

int radix_sort_seg_full(const uint64_t* kin, int n, const int32_t* seg_of, int nseg,
                        uint64_t* kout, int32_t* iout, void* tmp, hipStream_t stream) {
    return radix_sort_bits(kin, n, seg_of, nseg, 8, kout, iout, tmp, stream);
}

int radix_sort_full(const uint64_t* kin, int n, uint64_t* kout, int32_t* iout, void* tmp,
                    hipStream_t stream) {
    return radix_sort_bits(kin, n, nullptr, 1, 8, kout, iout, tmp, stream);
}

int radix_sort_top(const uint64_t* kin, int n, int key_bytes, uint64_t* kout, int32_t* iout,
                   void* tmp, hipStream_t stream) {
    if (key_bytes < 1 || key_bytes > 8) return fail(FS_E_INVALID, "radix sort: 1..8 key bytes");
    return radix_sort_bits(kin, n, nullptr, 1, key_bytes, kout, iout, tmp, stream);
}

}  // namespace fs
