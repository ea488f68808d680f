// Per-species diagnostics as device reductions (the quantities of the
// reference's standard_interface/diagnostic_plot.py:116-143).
//
// Sums over a species' points (charge, momentum, kinetic energy, entropy,
// L1, L2, negative area and count, min/max f) are blocked reductions in a
// fixed order: each thread adds its points in sequence, each workgroup its
// threads through LDS in a fixed tree, and one thread the workgroups in
// order, so a result depends only on the species' own points.
//
// The electric energy 1/2 sum_k dx_k E_k^2 runs over the distinct x in
// [0, L) in increasing order, each with E at its first occurrence, dx_k half
// the distance between its neighbours (periodically across the ends). The
// points are sorted by x with the library's stable 64-bit radix sort (equal
// x keep index order, so a run of equal keys starts at the first
// occurrence, as np.unique(..., return_index=True) finds it); the in-range
// values are one contiguous range of the sorted keys, and each run start
// finds its neighbouring runs by binary search, so no scan is needed.
#include <algorithm>
#include <cstring>
#include <vector>

#include "fs_field_batch.h"
#include "fs_internal.h"

namespace fs {
namespace diag {
namespace {

constexpr int kBlock = 256;
constexpr int kPer = 8;                  // points per thread
constexpr int kChunk = kBlock * kPer;    // points per workgroup
constexpr int kSums = 10;  // charge, mom, kin, entropy, l1, l2, neg, nneg, min, max

struct Part {
    double v[kSums];
};

__device__ __forceinline__ uint64_t order_key(double v) {
    if (v == 0.0) v = 0.0;  // one key for +0 and -0 (they compare equal)
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

// Fixed-order LDS reduction of one value per thread (sum, or min/max).
template <int kOp>
__device__ __forceinline__ double block_reduce(double v, double* sh) {
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            const double a = sh[threadIdx.x], b = sh[threadIdx.x + s];
            sh[threadIdx.x] = kOp == 0 ? a + b : kOp == 1 ? (b < a ? b : a) : (b > a ? b : a);
        }
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(kBlock) void sums_kernel(const double* __restrict__ ps,
                                                      const double* __restrict__ fs,
                                                      const double* __restrict__ qw, int n,
                                                      Part* __restrict__ parts) {
    __shared__ double sh[kBlock];
    double s[kSums] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    s[8] = __builtin_huge_val();
    s[9] = -__builtin_huge_val();
    const int base = blockIdx.x * kChunk + threadIdx.x;
    for (int j = 0; j < kPer; ++j) {
        const int i = base + j * kBlock;
        if (i >= n) break;
        const double p = ps[i], f = fs[i], w = qw[i];
        s[0] += w;
        s[1] += p * w;
        s[2] += p * p * w;
        if (f > 0) s[3] += f * log10(f) * w;
        s[4] += fabs(w);
        s[5] += w * f;
        if (f < 0) {
            s[6] += w;
            s[7] += 1.0;
        }
        s[8] = f < s[8] ? f : s[8];
        s[9] = f > s[9] ? f : s[9];
    }
    Part out;
    for (int k = 0; k < 8; ++k) out.v[k] = block_reduce<0>(s[k], sh);
    out.v[8] = block_reduce<1>(s[8], sh);
    out.v[9] = block_reduce<2>(s[9], sh);
    if (threadIdx.x == 0) parts[blockIdx.x] = out;
}

__global__ void key_kernel(const double* __restrict__ x, int n, uint64_t* __restrict__ key) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) key[i] = order_key(x[i]);
}

// first index in [lo, hi) whose key is >= k (lower bound)
__device__ __forceinline__ int lower(const uint64_t* __restrict__ a, int lo, int hi, uint64_t k) {
    while (lo < hi) {
        const int m = (lo + hi) >> 1;
        if (a[m] < k) lo = m + 1;
        else hi = m;
    }
    return lo;
}

// The in-range keys [key(0), key(L)) of the sorted array: ab = their first
// and one-past-last positions.
__global__ void bounds_kernel(const uint64_t* __restrict__ skey, int n, uint64_t k0, uint64_t kL,
                              int* __restrict__ ab) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    ab[0] = lower(skey, 0, n, k0);
    ab[1] = lower(skey, 0, n, kL);
}

// Per sorted position: dx E^2 of a run start in [a, b) (0 elsewhere) and the
// run-start count, as the two sums of a Part (the other slots zero).
__global__ __launch_bounds__(kBlock) void energy_kernel(const uint64_t* __restrict__ skey,
                                                        const int32_t* __restrict__ perm,
                                                        const double* __restrict__ x,
                                                        const double* __restrict__ e, int n,
                                                        const int* __restrict__ ab, double L,
                                                        Part* __restrict__ parts) {
    __shared__ double sh[kBlock];
    const int a = ab[0], b = ab[1];
    double term = 0.0, runs = 0.0;
    const int base = blockIdx.x * kChunk + threadIdx.x;
    for (int j = 0; j < kPer; ++j) {
        const int k = base + j * kBlock;
        if (k >= n) break;
        if (k < a || k >= b || (k > a && skey[k] == skey[k - 1])) continue;
        runs += 1.0;
        const double ek = e[perm[k]];
        // the neighbouring distinct values, periodically across [a, b)
        const int nxt = lower(skey, k + 1, b, skey[k] + 1);  // start of the next run
        const int last = lower(skey, a, b, skey[b - 1]);      // start of the last run
        double dx;
        if (k == a) {
            const int prev_start = last;
            dx = .5 * (x[perm[nxt < b ? nxt : a]] + L - x[perm[prev_start]]);
            if (nxt >= b) dx = .5 * L;  // a single distinct value (not counted: m < 3)
        } else if (nxt >= b) {
            const int prev_start = lower(skey, a, k, skey[k - 1]);
            dx = .5 * (x[perm[a]] + L - x[perm[prev_start]]);
        } else {
            const int prev_start = lower(skey, a, k, skey[k - 1]);
            dx = .5 * (x[perm[nxt]] - x[perm[prev_start]]);
        }
        term += dx * ek * ek;
    }
    Part out;
    for (int q = 0; q < kSums; ++q) out.v[q] = 0.0;
    out.v[0] = block_reduce<0>(term, sh);
    out.v[1] = block_reduce<0>(runs, sh);
    if (threadIdx.x == 0) parts[blockIdx.x] = out;
}

// Workgroup partials added in order (one thread); min/max in slots 8, 9.
__global__ void final_kernel(const Part* __restrict__ parts, int nb, double* __restrict__ out) {
    if (threadIdx.x != 0) return;
    double s[kSums] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    s[8] = __builtin_huge_val();
    s[9] = -__builtin_huge_val();
    for (int b = 0; b < nb; ++b) {
        for (int k = 0; k < 8; ++k) s[k] += parts[b].v[k];
        s[8] = parts[b].v[8] < s[8] ? parts[b].v[8] : s[8];
        s[9] = parts[b].v[9] > s[9] ? parts[b].v[9] : s[9];
    }
    for (int k = 0; k < kSums; ++k) out[k] = s[k];
}

}  // namespace

size_t work_bytes(int n) {
    const size_t nb = (size_t)(n + kChunk - 1) / kChunk + 1;
    return 2 * align_up(8 * (size_t)n, 256) + align_up(4 * (size_t)n, 256) +
           radix_sort_bytes(n) + 2 * align_up(sizeof(Part) * nb, 256) + 256;
}

// out: FS_DIAG_COUNT values (farrsight_hip.h). All pointers device arrays of
// one species' n points; work: work_bytes(n) bytes. Synchronizes `stream`.
int species(const double* x, const double* p, const double* f, const double* w,
            const double* e, int n, double q, double qm, double L, void* work, double* out,
            hipStream_t stream) {
    for (int k = 0; k < FS_DIAG_COUNT; ++k) out[k] = 0.0;
    out[13] = n;
    if (n <= 0) return FS_OK;
    const int nb = (n + kChunk - 1) / kChunk;
    char* c = static_cast<char*>(work);
    uint64_t* key = reinterpret_cast<uint64_t*>(c);
    c += align_up(8 * (size_t)n, 256);
    uint64_t* skey = reinterpret_cast<uint64_t*>(c);
    c += align_up(8 * (size_t)n, 256);
    int32_t* perm = reinterpret_cast<int32_t*>(c);
    c += align_up(4 * (size_t)n, 256);
    void* rtmp = c;
    c += radix_sort_bytes(n);
    Part* parts = reinterpret_cast<Part*>(c);
    c += align_up(sizeof(Part) * (nb + 1), 256);
    Part* eparts = reinterpret_cast<Part*>(c);
    c += align_up(sizeof(Part) * (nb + 1), 256);
    double* dres = reinterpret_cast<double*>(c);  // 2 x kSums doubles

    sums_kernel<<<nb, kBlock, 0, stream>>>(p, f, w, n, parts);
    FS_HIP_CHECK(hipGetLastError());
    final_kernel<<<1, 64, 0, stream>>>(parts, nb, dres);
    FS_HIP_CHECK(hipGetLastError());
    key_kernel<<<(n + kBlock - 1) / kBlock, kBlock, 0, stream>>>(x, n, key);
    FS_HIP_CHECK(hipGetLastError());
    FS_TRY(radix_sort_full(key, n, skey, perm, rtmp, stream));
    // the in-range keys [key(0), key(L)) of the sorted array, found on the
    // device; no in-range key leaves every term of the energy zero
    auto host_key = [](double v) {
        if (v == 0.0) v = 0.0;
        uint64_t b;
        std::memcpy(&b, &v, 8);
        return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
    };
    int* ab = reinterpret_cast<int*>(dres + 2 * kSums);
    bounds_kernel<<<1, 64, 0, stream>>>(skey, n, host_key(0.0), host_key(L), ab);
    FS_HIP_CHECK(hipGetLastError());
    energy_kernel<<<nb, kBlock, 0, stream>>>(skey, perm, x, e, n, ab, L, eparts);
    FS_HIP_CHECK(hipGetLastError());
    final_kernel<<<1, 64, 0, stream>>>(eparts, nb, dres + kSums);
    FS_HIP_CHECK(hipGetLastError());
    double s[2 * kSums];
    FS_HIP_CHECK(hipMemcpyAsync(s, dres, sizeof(double) * 2 * kSums, hipMemcpyDeviceToHost,
                                stream));
    FS_HIP_CHECK(hipStreamSynchronize(stream));
    const double pot = s[kSums + 1] >= 3 ? 0.5 * s[kSums] : 0.0;
    out[0] = s[0];
    out[1] = 1.0 / q * s[1];
    out[2] = 1.0 / 2.0 / qm * s[2];
    out[3] = pot;
    out[4] = 1.0 / q * s[3];
    out[5] = 1.0 / std::fabs(q) * s[4];
    out[6] = 1.0 / q * s[5];
    out[7] = s[8];
    out[8] = s[9];
    out[9] = 1.0 / q * s[6];
    out[10] = s[7];
    out[11] = s[7] / n;
    out[12] = s[4] > 0 ? std::fabs(s[6]) / s[4] : 0.0;
    out[13] = n;
    return FS_OK;
}

}  // namespace diag
}  // namespace fs
