// Periodic barycentric-Lagrange treecode on the device (gfx950).
//
// Replaces E_MQ_Treecode::operator() (reference src/FieldStructure.cpp:
// 230-1050). The tree, the MAC lists and the Chebyshev points come from the
// host builder (fs_tree.cpp); the device computes
//   moments  one workgroup per node: M_k = sum_i lambda_i l_k(x_i) with the
//            barycentric weights (-1)^k delta_k and the exact-hit rule
//            (:775-866); per-thread partials reduced in thread order.
//   sums     one wavefront per (leaf, 64 targets): far clusters
//            (Call_BL, :904-961) then near leaves (Call_Ds, :963-1020);
//            cluster points/moments and near sources are uniform across the
//            wavefront, so they arrive through scalar loads like the direct
//            sum's sources.
//   scatter  back to input order (:480-486).
// The treecode kernel -s (1 - sqrt((L^2+4eps^2)/(4 s^2 L^2 + 4 eps^2))) is
// sqrt(1 + 4 delta^2) times the direct-sum kernel (delta = eps/L), so it is
// evaluated as z (1/sqrt(z^2+delta^2) - C) with the sums scaled once by
// sqrt(1 + 4 delta^2)/2 at the end. z = 0 and z = +-1/2 give exactly 0, as the
// reference's skips do.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <vector>

#include "fs_field_batch.h"
#include "fs_internal.h"
#include "fs_tree.h"

namespace fs {
namespace {

constexpr int kMaxPP = 16;

__device__ __forceinline__ double rsq_f64(double a) {
    const double y0 = __builtin_amdgcn_rsq(a);
    const double t = a * y0;
    const double e = __builtin_fma(-t, y0, 1.0);
    const double r = y0 * e;
    const double p = __builtin_fma(e, 0.375, 0.5);
    return __builtin_fma(r, p, y0);
}

__device__ __forceinline__ double kern(double xi, double xj, double invL, double delta2,
                                       double C) {
    const double s = (xi - xj) * invL;
    const double z = s - __builtin_rint(s);
    const double y = rsq_f64(__builtin_fma(z, z, delta2));
    return z * (y - C);
}

__global__ __launch_bounds__(256) void moments_kernel(const double* __restrict__ px,
                                                      const double* __restrict__ lam,
                                                      const int2* __restrict__ range,
                                                      const double* __restrict__ cheb, int PP,
                                                      double* __restrict__ mom) {
    __shared__ double part[kMaxPP][256];
    const int node = blockIdx.x + 1;  // the root has no moments
    const int2 r = range[node];
    const double* tk = cheb + (size_t)node * PP;
    double acc[kMaxPP];
    double w1[kMaxPP];
#pragma unroll
    for (int j = 0; j < kMaxPP; ++j) {
        acc[j] = 0.0;
        // (-1)^j delta_j, delta = 1/2 at both ends (:802-808)
        w1[j] = (j < PP) ? (((j % 2 == 0) ? 1.0 : -1.0) * ((j == 0 || j == PP - 1) ? 0.5 : 1.0))
                         : 0.0;
    }
    for (int k = r.x + (int)threadIdx.x; k <= r.y; k += 256) {
        const double x = px[k];
        double a[kMaxPP];
        double sum = 0.0;
        int exact = -1;
#pragma unroll
        for (int j = 0; j < kMaxPP; ++j) {
            a[j] = 0.0;
            if (j < PP) {
                const double dx = x - tk[j];
                if (fabs(dx) <= DBL_MIN) {
                    exact = j;
                } else {
                    a[j] = w1[j] / dx;
                    sum += a[j];
                }
            }
        }
        if (exact > -1) {
            sum = 1.0;
#pragma unroll
            for (int j = 0; j < kMaxPP; ++j) a[j] = (j == exact) ? 1.0 : 0.0;
        }
        const double D = 1.0 / sum;
        const double l = lam[k];
#pragma unroll
        for (int j = 0; j < kMaxPP; ++j) acc[j] += (a[j] * D) * l;
    }
#pragma unroll
    for (int j = 0; j < kMaxPP; ++j) part[j][threadIdx.x] = acc[j];
    __syncthreads();
    if ((int)threadIdx.x < PP) {
        double s = 0.0;
        for (int t = 0; t < 256; ++t) s += part[threadIdx.x][t];
        mom[(size_t)node * PP + threadIdx.x] = s;
    }
}

// One wavefront per (leaf, up to 64 of its targets): blocks[b] = (leaf, first
// target, last target) in tree order.
__global__ __launch_bounds__(64) void tree_sum_kernel(
    const double* __restrict__ px, const double* __restrict__ lam,
    const int2* __restrict__ range, const int4* __restrict__ blocks,
    const int32_t* __restrict__ far_off, const int32_t* __restrict__ far_list,
    const int32_t* __restrict__ near_off, const int32_t* __restrict__ near_list,
    const double* __restrict__ cheb, const double* __restrict__ mom, int PP, double invL,
    double delta2, double scale, double* __restrict__ e_tree) {
    const int4 b = blocks[blockIdx.x];
    const int leaf = b.x;
    const int i = b.y + (int)threadIdx.x;
    const bool active = i <= b.z;
    const double C = rsq_f64(__builtin_fma(0.5, 0.5, delta2));
    const double xi = px[active ? i : b.y];
    double far_acc = 0.0;
    for (int f = far_off[leaf]; f < far_off[leaf + 1]; ++f) {
        const int node = far_list[f];
        const double* tk = cheb + (size_t)node * PP;
        const double* mk = mom + (size_t)node * PP;
        for (int k = 0; k < PP; ++k)
            far_acc = __builtin_fma(mk[k], kern(xi, tk[k], invL, delta2, C), far_acc);
    }
    double near_acc = 0.0;
    for (int q = near_off[leaf]; q < near_off[leaf + 1]; ++q) {
        const int2 nr = range[near_list[q]];
        for (int j = nr.x; j <= nr.y; ++j)
            near_acc = __builtin_fma(lam[j], kern(xi, px[j], invL, delta2, C), near_acc);
    }
    if (active) e_tree[i] = far_acc * scale + near_acc * scale;
}

__global__ __launch_bounds__(256) void scatter_perm_kernel(const double* __restrict__ e_tree,
                                                           const int32_t* __restrict__ perm,
                                                           int n, double* __restrict__ es) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k < n) es[perm[k]] = e_tree[k];
}

__global__ __launch_bounds__(256) void gather_kernel(const double* __restrict__ src,
                                                     const int32_t* __restrict__ perm, int n,
                                                     double* __restrict__ dst, double L,
                                                     int do_fmod) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    const double v = src[perm[k]];
    dst[k] = do_fmod ? fmod(v, L) : v;
}

// Order-preserving 64-bit image of fmod(x, L) (-0 taken as +0): the bit
// pattern with the sign bit set for positive values, inverted for negative.
__global__ __launch_bounds__(256) void tree_key_kernel(const double* __restrict__ x, int n,
                                                       double L, uint64_t* __restrict__ key) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    double v = fmod(x[k], L);
    if (v == 0.0) v = 0.0;
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    key[k] = (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

double key_value(uint64_t key) {
    uint64_t b = (key >> 63) ? (key ^ 0x8000000000000000ull) : ~key;
    double v;
    std::memcpy(&v, &b, 8);
    return v;
}

}  // namespace

// Device treecode: d_x, d_w (input order) -> d_es (input order). The points
// are sorted by fmod(x, L) on the device (a stable sort, so equal
// coordinates keep input order, as the reference's stable partitions do);
// the tree is built on the host from the sorted order; everything else runs
// on `stream`. evals (optional) receives the far and near kernel-evaluation
// counts.
int treecode_f64(const double* d_x, const double* d_w, int n, double L, double eps, double mac,
                 int degree, int max_source, double* d_es, DevBuf& work, hipStream_t stream,
                 int64_t* evals) {
    if (n <= 0) return FS_OK;
    if (degree < 1 || degree + 1 > kMaxPP)
        return fail(FS_E_INVALID, "treecode: degree must be in [1, %d]", kMaxPP - 1);
    if (max_source < 1) return fail(FS_E_INVALID, "treecode: max_source must be >= 1");
    if (!(L > 0)) return fail(FS_E_INVALID, "treecode: L must be > 0");
    // the points in increasing fmod(x, L) (FieldStructure.cpp:240-247), stable
    std::vector<int32_t> order(n);
    std::vector<double> sx(n);
    {
        const size_t b_k = align_up(sizeof(uint64_t) * n, 256), b_i = align_up(sizeof(int32_t) * n, 256);
        FS_TRY(work.ensure(2 * b_k + b_i + radix_sort_bytes(n), stream));
        char* p = work.as<char>();
        uint64_t* key = reinterpret_cast<uint64_t*>(p);
        uint64_t* skey = reinterpret_cast<uint64_t*>(p + b_k);
        int32_t* ord = reinterpret_cast<int32_t*>(p + 2 * b_k);
        tree_key_kernel<<<(unsigned)((n + 255) / 256), 256, 0, stream>>>(d_x, n, L, key);
        FS_HIP_CHECK(hipGetLastError());
        FS_TRY(radix_sort_full(key, n, skey, ord, p + 2 * b_k + b_i, stream));
        std::vector<uint64_t> hk(n);
        FS_HIP_CHECK(hipMemcpyAsync(hk.data(), skey, sizeof(uint64_t) * n, hipMemcpyDeviceToHost,
                                    stream));
        FS_HIP_CHECK(hipMemcpyAsync(order.data(), ord, sizeof(int32_t) * n, hipMemcpyDeviceToHost,
                                    stream));
        FS_HIP_CHECK(hipStreamSynchronize(stream));
        for (int k = 0; k < n; ++k) sx[k] = key_value(hk[k]);
    }
    tree::Tree t;
    if (!tree::build_sorted(order, sx, L, mac, degree, max_source, t))
        return fail(FS_E_RANGE, "treecode: more than max_source coincident points");
    const int PP = degree + 1;
    const int nn = (int)t.nodes.size();
    const int nleaf = (int)t.leaves.size();
    std::vector<int2> range(nn);
    for (int k = 0; k < nn; ++k) range[k] = make_int2(t.nodes[k].m0, t.nodes[k].m1);
    std::vector<int4> blocks;
    for (int li = 0; li < nleaf; ++li) {
        const tree::Node& lf = t.nodes[t.leaves[li]];
        for (int s = lf.m0; s <= lf.m1; s += 64)
            blocks.push_back(make_int4(li, s, std::min(lf.m1, s + 63), 0));
    }
    // device layout
    auto al = [](size_t b) { return align_up(b, 256); };
    const size_t b_x = al(sizeof(double) * n), b_range = al(sizeof(int2) * nn),
                 b_blocks = al(sizeof(int4) * blocks.size()),
                 b_foff = al(sizeof(int32_t) * (nleaf + 1)),
                 b_flist = al(sizeof(int32_t) * std::max<size_t>(1, t.far_list.size())),
                 b_noff = al(sizeof(int32_t) * (nleaf + 1)),
                 b_nlist = al(sizeof(int32_t) * std::max<size_t>(1, t.near_list.size())),
                 b_cheb = al(sizeof(double) * t.cheb.size()), b_perm = al(sizeof(int32_t) * n);
    const size_t total = 3 * b_x + b_range + b_blocks + b_foff + b_flist + b_noff + b_nlist +
                         2 * b_cheb + b_perm;
    FS_TRY(work.ensure(total, stream));
    char* p = work.as<char>();
    double* px_t = reinterpret_cast<double*>(p); p += b_x;
    double* lam_t = reinterpret_cast<double*>(p); p += b_x;
    double* e_t = reinterpret_cast<double*>(p); p += b_x;
    int2* d_range = reinterpret_cast<int2*>(p); p += b_range;
    int4* d_blocks = reinterpret_cast<int4*>(p); p += b_blocks;
    int32_t* d_foff = reinterpret_cast<int32_t*>(p); p += b_foff;
    int32_t* d_flist = reinterpret_cast<int32_t*>(p); p += b_flist;
    int32_t* d_noff = reinterpret_cast<int32_t*>(p); p += b_noff;
    int32_t* d_nlist = reinterpret_cast<int32_t*>(p); p += b_nlist;
    double* d_cheb = reinterpret_cast<double*>(p); p += b_cheb;
    double* d_mom = reinterpret_cast<double*>(p); p += b_cheb;
    int32_t* d_perm = reinterpret_cast<int32_t*>(p);
    FS_HIP_CHECK(hipMemcpyAsync(d_range, range.data(), sizeof(int2) * nn, hipMemcpyHostToDevice, stream));
    FS_HIP_CHECK(hipMemcpyAsync(d_blocks, blocks.data(), sizeof(int4) * blocks.size(),
                                hipMemcpyHostToDevice, stream));
    FS_HIP_CHECK(hipMemcpyAsync(d_foff, t.far_off.data(), sizeof(int32_t) * (nleaf + 1),
                                hipMemcpyHostToDevice, stream));
    if (!t.far_list.empty())
        FS_HIP_CHECK(hipMemcpyAsync(d_flist, t.far_list.data(), sizeof(int32_t) * t.far_list.size(),
                                    hipMemcpyHostToDevice, stream));
    FS_HIP_CHECK(hipMemcpyAsync(d_noff, t.near_off.data(), sizeof(int32_t) * (nleaf + 1),
                                hipMemcpyHostToDevice, stream));
    FS_HIP_CHECK(hipMemcpyAsync(d_nlist, t.near_list.data(), sizeof(int32_t) * t.near_list.size(),
                                hipMemcpyHostToDevice, stream));
    FS_HIP_CHECK(hipMemcpyAsync(d_cheb, t.cheb.data(), sizeof(double) * t.cheb.size(),
                                hipMemcpyHostToDevice, stream));
    FS_HIP_CHECK(hipMemcpyAsync(d_perm, t.perm.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice,
                                stream));
    const unsigned gb = (unsigned)((n + 255) / 256);
    gather_kernel<<<gb, 256, 0, stream>>>(d_x, d_perm, n, px_t, L, 1);
    gather_kernel<<<gb, 256, 0, stream>>>(d_w, d_perm, n, lam_t, L, 0);
    if (nn > 1) moments_kernel<<<nn - 1, 256, 0, stream>>>(px_t, lam_t, d_range, d_cheb, PP, d_mom);
    const double delta = eps / L, delta2 = delta * delta;
    const double scale = 0.5 * std::sqrt(1 + 4 * delta2);
    tree_sum_kernel<<<(unsigned)blocks.size(), 64, 0, stream>>>(
        px_t, lam_t, d_range, d_blocks, d_foff, d_flist, d_noff, d_nlist, d_cheb, d_mom, PP,
        1.0 / L, delta2, scale, e_t);
    scatter_perm_kernel<<<gb, 256, 0, stream>>>(e_t, d_perm, n, d_es);
    FS_HIP_CHECK(hipGetLastError());
    if (evals) {
        evals[0] = t.far_evals;
        evals[1] = t.near_evals;
    }
    // the host vectors above back the async copies
    FS_HIP_CHECK(hipStreamSynchronize(stream));
    return FS_OK;
}

}  // namespace fs

extern "C" {

int fs_emq_treecode_f64(fs_ctx* ctx, double* es, const double* sources, const double* q_ws,
                        int64_t ns, double L, double epsilon, double mac, int degree,
                        int max_source, int64_t* evals) {
    if (!ctx) return fs::fail(FS_E_INVALID, "null context");
    if (ns < 0 || ns > INT32_MAX) return fs::fail(FS_E_INVALID, "ns out of range");
    if (ns == 0) return FS_OK;
    if (!es || !sources || !q_ws) return fs::fail(FS_E_INVALID, "null pointer");
    FS_HIP_CHECK(hipSetDevice(ctx->device));
    const size_t b = fs::align_up(sizeof(double) * ns, 256);
    FS_TRY(ctx->io.ensure(3 * b, ctx->stream));
    double* d_x = ctx->io.as<double>(0);
    double* d_w = ctx->io.as<double>(b);
    double* d_e = ctx->io.as<double>(2 * b);
    FS_HIP_CHECK(hipMemcpyAsync(d_x, sources, sizeof(double) * ns, hipMemcpyHostToDevice, ctx->stream));
    FS_HIP_CHECK(hipMemcpyAsync(d_w, q_ws, sizeof(double) * ns, hipMemcpyHostToDevice, ctx->stream));
    FS_TRY(fs::treecode_f64(d_x, d_w, (int)ns, L, epsilon, mac, degree, max_source, d_e,
                            ctx->scratch, ctx->stream, evals));
    FS_HIP_CHECK(hipMemcpyAsync(es, d_e, sizeof(double) * ns, hipMemcpyDeviceToHost, ctx->stream));
    FS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return FS_OK;
}

}  // extern "C"
