// Periodic barycentric-Lagrange treecode on the device (gfx950).
//
// Replaces E_MQ_Treecode::operator() (reference src/FieldStructure.cpp:
// 230-1050). The tree, the MAC lists, the tree order and the Chebyshev points
// are built on the device (fs_tree_build.hip); then
//   moments  one workgroup per chunk of a node's members: M_k =
//            sum_i lambda_i l_k(x_i) with the barycentric weights
//            (-1)^k delta_k and the exact-hit rule (:775-866); per-thread
//            partials reduced in thread order, chunks in chunk order.
//   sums     one wavefront per (leaf, 64 targets): far clusters
//            (Call_BL, :904-961) then near leaves (Call_Ds, :963-1020);
//            cluster points/moments and near sources are uniform across the
//            wavefront, so they arrive through scalar loads like the direct
//            sum's sources.
//   scatter  back to input order (:480-486).
// The treecode kernel -s (1 - sqrt((L^2+4eps^2)/(4 s^2 L^2 + 4 eps^2))) is
// sqrt(1 + 4 delta^2) times the direct-sum kernel (delta = eps/L), so it is
// evaluated as z (1/sqrt(z^2+delta^2) - C) with the sums scaled once by
// sqrt(1 + 4 delta^2)/2 at the end (groups that share one wrap take it in x
// units with their linear part summed once, see tree_sum_kernel). On the
// per-pair form z = 0 and z = +-1/2 give exactly 0, as the reference's skips
// do; on the grouped form to rounding.
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

// x W - M with the product rounded on its own (not fused), as the sums below
// round theirs.
__device__ __forceinline__ double lin_part(double x, double W, double M) {
#pragma clang fp contract(off)
    return x * W - M;
}

// The wrap shared by every pair of a group of targets in [t0, t1] and a
// group of sources in [s0, s1]: 0 (none), 1 (targets one period back), -1
// (one period on), 2 (no single wrap: per pair). A margin keeps ranges that
// touch the half period on the per-pair form.
__device__ __forceinline__ int group_wrap(double t0, double t1, double s0, double s1,
                                          double invL) {
    const double lo = (t0 - s1) * invL, hi = (t1 - s0) * invL;
    constexpr double m = 1e-9;
    if (hi < 0.5 - m && lo > -0.5 + m) return 0;
    if (lo > 0.5 + m && hi < 1.5 - m) return 1;
    if (hi < -0.5 - m && lo > -1.5 + m) return -1;
    return 2;
}

// Per node: (sum m_k, sum m_k t_k, min t_k, max t_k) of its moments m_k at
// its Chebyshev points t_k, k in order.
__global__ __launch_bounds__(256) void far_sums_kernel(int nnodes, int PP,
                                                       const double* __restrict__ cheb,
                                                       const double* __restrict__ mom,
                                                       double4* __restrict__ fs) {
#pragma clang fp contract(off)
    const int node = blockIdx.x * 256 + threadIdx.x;
    if (node >= nnodes) return;
    const double* t = cheb + (size_t)node * PP;
    const double* m = mom + (size_t)node * PP;
    double a = 0.0, b = 0.0, lo = t[0], hi = t[0];
    for (int k = 0; k < PP; ++k) {
        a += m[k];
        b += m[k] * t[k];
        lo = fmin(lo, t[k]);
        hi = fmax(hi, t[k]);
    }
    fs[node] = make_double4(a, b, lo, hi);
}

// Per leaf, by its first block of targets: (sum lambda, sum lambda x, min x,
// max x) of its points; 64 lanes over strided points, then a fixed-order
// tree of shuffles (lane 0's result).
__global__ __launch_bounds__(64) void leaf_sums_kernel(const int4* __restrict__ blocks,
                                                       const int32_t* __restrict__ leaves,
                                                       const int2* __restrict__ range,
                                                       const double* __restrict__ px,
                                                       const double* __restrict__ lam,
                                                       double4* __restrict__ ls) {
#pragma clang fp contract(off)
    const int4 b = blocks[blockIdx.x];
    const int node = leaves[b.x];
    const int2 r = range[node];
    if (b.y != r.x) return;
    double a = 0.0, m = 0.0, lo = __builtin_inf(), hi = -__builtin_inf();
    for (int j = r.x + (int)threadIdx.x; j <= r.y; j += 64) {
        a += lam[j];
        m += lam[j] * px[j];
        lo = fmin(lo, px[j]);
        hi = fmax(hi, px[j]);
    }
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_down(a, off);
        m += __shfl_down(m, off);
        lo = fmin(lo, __shfl_down(lo, off));
        hi = fmax(hi, __shfl_down(hi, off));
    }
    if (threadIdx.x == 0) ls[node] = make_double4(a, m, lo, hi);
}

// Modified Chebyshev moments (:775-866), one workgroup per chunk of
// kMomChunk member points of a node (the node found by binary search over
// the nodes' first chunks): M_k = sum_i lambda_i l_k(x_i) with the
// barycentric weights (-1)^k delta_k and the exact-hit rule; per-thread
// partials reduced in thread order into the chunk's partial moments. The
// big nodes near the root spread over many workgroups.
__global__ __launch_bounds__(256) void moments_kernel(const double* __restrict__ px,
                                                      const double* __restrict__ lam,
                                                      const int2* __restrict__ range,
                                                      const int32_t* __restrict__ mom_off,
                                                      int nnodes, const double* __restrict__ cheb,
                                                      int PP, double* __restrict__ mompart) {
    __shared__ double part[kMaxPP][256];
    const int item = blockIdx.x;
    // the node whose chunks contain this item: the last with mom_off <= item
    int lo = 1, hi = nnodes - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (mom_off[mid] <= item) lo = mid;
        else hi = mid - 1;
    }
    const int node = lo;
    const int2 r = range[node];
    const int k0 = r.x + (item - mom_off[node]) * tree::kMomChunk;
    const int k1 = min(r.y + 1, k0 + tree::kMomChunk);
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
    for (int k = k0 + (int)threadIdx.x; k < k1; k += 256) {
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
        mompart[(size_t)item * PP + threadIdx.x] = s;
    }
}

// A node's moments: its chunks' partial moments added in chunk order.
__global__ __launch_bounds__(256) void moments_reduce_kernel(const int32_t* __restrict__ mom_off,
                                                             int nnodes, int PP,
                                                             const double* __restrict__ mompart,
                                                             double* __restrict__ mom) {
    const int g = blockIdx.x * 256 + threadIdx.x;
    const int node = 1 + g / PP, j = g % PP;
    if (node >= nnodes) return;
    double s = 0.0;
    for (int c = mom_off[node]; c < mom_off[node + 1]; ++c) s += mompart[(size_t)c * PP + j];
    mom[(size_t)node * PP + j] = s;
}

// One wavefront per (leaf, up to 64 of its targets): blocks[b] = (leaf, first
// target, last target) in tree order.
// A far cluster or near leaf whose x range gives every pair with the
// block's targets the same wrap is evaluated in x units, K = d/sqrt(d^2 +
// eps^2) - C_x d with d = x_i - x_j (the targets moved by the period where the
// wrap asks), eps = delta L, C_x = C/L -- the same K as z/sqrt(z^2 + delta^2)
// - C z -- and its linear part for the whole group at once, C_x (x_i W - M)
// from the group's sums (far_sums, leaf_sums): no rint and no C per pair.
// Groups that straddle the half period take the per-pair form.
__global__ __launch_bounds__(64) void tree_sum_kernel(
    const double* __restrict__ px, const double* __restrict__ lam,
    const int2* __restrict__ range, const int4* __restrict__ blocks,
    const int32_t* __restrict__ leaves,
    const int32_t* __restrict__ far_off, const int32_t* __restrict__ far_list,
    const int32_t* __restrict__ near_off, const int32_t* __restrict__ near_list,
    const double* __restrict__ cheb, const double* __restrict__ mom,
    const double4* __restrict__ far_sums, const double4* __restrict__ leaf_sums, int PP,
    double L, double invL, double delta2, double scale, double* __restrict__ e_tree) {
    const int4 b = blocks[blockIdx.x];
    const int leaf = b.x;
    const int i = b.y + (int)threadIdx.x;
    const bool active = i <= b.z;
    const double C = rsq_f64(__builtin_fma(0.5, 0.5, delta2));
    const double eps2 = delta2 * (L * L);
    const double Cx = rsq_f64(__builtin_fma(0.5 * L, 0.5 * L, eps2));
    const double xi = px[active ? i : b.y];
    // the block's targets lie in its leaf's x range
    const double4 tb = leaf_sums[leaves[leaf]];
    double far_acc = 0.0;
    for (int f = far_off[leaf]; f < far_off[leaf + 1]; ++f) {
        const int node = far_list[f];
        const double* tk = cheb + (size_t)node * PP;
        const double* mk = mom + (size_t)node * PP;
        const double4 fs = far_sums[node];
        const int wrap = group_wrap(tb.z, tb.w, fs.z, fs.w, invL);
        if (wrap != 2) {
            const double xs = wrap == 0 ? xi : (wrap == 1 ? xi - L : xi + L);
            for (int k = 0; k < PP; ++k) {
                const double d = xs - tk[k];
                far_acc = __builtin_fma(mk[k], d * rsq_f64(__builtin_fma(d, d, eps2)), far_acc);
            }
            far_acc = __builtin_fma(-Cx, lin_part(xs, fs.x, fs.y), far_acc);
        } else {
            for (int k = 0; k < PP; ++k)
                far_acc = __builtin_fma(mk[k], kern(xi, tk[k], invL, delta2, C), far_acc);
        }
    }
    double near_acc = 0.0;
    for (int q = near_off[leaf]; q < near_off[leaf + 1]; ++q) {
        const int node = near_list[q];
        const int2 nr = range[node];
        const double4 ls = leaf_sums[node];
        const int wrap = group_wrap(tb.z, tb.w, ls.z, ls.w, invL);
        int j = nr.x;
        if (wrap != 2) {
            const double xs = wrap == 0 ? xi : (wrap == 1 ? xi - L : xi + L);
            // eight sources per round: their scalar loads are issued
            // together, and the terms still add in source order
            for (; j + 7 <= nr.y; j += 8) {
                double xv[8], lv[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    xv[u] = px[j + u];
                    lv[u] = lam[j + u];
                }
                double kv[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const double d = xs - xv[u];
                    kv[u] = d * rsq_f64(__builtin_fma(d, d, eps2));
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) near_acc = __builtin_fma(lv[u], kv[u], near_acc);
            }
            for (; j <= nr.y; ++j) {
                const double d = xs - px[j];
                near_acc = __builtin_fma(lam[j], d * rsq_f64(__builtin_fma(d, d, eps2)), near_acc);
            }
            near_acc = __builtin_fma(-Cx, lin_part(xs, ls.x, ls.y), near_acc);
            continue;
        }
        for (; j + 7 <= nr.y; j += 8) {
            double xv[8], lv[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                xv[u] = px[j + u];
                lv[u] = lam[j + u];
            }
            double kv[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) kv[u] = kern(xi, xv[u], invL, delta2, C);
#pragma unroll
            for (int u = 0; u < 8; ++u) near_acc = __builtin_fma(lv[u], kv[u], near_acc);
        }
        for (; j <= nr.y; ++j)
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

}  // namespace

// Device treecode: d_x, d_w (input order) -> d_es (input order). The tree,
// its lists and the tree order are built on the device (fs_tree_build.hip;
// one host synchronization for the lists' sizes), then the moments, the far
// and near sums and the scatter back to input order run on `stream`. evals
// (optional) receives the far and near kernel-evaluation counts.
int treecode_f64(const double* d_x, const double* d_w, int n, double L, double eps, double mac,
                 int degree, int max_source, double* d_es, TreeWork& work, hipStream_t stream,
                 int64_t* evals) {
    if (n <= 0) return FS_OK;
    if (degree < 1 || degree + 1 > kMaxPP)
        return fail(FS_E_INVALID, "treecode: degree must be in [1, %d]", kMaxPP - 1);
    if (max_source < 1) return fail(FS_E_INVALID, "treecode: max_source must be >= 1");
    if (!(L > 0)) return fail(FS_E_INVALID, "treecode: L must be > 0");
    tree::Built T;
    FS_TRY(tree::build_device(d_x, n, L, mac, degree, max_source, work, T, stream));
    const int PP = degree + 1;
    const unsigned gb = (unsigned)((n + 255) / 256);
    gather_kernel<<<gb, 256, 0, stream>>>(d_x, T.perm, n, T.px, L, 1);
    gather_kernel<<<gb, 256, 0, stream>>>(d_w, T.perm, n, T.lam, L, 0);
    if (T.c.nnodes > 1) {
        moments_kernel<<<T.c.mom_items, 256, 0, stream>>>(T.px, T.lam, T.range, T.mom_off,
                                                           T.c.nnodes, T.cheb, PP, T.mompart);
        moments_reduce_kernel<<<(unsigned)(((int64_t)(T.c.nnodes - 1) * PP + 255) / 256), 256, 0,
                                stream>>>(T.mom_off, T.c.nnodes, PP, T.mompart, T.mom);
    }
    const double delta = eps / L, delta2 = delta * delta;
    const double scale = 0.5 * std::sqrt(1 + 4 * delta2);
    if (T.c.nnodes > 1)
        far_sums_kernel<<<(unsigned)((T.c.nnodes + 255) / 256), 256, 0, stream>>>(
            T.c.nnodes, PP, T.cheb, T.mom, T.far_sums);
    leaf_sums_kernel<<<(unsigned)T.c.blk_total, 64, 0, stream>>>(T.blocks, T.leaves, T.range,
                                                                  T.px, T.lam, T.leaf_sums);
    tree_sum_kernel<<<(unsigned)T.c.blk_total, 64, 0, stream>>>(
        T.px, T.lam, T.range, T.blocks, T.leaves, T.far_off, T.far_list, T.near_off,
        T.near_list, T.cheb, T.mom, T.far_sums, T.leaf_sums, PP, L, 1.0 / L, delta2, scale,
        T.e_tree);
    scatter_perm_kernel<<<gb, 256, 0, stream>>>(T.e_tree, T.perm, n, d_es);
    FS_HIP_CHECK(hipGetLastError());
    if (evals) {
        evals[0] = T.c.far_evals;
        evals[1] = T.c.near_evals;
    }
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
                            ctx->tree, ctx->stream, evals));
    FS_HIP_CHECK(hipMemcpyAsync(es, d_e, sizeof(double) * ns, hipMemcpyDeviceToHost, ctx->stream));
    FS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return FS_OK;
}

int fs_treecode_tree(fs_ctx* ctx, const double* sources, int64_t ns, double L, double mac,
                     int degree, int max_source, int64_t stats[4], int32_t* perm,
                     int32_t* leaf_nodes, int32_t* far_off, int32_t* far_nodes,
                     int32_t* near_off, int32_t* near_nodes) {
    if (!ctx || !sources || !stats) return fs::fail(FS_E_INVALID, "null pointer");
    if (ns <= 0 || ns > INT32_MAX) return fs::fail(FS_E_INVALID, "ns out of range");
    if (degree < 1 || degree + 1 > 16) return fs::fail(FS_E_INVALID, "degree must be in [1, 15]");
    if (max_source < 1) return fs::fail(FS_E_INVALID, "max_source must be >= 1");
    if (!(L > 0)) return fs::fail(FS_E_INVALID, "L must be > 0");
    FS_HIP_CHECK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    FS_TRY(ctx->io.ensure(fs::align_up(sizeof(double) * ns, 256), st));
    double* d_x = ctx->io.as<double>(0);
    FS_HIP_CHECK(hipMemcpyAsync(d_x, sources, sizeof(double) * ns, hipMemcpyHostToDevice, st));
    fs::tree::Built T;
    FS_TRY(fs::tree::build_device(d_x, (int)ns, L, mac, degree, max_source, ctx->tree, T, st));
    const fs::tree::Counts& C = T.c;
    stats[0] = C.nnodes;
    stats[1] = C.nleaf;
    stats[2] = C.far_total;
    stats[3] = C.near_total;
    std::vector<fs::tree::Node> nodes(C.nnodes);
    std::vector<int32_t> leaves(C.nleaf), foff(C.nleaf + 1), noff(C.nleaf + 1),
        flist(C.far_total), nlist(C.near_total), hperm(ns);
    FS_HIP_CHECK(hipMemcpyAsync(nodes.data(), T.nodes, sizeof(fs::tree::Node) * C.nnodes,
                                hipMemcpyDeviceToHost, st));
    FS_HIP_CHECK(hipMemcpyAsync(leaves.data(), T.leaves, 4 * C.nleaf, hipMemcpyDeviceToHost, st));
    FS_HIP_CHECK(hipMemcpyAsync(foff.data(), T.far_off, 4 * (C.nleaf + 1), hipMemcpyDeviceToHost, st));
    FS_HIP_CHECK(hipMemcpyAsync(noff.data(), T.near_off, 4 * (C.nleaf + 1), hipMemcpyDeviceToHost, st));
    if (C.far_total > 0)
        FS_HIP_CHECK(hipMemcpyAsync(flist.data(), T.far_list, 4 * C.far_total, hipMemcpyDeviceToHost, st));
    if (C.near_total > 0)
        FS_HIP_CHECK(hipMemcpyAsync(nlist.data(), T.near_list, 4 * C.near_total, hipMemcpyDeviceToHost, st));
    FS_HIP_CHECK(hipMemcpyAsync(hperm.data(), T.perm, 4 * ns, hipMemcpyDeviceToHost, st));
    FS_HIP_CHECK(hipStreamSynchronize(st));
    if (perm) std::copy(hperm.begin(), hperm.end(), perm);
    // leaves in increasing x (their sorted ranges are disjoint and in order)
    std::vector<int32_t> order(C.nleaf);
    for (int i = 0; i < C.nleaf; ++i) order[i] = i;
    std::sort(order.begin(), order.end(),
              [&](int32_t a, int32_t b) { return nodes[leaves[a]].s0 < nodes[leaves[b]].s0; });
    auto put = [&](int32_t* dst, int node) {
        dst[0] = nodes[node].s0;
        dst[1] = nodes[node].s1 - 1;
        dst[2] = nodes[node].depth;
    };
    int64_t fo = 0, no = 0;
    for (int i = 0; i < C.nleaf; ++i) {
        const int li = order[i];
        if (leaf_nodes) put(leaf_nodes + 3 * (int64_t)i, leaves[li]);
        if (far_off) far_off[i] = (int32_t)fo;
        if (near_off) near_off[i] = (int32_t)no;
        for (int k = foff[li]; k < foff[li + 1]; ++k, ++fo)
            if (far_nodes) put(far_nodes + 3 * fo, flist[k]);
        for (int k = noff[li]; k < noff[li + 1]; ++k, ++no)
            if (near_nodes) put(near_nodes + 3 * no, nlist[k]);
    }
    if (far_off) far_off[C.nleaf] = (int32_t)fo;
    if (near_off) near_off[C.nleaf] = (int32_t)no;
    return FS_OK;
}

}  // extern "C"
