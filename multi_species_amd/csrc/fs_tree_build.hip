// E_MQ_Treecode's tree and interaction lists built on the device (reference
// src/FieldStructure.cpp:506-702).
//
// The reference splits a node of at least max_source points at the midpoint
// of its interval, sends x <= mid to child 0 by a stable partition, drops
// empty children (:530-632) and collects each leaf's far clusters and near
// leaves by the MAC recursion (:663-702). Every partition is stable, so a
// node is a contiguous range of the points sorted stably by x, and its split
// is the first of them above the midpoint: the build needs the sorted keys
// only.
//
//   keys   order-preserving 64-bit image of fmod(x, L) (-0 as +0), stable
//          radix sort.
//   build  one workgroup, level by level: each node's cut by binary search
//          over its sorted keys; children and leaves numbered by workgroup
//          scans (a level's nodes are contiguous, in increasing x).
//   count  one thread per leaf: the MAC walk (by parent links, children in
//          order, so entries come out in the reference's order) counts far
//          and near entries and the leaf's 64-target blocks.
//   scan   one workgroup: offsets of the three per-leaf counts, evaluation
//          totals. The host reads the sizes here: the one synchronization.
//   fill   the same walk writes the lists and the blocks.
//   order  tree order (each leaf's points back in input order, as the
//          reference's stable partitions leave them): a stable sort of the
//          input indices keyed by their leaf's first sorted position (a
//          leaf's sorted range is its tree-order range).
//   nodes  member ranges and Chebyshev points of the 2nd kind (:792-823).
//
// Intervals, centres, the MAC test and the Chebyshev points are the
// reference's expressions, rounded the same way (contraction off; IEEE
// division and square root).
#pragma clang fp contract(off)

#include <algorithm>
#include <cmath>

#include "fs_field_batch.h"
#include "fs_tree.h"

namespace fs {
namespace tree {
namespace {

constexpr int kBuildThreads = 1024;
constexpr int kWaves = kBuildThreads / 64;
constexpr int kMaxDepth = 200;  // a node deeper than this that must split fails
constexpr int kBlock = 256;
constexpr int kMaxGrid = 512;
constexpr int kMaxPP = 16;

__device__ __forceinline__ uint64_t order_key(double v) {
    if (v == 0.0) v = 0.0;
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

__device__ __forceinline__ double key_value(uint64_t k) {
    const uint64_t b = (k >> 63) ? (k ^ 0x8000000000000000ull) : ~k;
    return __longlong_as_double((long long)b);
}

__global__ __launch_bounds__(kBlock) void key_kernel(const double* __restrict__ x, int n, double L,
                                                     uint64_t* __restrict__ key) {
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k < n) key[k] = order_key(fmod(x[k], L));
}

// Exclusive scan of one int per thread over a kBuildThreads workgroup.
__device__ __forceinline__ int block_scan(int v, int* s_wave, int& total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) s_wave[w] = x;
    __syncthreads();
    int base = 0, tot = 0;
    for (int k = 0; k < kWaves; ++k) {
        const int s = s_wave[k];
        base += k < w ? s : 0;
        tot += s;
    }
    __syncthreads();
    total = tot;
    return base + x - v;
}

__device__ __forceinline__ Node make_node(double x0, double x1, int s0, int s1, int parent,
                                          int depth) {
    Node c;
    c.x0 = x0;
    c.x1 = x1;
    c.xc = 0.5 * (x0 + x1);
    c.s0 = s0;
    c.s1 = s1;
    c.c0 = c.c1 = -1;
    c.parent = parent;
    c.leaf = -1;
    c.depth = depth;
    c.pad = 0;
    return c;
}

// build_tree_init (:506-527) with minval/maxval and their paddings
// (:1032-1050), then build_tree_1D_Recursive / split_2 (:530-632) level by
// level. Only the workgroup's own threads touch the nodes, so its barriers
// order every write before the next level's reads.
__global__ __launch_bounds__(kBuildThreads) void build_kernel(const uint64_t* __restrict__ skey,
                                                              int n, int max_source, int cap,
                                                              Node* __restrict__ nodes,
                                                              int32_t* __restrict__ leaves,
                                                              int32_t* __restrict__ mom_off,
                                                              Counts* __restrict__ cnt) {
    __shared__ int s_wave[kWaves];
    const int t = threadIdx.x;
    if (t == 0) {
        const double a = key_value(skey[0]) - 0.1234, b = key_value(skey[n - 1]) + 0.001;
        nodes[0] = make_node(a, b, 0, n, -1, 0);
    }
    __syncthreads();
    int lb = 0, le = 1, nl = 0, err = 0, mom_run = 0;
    for (int depth = 0; lb < le && err == 0; ++depth) {
        int next = le;
        for (int c0 = lb; c0 < le; c0 += kBuildThreads) {
            const int node = c0 + t;
            const bool mine = node < le;
            Node nd = {};
            int cut = 0, nch = 0, isleaf = 0;
            double mid = 0.0;
            if (mine) {
                nd = nodes[node];
                if (nd.s1 - nd.s0 >= max_source) {
                    mid = 0.5 * (nd.x0 + nd.x1);
                    // first sorted point above the midpoint (x <= mid stays left)
                    const uint64_t km = order_key(mid);
                    int lo = nd.s0, hi = nd.s1;
                    while (lo < hi) {
                        const int m = (lo + hi) >> 1;
                        if (skey[m] <= km) lo = m + 1;
                        else hi = m;
                    }
                    cut = lo;
                    nch = (cut > nd.s0 ? 1 : 0) + (nd.s1 > cut ? 1 : 0);
                } else {
                    isleaf = 1;
                }
            }
            int tot_ch = 0, tot_lf = 0, tot_mc = 0;
            const int ex_ch = block_scan(nch, s_wave, tot_ch);
            const int ex_lf = block_scan(isleaf, s_wave, tot_lf);
            // moment chunks of every node but the root, numbered in node order
            const int nmc = (mine && node > 0) ? (nd.s1 - nd.s0 + kMomChunk - 1) / kMomChunk : 0;
            const int ex_mc = block_scan(nmc, s_wave, tot_mc);
            if (mine) mom_off[node] = mom_run + ex_mc;
            mom_run += tot_mc;
            if (tot_ch > 0 && depth > kMaxDepth) {
                err = 1;
                break;
            }
            if (next + tot_ch > cap) {
                err = 2;
                break;
            }
            if (mine) {
                if (nch > 0) {
                    int id = next + ex_ch;
                    int c[2] = {-1, -1};
                    if (cut > nd.s0) {
                        nodes[id] = make_node(nd.x0, mid, nd.s0, cut, node, depth + 1);
                        c[0] = id++;
                    }
                    if (nd.s1 > cut) {
                        nodes[id] = make_node(mid, nd.x1, cut, nd.s1, node, depth + 1);
                        c[1] = id;
                    }
                    nodes[node].c0 = c[0];
                    nodes[node].c1 = c[1];
                }
                if (isleaf) {
                    const int li = nl + ex_lf;
                    leaves[li] = node;
                    nodes[node].leaf = li;
                }
            }
            next += tot_ch;
            nl += tot_lf;
        }
        __syncthreads();
        lb = le;
        le = next;
    }
    if (t == 0) {
        cnt->nnodes = le;
        cnt->nleaf = nl;
        cnt->err = err;
        cnt->mom_items = mom_run;
        if (err == 0) mom_off[le] = mom_run;
    }
}

// build_interaction_list (:663-702) for every leaf: the walk visits a node's
// children in order and climbs by parent links, so the entries come out in
// the reference's recursion order. kFill = false counts (far, near, blocks,
// evaluations); kFill = true writes the lists at the scanned offsets.
template <bool kFill>
__global__ __launch_bounds__(kBlock) void lists_kernel(
    const Node* __restrict__ nodes, const int32_t* __restrict__ leaves,
    const Counts* __restrict__ cnt, double L, double sqrt_theta, int PP,
    int32_t* __restrict__ far_cnt, int32_t* __restrict__ near_cnt, int32_t* __restrict__ blk_cnt,
    long long* __restrict__ fev, long long* __restrict__ nev, int32_t* __restrict__ far_list,
    int32_t* __restrict__ near_list, int4* __restrict__ blocks) {
    if (cnt->err != 0) return;
    const int nleaf = cnt->nleaf;
    for (int li = blockIdx.x * kBlock + threadIdx.x; li < nleaf; li += gridDim.x * kBlock) {
        const Node B = nodes[leaves[li]];
        const int fo = kFill ? far_cnt[li] : 0, no = kFill ? near_cnt[li] : 0;
        int nf = 0, nn = 0;
        long long npts = 0;
        int node = 0;
        for (;;) {
            const Node P = nodes[node];
            bool down = false;
            if (P.c0 < 0 && P.c1 < 0) {
                if (kFill) near_list[no + nn] = node;
                ++nn;
                npts += P.s1 - P.s0;
            } else {
                const double r_dist = fabs(B.xc - P.xc);
                const double dfbpx = fmin(r_dist, L - r_dist);
                const double dfbx = B.x1 - B.xc, dfpx = P.x1 - P.xc;
                const double sq_R = dfbpx * dfbpx, sq_rb = dfbx * dfbx, sq_rp = dfpx * dfpx;
                if ((sqrt(sq_rb) + sqrt(sq_rp)) / sqrt(sq_R) < sqrt_theta) {
                    if (kFill) far_list[fo + nf] = node;
                    ++nf;
                } else {
                    down = true;
                }
            }
            if (down) {
                node = P.c0 >= 0 ? P.c0 : P.c1;
                continue;
            }
            // the next sibling, climbing while a node is its parent's last child
            int cur = node;
            node = -1;
            while (cur != 0) {
                const int par = nodes[cur].parent;
                const int sib = nodes[par].c1;
                if (cur != sib && sib >= 0) {
                    node = sib;
                    break;
                }
                cur = par;
            }
            if (node < 0) break;
        }
        const int m = B.s1 - B.s0;
        if (!kFill) {
            far_cnt[li] = nf;
            near_cnt[li] = nn;
            blk_cnt[li] = (m + 63) / 64;
            fev[li] = (long long)m * nf * PP;
            nev[li] = (long long)m * npts;
        } else {
            const int b0 = blk_cnt[li];
            for (int b = 0; 64 * b < m; ++b)
                blocks[b0 + b] = make_int4(li, B.s0 + 64 * b, min(B.s1 - 1, B.s0 + 64 * b + 63), 0);
        }
    }
}

// Offsets of the per-leaf counts (entry nleaf = the total) and the
// evaluation totals.
__global__ __launch_bounds__(kBuildThreads) void scan_kernel(Counts* __restrict__ cnt,
                                                             int32_t* __restrict__ far_cnt,
                                                             int32_t* __restrict__ near_cnt,
                                                             int32_t* __restrict__ blk_cnt,
                                                             const long long* __restrict__ fev,
                                                             const long long* __restrict__ nev) {
    __shared__ int s_wave[kWaves];
    __shared__ long long s_ev[2][kBuildThreads];
    if (cnt->err != 0) return;
    const int nleaf = cnt->nleaf;
    const int t = threadIdx.x;
    long long bf = 0, bn = 0, bb = 0, ef = 0, en = 0;
    for (int c0 = 0; c0 < nleaf; c0 += kBuildThreads) {
        const int li = c0 + t;
        const bool mine = li < nleaf;
        const int vf = mine ? far_cnt[li] : 0, vn = mine ? near_cnt[li] : 0,
                  vb = mine ? blk_cnt[li] : 0;
        int tf = 0, tn = 0, tb = 0;
        const int xf = block_scan(vf, s_wave, tf);
        const int xn = block_scan(vn, s_wave, tn);
        const int xb = block_scan(vb, s_wave, tb);
        if (mine) {
            far_cnt[li] = (int)(bf + xf);
            near_cnt[li] = (int)(bn + xn);
            blk_cnt[li] = (int)(bb + xb);
            ef += fev[li];
            en += nev[li];
        }
        bf += tf;
        bn += tn;
        bb += tb;
    }
    s_ev[0][t] = ef;
    s_ev[1][t] = en;
    __syncthreads();
    for (int s = kBuildThreads / 2; s > 0; s >>= 1) {
        if (t < s) {
            s_ev[0][t] += s_ev[0][t + s];
            s_ev[1][t] += s_ev[1][t + s];
        }
        __syncthreads();
    }
    if (t == 0) {
        far_cnt[nleaf] = (int)bf;
        near_cnt[nleaf] = (int)bn;
        blk_cnt[nleaf] = (int)bb;
        cnt->far_total = bf;
        cnt->near_total = bn;
        cnt->blk_total = bb;
        cnt->far_evals = s_ev[0][0];
        cnt->near_evals = s_ev[1][0];
    }
}

// Sort keys of the tree order: every point keyed by its leaf's first sorted
// position, in the top bits.
__global__ __launch_bounds__(kBlock) void order_key_kernel(const Node* __restrict__ nodes,
                                                           const int32_t* __restrict__ leaves,
                                                           const Counts* __restrict__ cnt,
                                                           const int32_t* __restrict__ ord,
                                                           int shift, uint64_t* __restrict__ pkey) {
    const int nleaf = cnt->nleaf;
    for (int li = blockIdx.x; li < nleaf; li += gridDim.x) {
        const Node B = nodes[leaves[li]];
        for (int k = B.s0 + (int)threadIdx.x; k < B.s1; k += kBlock)
            pkey[ord[k]] = (uint64_t)B.s0 << shift;
    }
}

struct ChebT {
    double t[kMaxPP];
};

// Member ranges of every node and the Chebyshev points of the 2nd kind of
// every node but the root (alloc_set_cluster_list, :792-823).
__global__ __launch_bounds__(kBlock) void node_kernel(const Node* __restrict__ nodes,
                                                      const Counts* __restrict__ cnt, int PP,
                                                      ChebT tk, int2* __restrict__ range,
                                                      double* __restrict__ cheb) {
    const int nn = cnt->nnodes;
    for (int k = blockIdx.x * kBlock + threadIdx.x; k < nn; k += gridDim.x * kBlock) {
        const Node N = nodes[k];
        range[k] = make_int2(N.s0, N.s1 - 1);
        for (int j = 0; j < PP; ++j)
            cheb[(size_t)k * PP + j] = k > 0 ? N.x0 + (tk.t[j] + 1.0) * 0.5 * (N.x1 - N.x0) : 0.0;
    }
}

int grid_for(long long items) {
    return (int)std::max<long long>(1, std::min<long long>(kMaxGrid, (items + kBlock - 1) / kBlock));
}

}  // namespace

int build_device(const double* d_x, int n, double L, double mac, int degree, int max_source,
                 TreeWork& work, Built& out, hipStream_t stream) {
    if (n <= 0) return fail(FS_E_INVALID, "treecode: no points");
    const int PP = degree + 1;
    if (PP > kMaxPP) return fail(FS_E_INVALID, "treecode: degree must be in [1, %d]", kMaxPP - 1);
    // nodes that split: at most n / max_source per depth (disjoint ranges of
    // max_source points or more), depths 0..kMaxDepth, two children each
    const long long bound = 1 + 2LL * (kMaxDepth + 1) * (n / max_source);
    long long cap = std::max<long long>(work.node_cap, 64 + 8LL * (n / max_source));
    cap = std::min(cap, bound);
    if (cap > INT32_MAX / 2) return fail(FS_E_RANGE, "treecode: too many tree nodes");
    auto al = [](size_t b) { return align_up(b, 256); };
    const size_t b_k = al(8 * (size_t)n), b_i = al(4 * (size_t)n), b_sort = radix_sort_bytes(n);
    Counts* d_cnt = nullptr;
    uint64_t *key = nullptr, *skey = nullptr;
    int32_t *ord = nullptr, *perm = nullptr, *leaves = nullptr, *far_cnt = nullptr,
            *near_cnt = nullptr, *blk_cnt = nullptr, *mom_off = nullptr;
    long long *fev = nullptr, *nev = nullptr;
    Node* nodes = nullptr;
    int2* range = nullptr;
    double* cheb = nullptr;
    void* rtmp = nullptr;
    for (;;) {
        const size_t c = (size_t)cap;
        const size_t total = al(sizeof(Counts)) + 2 * b_k + 2 * b_i + b_sort + al(sizeof(Node) * c) +
                             al(4 * c) + 4 * al(4 * (c + 1)) + 2 * al(8 * c) + al(8 * c) +
                             al(8 * c * PP);
        FS_TRY(work.base.ensure(total, stream));
        char* p = work.base.as<char>();
        d_cnt = reinterpret_cast<Counts*>(p); p += al(sizeof(Counts));
        key = reinterpret_cast<uint64_t*>(p); p += b_k;
        skey = reinterpret_cast<uint64_t*>(p); p += b_k;
        ord = reinterpret_cast<int32_t*>(p); p += b_i;
        perm = reinterpret_cast<int32_t*>(p); p += b_i;
        rtmp = p; p += b_sort;
        nodes = reinterpret_cast<Node*>(p); p += al(sizeof(Node) * c);
        leaves = reinterpret_cast<int32_t*>(p); p += al(4 * c);
        far_cnt = reinterpret_cast<int32_t*>(p); p += al(4 * (c + 1));
        near_cnt = reinterpret_cast<int32_t*>(p); p += al(4 * (c + 1));
        blk_cnt = reinterpret_cast<int32_t*>(p); p += al(4 * (c + 1));
        mom_off = reinterpret_cast<int32_t*>(p); p += al(4 * (c + 1));
        fev = reinterpret_cast<long long*>(p); p += al(8 * c);
        nev = reinterpret_cast<long long*>(p); p += al(8 * c);
        range = reinterpret_cast<int2*>(p); p += al(8 * c);
        cheb = reinterpret_cast<double*>(p);

        key_kernel<<<(unsigned)((n + kBlock - 1) / kBlock), kBlock, 0, stream>>>(d_x, n, L, key);
        FS_HIP_CHECK(hipGetLastError());
        FS_TRY(radix_sort_full(key, n, skey, ord, rtmp, stream));
        build_kernel<<<1, kBuildThreads, 0, stream>>>(skey, n, max_source, (int)cap, nodes, leaves,
                                                      mom_off, d_cnt);
        FS_HIP_CHECK(hipGetLastError());
        lists_kernel<false><<<grid_for(cap), kBlock, 0, stream>>>(
            nodes, leaves, d_cnt, L, std::sqrt(mac * mac), PP, far_cnt, near_cnt, blk_cnt, fev,
            nev, nullptr, nullptr, nullptr);
        FS_HIP_CHECK(hipGetLastError());
        scan_kernel<<<1, kBuildThreads, 0, stream>>>(d_cnt, far_cnt, near_cnt, blk_cnt, fev, nev);
        FS_HIP_CHECK(hipGetLastError());
        FS_HIP_CHECK(hipMemcpyAsync(&out.c, d_cnt, sizeof(Counts), hipMemcpyDeviceToHost, stream));
        FS_HIP_CHECK(hipStreamSynchronize(stream));
        if (out.c.err == 2 && cap < bound) {
            cap = std::min(bound, cap * 4);
            continue;
        }
        if (out.c.err == 1) return fail(FS_E_RANGE, "treecode: more than max_source coincident points");
        if (out.c.err != 0) return fail(FS_E_INTERNAL, "treecode: node capacity %lld exceeded", cap);
        break;
    }
    work.node_cap = (int)cap;
    const Counts& C = out.c;
    if (C.far_total < 0 || C.near_total < 0 || C.blk_total < 0 || C.far_total > INT32_MAX ||
        C.near_total > INT32_MAX || C.blk_total > INT32_MAX)
        return fail(FS_E_RANGE, "treecode: interaction lists too long");
    // the lists and the evaluation's work arrays
    const size_t b_far = al(4 * std::max<long long>(1, C.far_total)),
                 b_near = al(4 * std::max<long long>(1, C.near_total)),
                 b_blk = al(sizeof(int4) * std::max<long long>(1, C.blk_total)),
                 b_mom = al(8 * (size_t)C.nnodes * PP), b_x = al(8 * (size_t)n),
                 b_mp = al(8 * (size_t)std::max(C.mom_items, 1) * PP),
                 b_ns = al(sizeof(double4) * (size_t)std::max(C.nnodes, 1));
    FS_TRY(work.lists.ensure(b_far + b_near + b_blk + b_mom + 3 * b_x + b_mp + 2 * b_ns, stream));
    char* q = work.lists.as<char>();
    int32_t* far_list = reinterpret_cast<int32_t*>(q); q += b_far;
    int32_t* near_list = reinterpret_cast<int32_t*>(q); q += b_near;
    int4* blocks = reinterpret_cast<int4*>(q); q += b_blk;
    out.mom = reinterpret_cast<double*>(q); q += b_mom;
    out.px = reinterpret_cast<double*>(q); q += b_x;
    out.lam = reinterpret_cast<double*>(q); q += b_x;
    out.e_tree = reinterpret_cast<double*>(q); q += b_x;
    out.mompart = reinterpret_cast<double*>(q); q += b_mp;
    out.far_sums = reinterpret_cast<double4*>(q); q += b_ns;
    out.leaf_sums = reinterpret_cast<double4*>(q);

    lists_kernel<true><<<grid_for(C.nleaf), kBlock, 0, stream>>>(
        nodes, leaves, d_cnt, L, std::sqrt(mac * mac), PP, far_cnt, near_cnt, blk_cnt, fev, nev,
        far_list, near_list, blocks);
    FS_HIP_CHECK(hipGetLastError());
    // tree order: a stable sort of the input indices by their leaf's first
    // sorted position (bytes enough for n - 1)
    int bytes = 1;
    while (bytes < 4 && ((long long)(n - 1) >> (8 * bytes)) != 0) ++bytes;
    order_key_kernel<<<(unsigned)std::max(1, std::min(C.nleaf, 4 * kMaxGrid)), kBlock, 0, stream>>>(
        nodes, leaves, d_cnt, ord, 64 - 8 * bytes, key);
    FS_HIP_CHECK(hipGetLastError());
    FS_TRY(radix_sort_top(key, n, bytes, skey, perm, rtmp, stream));
    ChebT tk{};
    const double pi = 3.14159265358979323846;
    const double h = pi / degree;
    for (int i = 0; i < PP; ++i) tk.t[i] = std::cos(i * h);
    node_kernel<<<grid_for(C.nnodes), kBlock, 0, stream>>>(nodes, d_cnt, PP, tk, range, cheb);
    FS_HIP_CHECK(hipGetLastError());

    out.nodes = nodes;
    out.leaves = leaves;
    out.perm = perm;
    out.range = range;
    out.cheb = cheb;
    out.far_off = far_cnt;
    out.far_list = far_list;
    out.near_off = near_cnt;
    out.near_list = near_list;
    out.blocks = blocks;
    out.mom_off = mom_off;
    return FS_OK;
}

}  // namespace tree
}  // namespace fs
