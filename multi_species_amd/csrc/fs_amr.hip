// Adaptive refinement of panel trees on the device (fs_amr.h).
//
// refine_panels (reference src/structure_files/AMRStructure_amr.cpp:506-839)
// splits the flagged panels of a pass one after another in index order, and
// what a panel does depends on its neighbours' state at that moment. Each of
// those effects has a closed form in the state at the start of the pass, so
// a pass runs in parallel and still numbers panels and points as the
// sequential one does:
//
//  * The set R of panels split in the pass: the flagged panels, plus every
//    panel q that a member a < q flags on its way (a side of a with no
//    neighbour at a's level: q is a's parent's neighbour there, not yet
//    split, so the pass reaches and splits it later). The chains are
//    followed from every flagged panel; each step goes one level up.
//  * When the pass reaches member jj, a neighbour nb counts as split iff nb
//    was split before the pass or nb is in R with nb < jj. A side whose
//    neighbour is split takes that neighbour's children's edge points
//    (except across the periodic seam); every other side makes two fresh
//    points. A member adds 4 panels and 8 + 2 (fresh sides) points, so one
//    exclusive scan over R gives every member its panel and point numbers.
//  * A side without a neighbour at jj's level (-1) may have been linked, by
//    then, to a child of the coarser neighbour G (G in R, G < jj; G links
//    the children of the panel it faces). The side is then fresh and flags
//    nothing; otherwise it is fresh and flags the parent's neighbour there
//    (itself possibly such a child of a panel split earlier in the pass).
//    A flag on a panel of R is cleared when the pass reaches it, so only
//    flags on panels outside R remain (the next pass splits them).
//  * Neighbour links of same-level neighbours are symmetric; the host
//    links a member's children to nb's children when nb was split before
//    it, and nb, split after it in the same pass, links them back: either
//    way the children end linked to nb's children whenever nb is split by
//    the end of the pass. Children of panels split before the pass are
//    linked to the new children (the host's rewiring).
//
// The build launch writes the new records and fresh points from the
// pass-start state; the link launch then fills the points shared with
// children made in the same pass, links older children, writes the members'
// own flags and sets the flags for the next pass, so no thread reads what
// another writes. f at the new points comes from the grid-search module's
// point kernel on the old mesh (interpolate_from_mesh), and test_panel
// (:971-1081) runs on the new panels only (an older leaf's test cannot
// change: its points, f and level are the same).
//
// set_leaves_weights (:1084-1151) adds each leaf's weights to its nine
// points in depth-first leaf order; here every leaf finds its depth-first
// rank from subtree leaf counts, scatters its nine terms into per-point
// slots, and each point adds its terms in rank order. The reduced-x map
// (AMRStructure.cpp:147-181) comes from a stable sort of the x keys.
#pragma clang fp contract(off)

#include "fs_amr.h"

#include <algorithm>
#include <cstring>

#include "fs_amr_pass.h"
#include "fs_field_batch.h"

namespace fs {
namespace amr {
namespace {

using namespace pass;

constexpr int kB = 256;
constexpr int kTile = kB * 8;  // scan items per workgroup
constexpr int kMaxPasses = 256;

struct DevOps {
    __device__ void or_(int* a, int v) const { atomicOr(a, v); }
    __device__ int add(int* a, int v) const { return atomicAdd(a, v); }
};

__device__ __forceinline__ int tid() { return (int)(blockIdx.x * kB + threadIdx.x); }

// Tests the panels made by the last pass [P0, P) and counts the flagged
// panels of [mu, P).
__global__ __launch_bounds__(kB) void test_kernel(const Job* __restrict__ jobs) {
    const Job& J = jobs[blockIdx.y];
    const int flagged = test_one(J, J.mu + tid(), DevOps{});
    __shared__ int red;
    if (threadIdx.x == 0) red = 0;
    __syncthreads();
    if (flagged) atomicAdd(&red, 1);
    __syncthreads();
    if (threadIdx.x == 0 && red) atomicAdd(&J.ctr->flagged, red);
}

__global__ __launch_bounds__(kB) void chain_clear_kernel(const Job* __restrict__ jobs) {
    const Job& J = jobs[blockIdx.y];
    chain_clear_one(J, J.mu + tid());
}

__global__ __launch_bounds__(kB) void chain_kernel(const Job* __restrict__ jobs) {
    const Job& J = jobs[blockIdx.y];
    chain_one(J, J.mu + tid(), DevOps{});
}

__global__ __launch_bounds__(kB) void count_kernel(const Job* __restrict__ jobs) {
    const Job& J = jobs[blockIdx.y];
    count_one(J, J.mu + tid());
}

// ---- exclusive scans of int2 items, per job ---------------------------------
// mode 0: the pass' (member, points) over panels [mu, P) -> nR, newV;
// mode 1: run heads over points [0, V) -> U.
__device__ __forceinline__ int scan_n(const Job& J, int mode) {
    return mode == 0 ? J.ctr->P - J.mu : J.ctr->V;
}

__device__ __forceinline__ int2 block_inclusive(int2 v, int2* sh) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int o = 1; o < kB; o <<= 1) {
        const int2 a = t >= o ? sh[t - o] : make_int2(0, 0);
        __syncthreads();
        sh[t] = add2(sh[t], a);
        __syncthreads();
    }
    const int2 r = sh[t];
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(kB) void scan_tile_kernel(const Job* __restrict__ jobs, int mode) {
    const Job& J = jobs[blockIdx.y];
    const int n = scan_n(J, mode);
    const int b0 = (int)blockIdx.x * kTile;
    if (b0 >= n) return;
    const int2* in = mode == 0 ? J.cnt : J.vcnt;
    int2* out = mode == 0 ? J.off : J.voff;
    __shared__ int2 sh[kB];
    int2 v[8];
    int2 s = make_int2(0, 0);
    const int i0 = b0 + (int)threadIdx.x * 8;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        v[k] = (i0 + k < n) ? in[i0 + k] : make_int2(0, 0);
        s = add2(s, v[k]);
    }
    const int2 incl = block_inclusive(s, sh);
    int2 ex = sub2(incl, s);
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        if (i0 + k < n) out[i0 + k] = ex;
        ex = add2(ex, v[k]);
    }
    if (threadIdx.x == kB - 1) J.bsum[blockIdx.x] = incl;
}

__global__ __launch_bounds__(kB) void scan_blocks_kernel(const Job* __restrict__ jobs, int mode) {
    const Job& J = jobs[blockIdx.y];
    const int n = scan_n(J, mode);
    const int nb = (n + kTile - 1) / kTile;
    __shared__ int2 sh[kB];
    int2 run = make_int2(0, 0);
    for (int c = 0; c < nb; c += kB) {
        const int i = c + (int)threadIdx.x;
        const int2 v = i < nb ? J.bsum[i] : make_int2(0, 0);
        const int2 incl = block_inclusive(v, sh);
        if (i < nb) J.bsum[i] = add2(run, sub2(incl, v));
        // every thread reads the chunk total before the next chunk's scan
        sh[threadIdx.x] = incl;
        __syncthreads();
        run = add2(run, sh[kB - 1]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (mode == 0) {
            J.ctr->nR = run.x;
            J.ctr->newV = run.y;
        } else {
            J.ctr->U = run.x;
        }
    }
}

__global__ __launch_bounds__(kB) void scan_add_kernel(const Job* __restrict__ jobs, int mode) {
    const Job& J = jobs[blockIdx.y];
    const int n = scan_n(J, mode);
    const int b0 = (int)blockIdx.x * kTile;
    if (b0 >= n || blockIdx.x == 0) return;
    int2* out = mode == 0 ? J.off : J.voff;
    const int2 a = J.bsum[blockIdx.x];
    for (int k = (int)threadIdx.x; k < kTile; k += kB) {
        const int i = b0 + k;
        if (i < n) out[i] = add2(out[i], a);
    }
}

// ---- one pass, weights and the map ----------------------------------------

__global__ __launch_bounds__(kB) void build_kernel(const Job* __restrict__ jobs) {
    const Job& J = jobs[blockIdx.y];
    build_one(J, J.mu + tid(), DevOps{});
}

__global__ __launch_bounds__(kB) void link_kernel(const Job* __restrict__ jobs) {
    const Job& J = jobs[blockIdx.y];
    link_one(J, J.mu + tid(), DevOps{});
}

__global__ void commit_kernel(const Job* __restrict__ jobs, int njobs) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < njobs) commit_one(jobs[j]);
}

__global__ __launch_bounds__(kB) void leaf_init_kernel(const Job* __restrict__ jobs) {
    leaf_init_one(jobs[blockIdx.y], tid());
}

__global__ __launch_bounds__(kB) void leaf_terms_kernel(const Job* __restrict__ jobs) {
    leaf_terms_one(jobs[blockIdx.y], tid(), DevOps{});
}

__global__ __launch_bounds__(kB) void weight_kernel(const Job* __restrict__ jobs) {
    weight_one(jobs[blockIdx.y], tid());
}

__global__ __launch_bounds__(kB) void key_kernel(const Job* __restrict__ jobs) {
    key_one(jobs[blockIdx.y], tid());
}

__global__ __launch_bounds__(kB) void head_kernel(const Job* __restrict__ jobs) {
    head_one(jobs[blockIdx.y], tid());
}

__global__ __launch_bounds__(kB) void map_kernel(const Job* __restrict__ jobs) {
    map_one(jobs[blockIdx.y], tid());
}

// The base tree and grid points of every species copied into its work
// buffers (one launch for all of them).
struct BaseCopy {
    const int32_t* rec;
    const double* x;
    const double* p;
    const double* f;
    int32_t* rec_to;
    double* x_to;
    double* p_to;
    double* f_to;
    int nrec, nb;  // int32 words of the records, points
};

__global__ __launch_bounds__(kB) void base_copy_kernel(const BaseCopy* __restrict__ bc) {
    const BaseCopy& B = bc[blockIdx.y];
    const int i = tid();
    if (i < B.nrec) B.rec_to[i] = B.rec[i];
    if (i < B.nb) {
        B.x_to[i] = B.x[i];
        B.p_to[i] = B.p[i];
        B.f_to[i] = B.f[i];
    }
}

// The finishing sort of all species at once: species k's keys sit at
// base[k] of the concatenated arrays, its segment number k beside them, and
// its sorted slice comes back at the same place (the segment is the most
// significant digit); the permutation is then made species-local.
__global__ __launch_bounds__(kB) void seg_fill_kernel(const Job* __restrict__ jobs,
                                                      const int* __restrict__ base,
                                                      int32_t* __restrict__ seg_of) {
    const int i = tid();
    if (i < jobs[blockIdx.y].ctr->V) seg_of[base[blockIdx.y] + i] = (int32_t)blockIdx.y;
}

__global__ __launch_bounds__(kB) void perm_local_kernel(const Job* __restrict__ jobs,
                                                        const int* __restrict__ base) {
    const Job& J = jobs[blockIdx.y];
    const int i = tid();
    if (i < J.ctr->V) J.perm[i] -= base[blockIdx.y];
}

// ---- host side --------------------------------------------------------------

unsigned cdiv(long a, long b) { return (unsigned)((a + b - 1) / b); }

// Grows a buffer of w to `need` bytes keeping its first `used` bytes (the
// old block is let go after the copy is enqueued: a pooled block is reused
// only by later work on the same stream, a freed one after the device's).
int grow_keep(Work& w, DevBuf& b, size_t used, size_t need, hipStream_t st) {
    if (need <= b.cap) return FS_OK;
    DevBuf nb;
    FS_TRY(w.ensure(nb, std::max(need, b.cap + b.cap / 2), st));
    if (used > 0) FS_HIP_CHECK(hipMemcpyAsync(nb.ptr, b.ptr, used, hipMemcpyDeviceToDevice, st));
    if (!w.pool) FS_HIP_CHECK(hipStreamSynchronize(st));
    w.put(b);
    b = nb;
    nb.ptr = nullptr;
    nb.cap = 0;
    return FS_OK;
}

// Tree and pass-state capacity for p panels and v points (contents kept).
int ensure_tree(Work& w, int used_p, int used_v, int p, int v, hipStream_t st) {
    FS_TRY(grow_keep(w, w.rec, sizeof(int32_t) * kI * (size_t)used_p,
                     sizeof(int32_t) * kI * (size_t)p, st));
    FS_TRY(grow_keep(w, w.chain, sizeof(int32_t) * (size_t)used_p, sizeof(int32_t) * (size_t)p, st));
    FS_TRY(grow_keep(w, w.cnt, sizeof(int2) * (size_t)used_p, sizeof(int2) * (size_t)p, st));
    FS_TRY(grow_keep(w, w.off, sizeof(int2) * (size_t)used_p, sizeof(int2) * (size_t)p, st));
    for (DevBuf* b : {&w.x, &w.p, &w.f})
        FS_TRY(grow_keep(w, *b, sizeof(double) * (size_t)used_v, sizeof(double) * (size_t)v, st));
    w.cap_p = (int)std::min<size_t>(w.rec.cap / (sizeof(int32_t) * kI),
                                    std::min(w.chain.cap / sizeof(int32_t), w.cnt.cap / sizeof(int2)));
    w.cap_p = (int)std::min<size_t>(w.cap_p, w.off.cap / sizeof(int2));
    w.cap_v = (int)(std::min(w.x.cap, std::min(w.p.cap, w.f.cap)) / sizeof(double));
    return FS_OK;
}

void set_views(Job& j, Work& w) {
    j.rec = w.rec.as<int32_t>();
    j.x = w.x.as<double>();
    j.p = w.p.as<double>();
    j.f = w.f.as<double>();
    j.cap_p = w.cap_p;
    j.cap_v = w.cap_v;
    j.chain = w.chain.as<int32_t>();
    j.cnt = w.cnt.as<int2>();
    j.off = w.off.as<int2>();
}

// Pass temporaries for p panels (plans of up to p members, scan block sums
// of p or v items).
int ensure_tmp(Work& w, Job& j, int p, int v, hipStream_t st) {
    const size_t bp = align_up(sizeof(int32_t) * kPlanInts * (size_t)p, 256);
    const size_t bs = align_up(sizeof(int2) * (cdiv(std::max(p, v), kTile) + 1), 256);
    FS_TRY(w.ensure(w.tmp, bp + bs, st));
    j.plan = w.tmp.as<int32_t>(0);
    j.bsum = w.tmp.as<int2>(bp);
    return FS_OK;
}

// Finishing work space and outputs for v points (the keys, sorted keys and
// permutation are slices of the shared sort buffer).
int ensure_finish(Work& w, Job& j, int v, hipStream_t st) {
    const size_t n = (size_t)std::max(v, 1);
    const size_t b4 = align_up(4 * n, 256), b8 = align_up(8 * n, 256);
    const size_t bsl8 = align_up(8 * n * kSlots, 256);
    const size_t b2 = align_up(sizeof(int2) * n, 256);
    FS_TRY(w.ensure(w.fin, b4 + 2 * bsl8 + 2 * b2, st));
    char* c = w.fin.as<char>();
    j.nslot = reinterpret_cast<int*>(c);
    c += b4;
    j.spath = reinterpret_cast<uint64_t*>(c);
    c += bsl8;
    j.sval = reinterpret_cast<double*>(c);
    c += bsl8;
    j.vcnt = reinterpret_cast<int2*>(c);
    c += b2;
    j.voff = reinterpret_cast<int2*>(c);
    const size_t bo = align_up(4 * (n + 1), 256);
    FS_TRY(w.ensure(w.out, 2 * b8 + bo + 2 * b4, st));
    c = w.out.as<char>();
    j.bw = reinterpret_cast<double*>(c);
    c += b8;
    j.ux = reinterpret_cast<double*>(c);
    c += b8;
    j.roff = reinterpret_cast<int32_t*>(c);
    c += bo;
    j.members = reinterpret_cast<int32_t*>(c);
    c += b4;
    j.uof = reinterpret_cast<int32_t*>(c);
    return FS_OK;
}

int status_error(int st) {
    if (st & kStatTopology)
        return fail(FS_E_INTERNAL,
                    "adaptive refinement: a neighbour link the device passes do not follow");
    if (st & kStatCapacity) return fail(FS_E_INTERNAL, "adaptive refinement: pass buffers too small");
    if (st & kStatSlots)
        return fail(FS_E_INTERNAL, "adaptive refinement: a point on more than %d leaves", kSlots);
    if (st & kStatDepth) return fail(FS_E_INTERNAL, "adaptive refinement: tree deeper than expected");
    return FS_OK;
}

}  // namespace

int refine(std::vector<Input>& in, std::vector<Output>& out, DevBuf& jobs_buf,
           DevBuf& sort_buf, hipStream_t st) {
    const int nj = (int)in.size();
    out.assign(nj, Output{});
    if (nj == 0) return FS_OK;
    if (nj > 65536) return fail(FS_E_INVALID, "adaptive refinement: more than 65536 species");
    std::vector<Job> hj(nj);
    std::vector<Counters> hc(nj);
    std::vector<remesh::Job> hp(nj);
    std::vector<BaseCopy> hb(nj);
    // job tables, the counters of every species (one copy each way), the
    // base copies
    const size_t bjobs = align_up(sizeof(Job) * nj, 256);
    const size_t bpj = align_up(sizeof(remesh::Job) * nj, 256);
    const size_t bctr = align_up(sizeof(Counters) * nj, 256);
    FS_TRY(jobs_buf.ensure(bjobs + bpj + bctr + sizeof(BaseCopy) * nj, st));
    Job* dj = jobs_buf.as<Job>(0);
    remesh::Job* dpj = jobs_buf.as<remesh::Job>(bjobs);
    Counters* dctr = jobs_buf.as<Counters>(bjobs + bpj);
    BaseCopy* dbc = jobs_buf.as<BaseCopy>(bjobs + bpj + bctr);

    // the base tree with the grid interpolation's f
    int max_base = 0;
    for (int k = 0; k < nj; ++k) {
        Input& I = in[k];
        Work& w = *I.w;
        if (I.eps.size() > 5 || I.mu < 1 || I.mu > I.base_panels)
            return fail(FS_E_INVALID, "adaptive refinement: bad species input");
        FS_TRY(ensure_tree(w, 0, 0, I.base_panels, I.nb, st));
        hb[k] = BaseCopy{I.base_rec, I.gx, I.gp, I.gf, w.rec.as<int32_t>(), w.x.as<double>(),
                         w.p.as<double>(), w.f.as<double>(), kI * I.base_panels, I.nb};
        max_base = std::max(max_base, std::max(kI * I.base_panels, I.nb));
        hc[k] = Counters{I.base_panels, I.nb, 0, 0, 0, 0, I.mu, 0};
        Job& j = hj[k];
        j = Job{};
        set_views(j, w);
        j.ctr = dctr + k;
        j.mu = I.mu;
        j.max_height = I.max_height;
        j.periodic = I.periodic ? 1 : 0;
        j.n_eps = (int)I.eps.size();
        for (size_t e = 0; e < I.eps.size(); ++e) j.eps[e] = I.eps[e];
        j.q = I.q;
        j.simpson = I.simpson ? 1 : 0;
        j.pj = dpj + k;
        FS_TRY(ensure_tmp(w, j, I.base_panels, I.nb, st));
        hp[k] = I.old;
        hp[k].on = nullptr;
    }
    FS_HIP_CHECK(hipMemcpyAsync(dbc, hb.data(), sizeof(BaseCopy) * nj, hipMemcpyHostToDevice, st));
    base_copy_kernel<<<dim3(cdiv(std::max(max_base, 1), kB), (unsigned)nj), kB, 0, st>>>(dbc);
    FS_HIP_CHECK(hipGetLastError());
    FS_HIP_CHECK(hipMemcpyAsync(dctr, hc.data(), sizeof(Counters) * nj, hipMemcpyHostToDevice, st));
    FS_HIP_CHECK(hipMemcpyAsync(dpj, hp.data(), sizeof(remesh::Job) * nj, hipMemcpyHostToDevice, st));

    auto upload = [&]() -> int {
        FS_HIP_CHECK(hipMemcpyAsync(dj, hj.data(), sizeof(Job) * nj, hipMemcpyHostToDevice, st));
        return FS_OK;
    };
    auto readback = [&]() -> int {
        FS_HIP_CHECK(hipMemcpyAsync(hc.data(), dctr, sizeof(Counters) * nj, hipMemcpyDeviceToHost, st));
        FS_HIP_CHECK(hipStreamSynchronize(st));
        int all = 0;
        for (int k = 0; k < nj; ++k) all |= hc[k].status;
        return status_error(all);
    };
    // test [P0, P), then R, counts and numbers of the next pass over [mu, P)
    auto test_and_plan = [&](int span) -> int {
        const dim3 g(cdiv(std::max(span, 1), kB), (unsigned)nj);
        test_kernel<<<g, kB, 0, st>>>(dj);
        chain_clear_kernel<<<g, kB, 0, st>>>(dj);
        chain_kernel<<<g, kB, 0, st>>>(dj);
        count_kernel<<<g, kB, 0, st>>>(dj);
        const dim3 gs(cdiv(std::max(span, 1), kTile), (unsigned)nj);
        scan_tile_kernel<<<gs, kB, 0, st>>>(dj, 0);
        scan_blocks_kernel<<<dim3(1, (unsigned)nj), kB, 0, st>>>(dj, 0);
        scan_add_kernel<<<gs, kB, 0, st>>>(dj, 0);
        FS_HIP_CHECK(hipGetLastError());
        return FS_OK;
    };

    FS_TRY(upload());
    int span = 0;
    for (int k = 0; k < nj; ++k) span = std::max(span, hc[k].P - in[k].mu);
    FS_TRY(test_and_plan(span));
    FS_TRY(readback());

    for (int pass = 0;; ++pass) {
        int any = 0;
        for (int k = 0; k < nj; ++k) any |= hc[k].nR;
        if (!any) break;
        if (pass >= kMaxPasses) return fail(FS_E_INTERNAL, "adaptive refinement: too many passes");
        int span_old = 0, span_new = 0, max_new_v = 0;
        for (int k = 0; k < nj; ++k) {
            Work& w = *in[k].w;
            const Counters& C = hc[k];
            const int np_ = C.P + 4 * C.nR, nv_ = C.V + C.newV;
            if ((int64_t)C.P + 4 * (int64_t)C.nR > INT32_MAX / kI ||
                (int64_t)C.V + C.newV > INT32_MAX / 8)
                return fail(FS_E_RANGE, "adaptive refinement: tree too large");
            FS_TRY(ensure_tree(w, C.P, C.V, np_, nv_, st));
            set_views(hj[k], w);
            FS_TRY(ensure_tmp(w, hj[k], np_, nv_, st));
            span_old = std::max(span_old, C.P - in[k].mu);
            span_new = std::max(span_new, np_ - in[k].mu);
            max_new_v = std::max(max_new_v, C.newV);
        }
        FS_TRY(upload());
        const dim3 g(cdiv(std::max(span_old, 1), kB), (unsigned)nj);
        build_kernel<<<g, kB, 0, st>>>(dj);
        link_kernel<<<g, kB, 0, st>>>(dj);
        commit_kernel<<<cdiv(nj, 64), 64, 0, st>>>(dj, nj);
        FS_HIP_CHECK(hipGetLastError());
        remesh::Launch l;
        l.njobs = nj;
        l.max_points = max_new_v;
        FS_TRY(remesh::interp_points(dpj, l, st));
        FS_TRY(test_and_plan(span_new));
        FS_TRY(readback());
    }

    // weights in depth-first leaf order, and the reduced-x map
    int maxp = 0, maxv = 0;
    std::vector<int> base(nj);
    int64_t total_v = 0;
    for (int k = 0; k < nj; ++k) {
        Work& w = *in[k].w;
        FS_TRY(ensure_tmp(w, hj[k], hc[k].P, hc[k].V, st));
        FS_TRY(ensure_finish(w, hj[k], hc[k].V, st));
        maxp = std::max(maxp, hc[k].P);
        maxv = std::max(maxv, hc[k].V);
        base[k] = (int)total_v;
        total_v += hc[k].V;
    }
    if (total_v > INT32_MAX / 8) return fail(FS_E_RANGE, "adaptive refinement: too many points");
    const int nv = (int)std::max<int64_t>(total_v, 1);
    const size_t s8 = align_up(8 * (size_t)nv, 256), s4 = align_up(4 * (size_t)nv, 256);
    const size_t sb = align_up(sizeof(int) * nj, 256);
    FS_TRY(sort_buf.ensure(2 * s8 + 2 * s4 + sb + radix_sort_bytes(nv), st));
    char* sc = sort_buf.as<char>();
    uint64_t* cat_key = reinterpret_cast<uint64_t*>(sc);
    uint64_t* cat_skey = reinterpret_cast<uint64_t*>(sc + s8);
    int32_t* cat_perm = reinterpret_cast<int32_t*>(sc + 2 * s8);
    int32_t* seg_of = reinterpret_cast<int32_t*>(sc + 2 * s8 + s4);
    int* dbase = reinterpret_cast<int*>(sc + 2 * s8 + 2 * s4);
    void* sort_tmp = sc + 2 * s8 + 2 * s4 + sb;
    for (int k = 0; k < nj; ++k) {
        hj[k].key = cat_key + base[k];
        hj[k].skey = cat_skey + base[k];
        hj[k].perm = cat_perm + base[k];
    }
    FS_HIP_CHECK(hipMemcpyAsync(dbase, base.data(), sizeof(int) * nj, hipMemcpyHostToDevice, st));
    FS_TRY(upload());
    const dim3 gp(cdiv(std::max(maxp, 1), kB), (unsigned)nj);
    const dim3 gv(cdiv(std::max(maxv, 1), kB), (unsigned)nj);
    leaf_init_kernel<<<gv, kB, 0, st>>>(dj);
    leaf_terms_kernel<<<gp, kB, 0, st>>>(dj);
    weight_kernel<<<gv, kB, 0, st>>>(dj);
    key_kernel<<<gv, kB, 0, st>>>(dj);
    seg_fill_kernel<<<gv, kB, 0, st>>>(dj, dbase, seg_of);
    FS_HIP_CHECK(hipGetLastError());
    if (total_v > 0)
        FS_TRY(radix_sort_seg_full(cat_key, (int)total_v, seg_of, nj, cat_skey, cat_perm, sort_tmp,
                                   st));
    perm_local_kernel<<<gv, kB, 0, st>>>(dj, dbase);
    head_kernel<<<gv, kB, 0, st>>>(dj);
    const dim3 gs(cdiv(std::max(maxv, 1), kTile), (unsigned)nj);
    scan_tile_kernel<<<gs, kB, 0, st>>>(dj, 1);
    scan_blocks_kernel<<<dim3(1, (unsigned)nj), kB, 0, st>>>(dj, 1);
    scan_add_kernel<<<gs, kB, 0, st>>>(dj, 1);
    map_kernel<<<gv, kB, 0, st>>>(dj);
    FS_HIP_CHECK(hipGetLastError());
    FS_TRY(readback());

    for (int k = 0; k < nj; ++k) {
        const Job& j = hj[k];
        Output& o = out[k];
        o.P = hc[k].P;
        o.V = hc[k].V;
        o.U = hc[k].U;
        o.rec = j.rec;
        o.x = j.x;
        o.p = j.p;
        o.f = j.f;
        o.bw = j.bw;
        o.ux = j.ux;
        o.roff = j.roff;
        o.members = j.members;
        o.uof = j.uof;
    }
    return FS_OK;
}

}  // namespace amr
}  // namespace fs
