// The FARRSIGHT step loop on the device: AMRSimulation (reference
// src/simulation_files/AMRSimulation*.cpp) with every species' state
// resident in HBM, for one deck or a batch of independent decks.
//
// Layout: the species of all decks are concatenated (deck by deck, each in
// AMRSimulation::gather order) in one set of arrays xs, ps, q_ws, es, fs;
// species s owns [off_s, off_s+n_s), deck d the union of its species. Each
// point carries its species index, which gives q/m and the deck's time step
// to the RK4 kernels. The field is a batched direct sum (fs_field_batch.hip)
// with one segment per deck, so every deck sees only its own charges. A step
// enqueues, on the context's stream and without host round trips for
// non-adaptive decks:
//   rk4 stage 1 | field | stage 2 | field | stage 3 | field | final update
//   (AMRSimulation_step.cpp:122-263; contraction off, so every update rounds
//   as the reference's expression), then either
//   remesh of every species of every deck (old-mesh coefficients and search
//   records, grid interpolation, reset of x/p to the base grid, q_ws = base
//   weight * f) and the reduced-x field (segment sums in point order, direct
//   sum over each deck's unique x, scatter back; AMRSimulation_field.cpp:
//   5-162, AMRStructure.cpp:147-196), or a full field evaluation.
// That sequence is captured once in a HIP graph and replayed every step
// (the plans it uses are built at layout time; nothing in it allocates,
// clears with memset or synchronizes).
// Adaptive decks refine their tree on the device after the grid
// interpolation (fs_amr.hip: one pass per refinement level, a few counters
// read back per pass); the sequential host builder stays available as the
// selectable second mode and as the fallback when a device pass reports a
// status. A step whose remesh refines adaptively runs eagerly; every other
// step (no remesh, or a remesh in which only static decks reset to their
// grids) replays the graph. Treecode decks read their list sizes back and
// run eagerly.
#pragma clang fp contract(off)

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "fs_amr.h"
#include "fs_field_batch.h"
#include "fs_mesh.h"
#include "fs_remesh.h"
#include "fs_sim_internal.h"

namespace fs {
namespace sim {

constexpr int kBlock = 256;

// Per species: q/m and the deck's time step (h = dt/2, dt, dt/6 rounded as the
// reference's host expressions).
struct SpeciesInfo {
    double qm, h, dt, dt6;
};

// The external drive of each species' deck at the two times a step needs:
// slot 0 the RK4 mid stages (t + dt/2), slot 1 the last stage and the field
// after the step (t + dt), or the deck's current time outside a step.
struct DriveDev {
    int on, pad;
    double k;
    double a[2], omt[2];
};

// (*calculate_e_external) (AMRSimulation_field.cpp:70-72, 169-171):
// e += a(t) sin(k x - omega t) with the reference's rounding (k x and the
// product each rounded, no FMA).
__global__ __launch_bounds__(kBlock) void drive_kernel(int n, const int32_t* __restrict__ sid,
                                                       const DriveDev* __restrict__ tab, int slot,
                                                       const double* __restrict__ x,
                                                       double* __restrict__ e) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const DriveDev D = tab[sid[i]];
    if (D.on) e[i] += D.a[slot] * sin(D.k * x[i] - D.omt[slot]);
}

// k1 -> k2 (AMRSimulation_step.cpp:148-164): ef1 = es q/m, p2 = p + dt/2 ef1,
// x_k = x + dt/2 p1 (p is the velocity).
__global__ __launch_bounds__(kBlock) void rk4_stage1(int n, const int32_t* __restrict__ sid,
                                                     const SpeciesInfo* __restrict__ si,
                                                     const double* __restrict__ xs,
                                                     const double* __restrict__ ps,
                                                     const double* __restrict__ es,
                                                     double* __restrict__ p2,
                                                     double* __restrict__ xk) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const SpeciesInfo I = si[sid[i]];
    const double ef1 = es[i] * I.qm;
    p2[i] = ps[i] + I.h * ef1;
    xk[i] = xs[i] + I.h * ps[i];
}

// k2 -> k3 (:179-189)
__global__ __launch_bounds__(kBlock) void rk4_stage2(int n, const int32_t* __restrict__ sid,
                                                     const SpeciesInfo* __restrict__ si,
                                                     const double* __restrict__ xs,
                                                     const double* __restrict__ ps,
                                                     const double* __restrict__ e2,
                                                     const double* __restrict__ p2,
                                                     double* __restrict__ p3,
                                                     double* __restrict__ xk) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const SpeciesInfo I = si[sid[i]];
    const double ef2 = e2[i] * I.qm;
    p3[i] = ps[i] + I.h * ef2;
    xk[i] = xs[i] + I.h * p2[i];
}

// k3 -> k4 (:212-220)
__global__ __launch_bounds__(kBlock) void rk4_stage3(int n, const int32_t* __restrict__ sid,
                                                     const SpeciesInfo* __restrict__ si,
                                                     const double* __restrict__ xs,
                                                     const double* __restrict__ ps,
                                                     const double* __restrict__ e3,
                                                     const double* __restrict__ p3,
                                                     double* __restrict__ p4,
                                                     double* __restrict__ xk) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const SpeciesInfo I = si[sid[i]];
    const double ef3 = e3[i] * I.qm;
    p4[i] = ps[i] + I.dt * ef3;
    xk[i] = xs[i] + I.dt * p3[i];
}

// u_{n+1} = u_n + dt/6 (k1 + 2 k2 + 2 k3 + k4) (:236-248)
__global__ __launch_bounds__(kBlock) void rk4_final(
    int n, const int32_t* __restrict__ sid, const SpeciesInfo* __restrict__ si,
    double* __restrict__ xs, double* __restrict__ ps, const double* __restrict__ es,
    const double* __restrict__ e2, const double* __restrict__ e3, const double* __restrict__ e4,
    const double* __restrict__ p2, const double* __restrict__ p3, const double* __restrict__ p4) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const SpeciesInfo I = si[sid[i]];
    const double qm = I.qm, dt6 = I.dt6;
    const double p1 = ps[i];
    const double ef1 = es[i] * qm, ef2 = e2[i] * qm, ef3 = e3[i] * qm, ef4 = e4[i] * qm;
    xs[i] = xs[i] + dt6 * (p1 + 2 * p2[i] + 2 * p3[i] + p4[i]);
    ps[i] = p1 + dt6 * (ef1 + 2 * ef2 + 2 * ef3 + ef4);
}

// After a remesh: points back on the base grid, f interpolated, q_ws = base
// weight * f (AMRStructure_amr.cpp:1096-1099), every species in one launch.
struct ResetJob {
    int n;
    const double* gx;
    const double* gp;
    const double* fnew;
    const double* bw;
    double* xs;
    double* ps;
    double* fs;
    double* qw;
    const int* on;  // the species remeshes in this step (its deck's period)
};

__global__ __launch_bounds__(kBlock) void reset_points_jobs(const ResetJob* __restrict__ jobs) {
    const ResetJob& J = jobs[blockIdx.y];
    if (*J.on == 0) return;
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= J.n) return;
    const double f = J.fnew[i];
    J.xs[i] = J.gx[i];
    J.ps[i] = J.gp[i];
    J.fs[i] = f;
    J.qw[i] = J.bw[i] * f;
}

// reduced_ws[u] += q_ws[i] over the points of each unique x, in point order
// (AMRStructure.cpp:177-180).
// One wavefront per unique x: the lanes fetch 64 members at a time, lane 0
// adds them in point order.
__global__ __launch_bounds__(64) void reduced_sum(int U, const int32_t* __restrict__ offsets,
                                                  const int32_t* __restrict__ members,
                                                  const double* __restrict__ qw,
                                                  double* __restrict__ rw) {
    __shared__ double vals[64];
    const int u = blockIdx.x;
    if (u >= U) return;
    const int k0 = offsets[u], k1 = offsets[u + 1];
    double s = 0.0;
    for (int c = k0; c < k1; c += 64) {
        const int k = c + (int)threadIdx.x;
        vals[threadIdx.x] = (k < k1) ? qw[members[k]] : 0.0;
        __syncthreads();
        if (threadIdx.x == 0) {
            const int m = min(64, k1 - c);
            for (int t = 0; t < m; ++t) s += vals[t];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) rw[u] = s;
}

// es[i] = reduced_es[unique(i)] (AMRStructure.cpp:191-195), at the points of
// the species that remeshed (on[species]).
__global__ __launch_bounds__(kBlock) void scatter_reduced(int n,
                                                          const int32_t* __restrict__ unique_of,
                                                          const int32_t* __restrict__ sid,
                                                          const int* __restrict__ on,
                                                          const double* __restrict__ re,
                                                          double* __restrict__ es) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n || on[sid[i]] == 0) return;
    es[i] = re[unique_of[i]];
}

// A tree refined on the device joins the batch arrays: q_ws = weight * f.
__global__ __launch_bounds__(kBlock) void qw_kernel(int n, const double* __restrict__ bw,
                                                    const double* __restrict__ f,
                                                    double* __restrict__ qw) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) qw[i] = bw[i] * f[i];
}

// ... and its reduced-x map, placed at the species' offsets: points from o,
// distinct x from ub.
__global__ __launch_bounds__(kBlock) void place_map_kernel(
    int n, int U, int o, int ub, int k, const int32_t* __restrict__ roff,
    const int32_t* __restrict__ members, const int32_t* __restrict__ uof,
    const double* __restrict__ ux, int32_t* __restrict__ r_off, int32_t* __restrict__ r_mem,
    int32_t* __restrict__ r_uof, double* __restrict__ rx, int32_t* __restrict__ rsid) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        r_mem[o + i] = o + members[i];
        r_uof[o + i] = ub + uof[i];
    }
    if (i < U) {
        r_off[ub + i] = o + roff[i];
        rx[ub + i] = ux[i];
        rsid[ub + i] = k;
    }
}

__global__ void set_int_kernel(int32_t* a, int v) { *a = v; }

// The species of every point: the last species whose range starts at or
// before it (ranges ascending and contiguous).
__global__ __launch_bounds__(kBlock) void sid_kernel(int n, const int2* __restrict__ range,
                                                     int nsp, int32_t* __restrict__ sid) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    int lo = 0, hi = nsp - 1;
    while (lo < hi) {
        const int m = (lo + hi + 1) >> 1;
        if (range[m].x <= i) lo = m;
        else hi = m - 1;
    }
    sid[i] = lo;
}

__global__ __launch_bounds__(kBlock) void zero_ints(int* a, int n) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) a[i] = 0;
}

inline unsigned blocks(int n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// f0 of the deck (src/structure_files/initial_distributions.cpp:12-152 via
// AMRSimulation_load.cpp:149-221), same expression order.
double initial_f(const fs_species_params& sp, double kx, double x, double p) {
    const double pth = sp.pth, pstr = sp.pstr, amp = sp.amp;
    auto maxw = [](double v, double th, double c) {
        return 1.0 / std::sqrt(2.0 * M_PI) / th * std::exp(-(v - c) * (v - c) / 2 / th / th);
    };
    switch (sp.ics_type) {
        case 3:
            return 1.0 / std::sqrt(2.0 * M_PI) / pth *
                   std::exp(-(p - 0.0) * (p - 0.0) / 2 / pth / pth) *
                   (1 + amp * std::cos(kx * x)) * p * p;
        case 4:
            return .5 * (maxw(p - pstr, pth, 0.0) + maxw(p + pstr, pth, 0.0)) *
                   (1 + amp * std::cos(kx * x));
        case 11: {
            const double nb = sp.beam_frac;
            return ((1.0 - nb) * maxw(p, pth, pstr) + nb * maxw(p, sp.beam_vth, sp.beam_v)) *
                   (1 + amp * std::cos(kx * x));
        }
        default:
            return 1.0 / std::sqrt(2.0 * M_PI) / pth *
                   std::exp(-(p - pstr) * (p - pstr) / 2 / pth / pth) *
                   (1 + amp * std::cos(kx * x));
    }
}

struct Species {
    fs_species_params prm{};
    int deck = 0;
    double q = -1, qm = -1, m = 1, kx = 0;
    mesh::Spec spec;
    bool adaptive = false;
    mesh::PanelTree base;  // the uniform base mesh (structure and grid)
    mesh::PanelTree tree;  // the current mesh
    mesh::PanelTree next;  // the adaptive remesh's work tree (buffers kept between remeshes)
    int nx = 0, np = 0;
    std::vector<int32_t> rows;
    std::vector<double> base_w;
    mesh::ReducedMap red;
    double f_beyond = 0;
    int off = -1;          // offset in the concatenated arrays (-1: not resident)

    DevBuf d_rec, d_grid, d_bw, d_fnew, d_old, d_work;
    DevBuf d_red;  // the reduced-x map (ux, offsets, members, unique_of) of a host tree
    double *gx = nullptr, *gp = nullptr;
    int32_t* rows_d = nullptr;
    remesh::Job job{};  // this species' grid interpolation (device pointers)

    // Adaptive species refined on the device: the current tree lives in
    // `aw` (records, and until the next layout its points), with its weights
    // and reduced-x map; `tree` on the host is stale until host_tree()
    // downloads it. The base mesh's records stay on the device.
    amr::Work aw;
    DevBuf d_base_rec;
    bool dev_tree = false;   // the current tree is aw's
    bool host_valid = true;  // `tree` holds the current tree
    bool from_amr = false;   // the state is in aw, not yet in the batch arrays
    amr::Output amr_out;

    int n() const { return dev_tree ? amr_out.V : (int)tree.xs.size(); }
    int npanels() const { return dev_tree ? amr_out.P : tree.size(); }
    int nunique() const { return dev_tree ? amr_out.U : (int)red.ux.size(); }
    // the device blocks back to the context's pool (every use of them is on
    // the context's stream)
    void release(BlockPool& pool) {
        for (DevBuf* b : {&d_rec, &d_grid, &d_bw, &d_fnew, &d_old, &d_work, &d_base_rec, &d_red})
            pool.put(*b);
        aw.release(&pool);
    }
    void release() {
        for (DevBuf* b : {&d_rec, &d_grid, &d_bw, &d_fnew, &d_old, &d_work, &d_base_rec, &d_red})
            b->release();
        aw.release(nullptr);
    }
};

// The field of a deck (make_field_return_ptr, AMRSimulation_load.cpp:49-96):
// the treecode for periodic decks that ask for it, the open-boundary direct
// sum for open decks without it, the periodic direct sum otherwise (open
// decks with use_treecode included, :61-73).
inline bool uses_treecode(const fs_sim_params& p) { return p.use_treecode > 0 && p.bcs == 0; }
inline bool open_field(const fs_sim_params& p) { return p.bcs != 0 && p.use_treecode <= 0; }

struct Deck {
    fs_sim_params prm{};
    std::vector<std::unique_ptr<Species>> sp;
    int off = 0, n = 0;    // points
    int uoff = 0, u = 0;   // reduced points
    int iter = 0;
    double t = 0;
    bool adaptive = false;
    double Lx() const { return prm.x_max - prm.x_min; }
};

}  // namespace sim
}  // namespace fs

using fs::DevBuf;
namespace msh = fs::mesh;
namespace rm = fs::remesh;

struct fs_sim {
    fs_ctx* ctx = nullptr;
    std::vector<std::unique_ptr<fs::sim::Deck>> decks;
    std::vector<fs::sim::Species*> sp;  // all species, deck by deck
    int N = 0;
    int U = 0;
    int iter = 0;
    int precision = FS_PRECISION_F64;
    bool use_graph = true;
    bool tune_pending = true;  // autotune the field launch shape at the next layout
    int tuned_tpt = 2, tuned_wgs = 8;
    bool any_adaptive = false, any_treecode = false;
    // concatenated state: xs, ps, qw, es, fs; RK4 work: xk, p2, p3, p4, e2, e3, e4
    DevBuf state, rk, red, status, jobs, pinfo;
    fs::TreeWork tree_work;  // treecode work space (grows to the largest treecode deck)
    DevBuf diag_work;  // diagnostics work space
    fs::FieldBatch ffull, fred;
    rm::Job* d_jobs = nullptr;            // grid jobs, one per species
    fs::sim::ResetJob* d_reset = nullptr; // reset descriptors, one per species
    rm::Launch launch;
    double *xs = nullptr, *ps = nullptr, *qw = nullptr, *es = nullptr, *fs_ = nullptr;
    double *xk = nullptr, *p2 = nullptr, *p3 = nullptr, *p4 = nullptr, *e2 = nullptr,
           *e3 = nullptr, *e4 = nullptr;
    double *rx = nullptr, *rw = nullptr, *re = nullptr;
    int32_t *r_off = nullptr, *r_mem = nullptr, *r_uof = nullptr;
    int32_t* rsid = nullptr;  // species of each reduced point
    int32_t* sid = nullptr;
    fs::sim::SpeciesInfo* sinfo = nullptr;
    // external drives: per-species table, its host copy (kept until the copy
    // to the device has run: every step synchronizes)
    bool any_drive = false;
    DevBuf drive;
    fs::sim::DriveDev* drive_d = nullptr;
    std::vector<fs::sim::DriveDev> drive_h;
    int* d_status = nullptr;
    hipGraphExec_t g_remesh = nullptr, g_plain = nullptr;
    // species that remesh in the current step (decks of different
    // remesh_period share a batch): device flags read by the remesh jobs,
    // the reset and the scatter of the reduced field; the host copy last
    // written to the device
    DevBuf onbuf;
    int* d_on = nullptr;
    std::vector<int> on_dev;
    std::vector<char> deck_on;  // the decks that remesh in this step (empty: all)
    hipGraphExec_t g_mixed = nullptr;
    // remesh timing (fs_sim_set_timing): events around the remesh's device
    // work of eager steps
    bool timing = false;
    bool timed = false;  // the events of the last step enclose a remesh
    hipEvent_t ev_r0 = nullptr, ev_r1 = nullptr;
    // adaptive refinement: 0 on the device, 1 on the host (the host
    // builder, also taken when the device passes meet a link they do not
    // reproduce); remeshes done each way
    int amr_mode = 0;
    int64_t amr_device = 0, amr_host = 0;
    DevBuf amr_jobs, amr_sort;
    int64_t steps_replayed = 0, steps_eager = 0;  // graph replays, eager steps

    hipStream_t stream() const { return ctx->stream; }
};

namespace fs {
namespace sim {
namespace {

using Sim = ::fs_sim;

void drop_graphs(Sim& S) {
    if (S.g_remesh) (void)hipGraphExecDestroy(S.g_remesh);
    if (S.g_plain) (void)hipGraphExecDestroy(S.g_plain);
    if (S.g_mixed) (void)hipGraphExecDestroy(S.g_mixed);
    S.g_remesh = S.g_plain = S.g_mixed = nullptr;
}

// Switches the species of the decks with deck_on[d] on for the remesh and
// the reduced field (all on when deck_on is empty); copies the flags to the
// device when they changed.
int set_active(Sim& S, const std::vector<char>& deck_on) {
    std::vector<int> h(std::max<size_t>(S.sp.size(), 1), 1);
    for (size_t k = 0; k < S.sp.size(); ++k)
        h[k] = deck_on.empty() ? 1 : (deck_on[S.sp[k]->deck] ? 1 : 0);
    if (h == S.on_dev) return FS_OK;
    S.on_dev = h;
    FS_HIP_CHECK(hipMemcpyAsync(S.d_on, S.on_dev.data(), sizeof(int) * S.on_dev.size(),
                                hipMemcpyHostToDevice, S.stream()));
    return FS_OK;
}

// E_MQ_Treecode on one deck's slice (tree, lists and sums on the device).
int treecode_deck(Sim& S, const Deck& D, double* out, const double* x, const double* w, int n) {
    return treecode_f64(x, w, n, D.Lx(), D.prm.greens_epsilon, D.prm.mac, D.prm.degree,
                        D.prm.max_source, out, S.tree_work, S.stream(), nullptr);
}

// The drive table for the next field evaluations: in a step, slot 0 at
// t + dt/2 (AMRSimulation_step.cpp:170, 201) and slot 1 at t + dt (:224, and
// the field after the step at the advanced t); outside a step, slot 1 at the
// deck's current t. Amplitudes on the host with the reference's
// expressions; a replayed graph reads them from device memory.
int set_drive(Sim& S, bool in_step) {
    if (!S.any_drive) return FS_OK;
    for (size_t k = 0; k < S.sp.size(); ++k) {
        const Deck& D = *S.decks[S.sp[k]->deck];
        DriveDev v{};
        if (D.prm.use_external_field > 0) {
            const fs_sim_params& P = D.prm;
            const double t0 = D.t + P.dt * 0.5;
            const double t1 = in_step ? D.t + P.dt : D.t;
            v.on = 1;
            v.k = P.k;
            v.a[0] = external_amplitude(P.external_type, t0, P.Am, P.k);
            v.a[1] = external_amplitude(P.external_type, t1, P.Am, P.k);
            v.omt[0] = P.omega * t0;
            v.omt[1] = P.omega * t1;
        }
        S.drive_h[k] = v;
    }
    FS_HIP_CHECK(hipMemcpyAsync(S.drive_d, S.drive_h.data(), sizeof(DriveDev) * S.drive_h.size(),
                                hipMemcpyHostToDevice, S.stream()));
    return FS_OK;
}

// E at every point: the batched direct sums over all direct-sum decks
// (periodic and open), then the treecode decks one by one, then the drives
// (slot: set_drive). Mixed precision: probe = false keeps each deck's FP64
// band from the last evaluation that measured it: RK4 stages 2 and 4
// measure, stage 3 (positions half a stage from stage 2's) and the full
// field after a step without remesh (a sixth of a stage from stage 4's)
// reuse theirs.
int field_full(Sim& S, double* out, const double* x, int slot, bool probe = true) {
    S.ffull.set_probe(probe);
    const int rc = S.ffull.run(x, S.qw, out, S.stream());
    S.ffull.set_probe(true);
    FS_TRY(rc);
    if (S.any_treecode)
        for (auto& D : S.decks)
            if (uses_treecode(D->prm))
                FS_TRY(treecode_deck(S, *D, out + D->off, x + D->off, S.qw + D->off, D->n));
    if (S.any_drive && S.N > 0) {
        drive_kernel<<<blocks(S.N), kBlock, 0, S.stream()>>>(S.N, S.sid, S.drive_d, slot, x, out);
        FS_HIP_CHECK(hipGetLastError());
    }
    return FS_OK;
}

// evaluate_field_uniform_grid (AMRSimulation_field.cpp:5-162)
int field_reduced(Sim& S) {
    hipStream_t st = S.stream();
    if (S.U > 0) {
        reduced_sum<<<(unsigned)S.U, 64, 0, st>>>(S.U, S.r_off, S.r_mem, S.qw, S.rw);
        FS_HIP_CHECK(hipGetLastError());
    }
    FS_TRY(S.fred.run(S.rx, S.rw, S.re, st));
    if (S.any_treecode)
        for (auto& D : S.decks)
            if (uses_treecode(D->prm))
                FS_TRY(treecode_deck(S, *D, S.re + D->uoff, S.rx + D->uoff, S.rw + D->uoff, D->u));
    // the drive at the reduced points (slot 1), before they are scattered
    // (AMRSimulation_field.cpp:70-72)
    if (S.any_drive && S.U > 0) {
        drive_kernel<<<blocks(S.U), kBlock, 0, st>>>(S.U, S.rsid, S.drive_d, 1, S.rx, S.re);
        FS_HIP_CHECK(hipGetLastError());
    }
    scatter_reduced<<<blocks(S.N), kBlock, 0, st>>>(S.N, S.r_uof, S.sid, S.d_on, S.re, S.es);
    FS_HIP_CHECK(hipGetLastError());
    return FS_OK;
}

// rk4_step (AMRSimulation_step.cpp:122-263)
int rk4(Sim& S) {
    hipStream_t st = S.stream();
    const int N = S.N;
    if (N == 0) return FS_OK;
    rk4_stage1<<<blocks(N), kBlock, 0, st>>>(N, S.sid, S.sinfo, S.xs, S.ps, S.es, S.p2, S.xk);
    FS_TRY(field_full(S, S.e2, S.xk, 0));
    rk4_stage2<<<blocks(N), kBlock, 0, st>>>(N, S.sid, S.sinfo, S.xs, S.ps, S.e2, S.p2, S.p3,
                                             S.xk);
    FS_TRY(field_full(S, S.e3, S.xk, 0, false));
    rk4_stage3<<<blocks(N), kBlock, 0, st>>>(N, S.sid, S.sinfo, S.xs, S.ps, S.e3, S.p3, S.p4,
                                             S.xk);
    FS_TRY(field_full(S, S.e4, S.xk, 1));
    rk4_final<<<blocks(N), kBlock, 0, st>>>(N, S.sid, S.sinfo, S.xs, S.ps, S.es, S.e2, S.e3,
                                            S.e4, S.p2, S.p3, S.p4);
    FS_HIP_CHECK(hipGetLastError());
    return FS_OK;
}

int read_status(Sim& S) {
    const int n = (int)S.sp.size();
    if (n == 0) return FS_OK;
    std::vector<int> h(n, 0);
    FS_HIP_CHECK(hipMemcpyAsync(h.data(), S.d_status, sizeof(int) * n, hipMemcpyDeviceToHost,
                                S.stream()));
    FS_HIP_CHECK(hipStreamSynchronize(S.stream()));
    int all = 0;
    for (int k = 0; k < n; ++k) all |= h[k];
    return rm::status_to_error(all);
}

int clear_status_all(Sim& S) {
    const int n = (int)S.sp.size();
    if (n == 0) return FS_OK;
    zero_ints<<<blocks(n), kBlock, 0, S.stream()>>>(S.d_status, n);
    FS_HIP_CHECK(hipGetLastError());
    return FS_OK;
}

// The host copy of a device-refined tree's panel records.
int host_tree(Sim& S, Species& s) {
    if (!s.dev_tree || s.host_valid) return FS_OK;
    s.tree.rec.resize((size_t)s.npanels() * msh::kInts);
    FS_HIP_CHECK(hipMemcpyAsync(s.tree.rec.data(), s.amr_out.rec,
                                sizeof(int32_t) * s.tree.rec.size(), hipMemcpyDeviceToHost,
                                S.stream()));
    FS_HIP_CHECK(hipStreamSynchronize(S.stream()));
    s.host_valid = true;
    return FS_OK;
}

// AMRSimulation::remesh (AMRSimulation_remesh.cpp:3-48), device-only path:
// every species' grid interpolation in one launch per kernel, then the
// points reset to the grid and the reduced-x field.
int remesh_static(Sim& S) {
    hipStream_t st = S.stream();
    if (S.sp.empty()) return FS_OK;
    if (S.timing) FS_HIP_CHECK(hipEventRecord(S.ev_r0, st));
    FS_TRY(rm::clear_status(S.d_jobs, S.launch, st));
    FS_TRY(rm::prepare_old_meshes(S.d_jobs, S.launch, st));
    FS_TRY(rm::interp_grids(S.d_jobs, S.launch, st));
    reset_points_jobs<<<dim3(blocks(S.launch.max_points), (unsigned)S.sp.size()), kBlock, 0, st>>>(
        S.d_reset);
    FS_HIP_CHECK(hipGetLastError());
    if (S.timing) {
        FS_HIP_CHECK(hipEventRecord(S.ev_r1, st));
        S.timed = true;
    }
    return field_reduced(S);
}

// Per-species device copies of what a remesh reads (grid, rows, base weights,
// panel records) and its work space.
int upload_species(Sim& S, const Deck& D, Species& s, bool grid_too) {
    hipStream_t st = S.stream();
    const int n = s.n(), nb = s.nx * s.np, P = s.npanels();
    s.aw.pool = &S.ctx->pool;
    // a tree refined on the device is already there (records, weights); a
    // static species' tree and weights never change once uploaded
    if (!s.dev_tree && (grid_too || s.adaptive)) {
        FS_TRY(S.ctx->pool.ensure(s.d_rec, sizeof(int32_t) * s.tree.rec.size(), st));
        FS_HIP_CHECK(hipMemcpyAsync(s.d_rec.ptr, s.tree.rec.data(),
                                    sizeof(int32_t) * s.tree.rec.size(), hipMemcpyHostToDevice, st));
        FS_TRY(S.ctx->pool.ensure(s.d_bw, sizeof(double) * n, st));
        FS_HIP_CHECK(hipMemcpyAsync(s.d_bw.ptr, s.base_w.data(), sizeof(double) * n,
                                    hipMemcpyHostToDevice, st));
        const mesh::ReducedMap& r = s.red;
        const size_t U = r.ux.size();
        const size_t bu = align_up(sizeof(double) * U, 256);
        const size_t bo = align_up(sizeof(int32_t) * (U + 1), 256);
        const size_t bm = align_up(sizeof(int32_t) * r.members.size(), 256);
        FS_TRY(S.ctx->pool.ensure(s.d_red, bu + bo + bm + sizeof(int32_t) * r.unique_of.size() + 256,
                                  st));
        if (U > 0)
            FS_HIP_CHECK(hipMemcpyAsync(s.d_red.ptr, r.ux.data(), sizeof(double) * U,
                                        hipMemcpyHostToDevice, st));
        FS_HIP_CHECK(hipMemcpyAsync(s.d_red.as<int32_t>(bu), r.offsets.data(),
                                    sizeof(int32_t) * r.offsets.size(), hipMemcpyHostToDevice, st));
        if (!r.members.empty())
            FS_HIP_CHECK(hipMemcpyAsync(s.d_red.as<int32_t>(bu + bo), r.members.data(),
                                        sizeof(int32_t) * r.members.size(), hipMemcpyHostToDevice,
                                        st));
        if (!r.unique_of.empty())
            FS_HIP_CHECK(hipMemcpyAsync(s.d_red.as<int32_t>(bu + bo + bm), r.unique_of.data(),
                                        sizeof(int32_t) * r.unique_of.size(),
                                        hipMemcpyHostToDevice, st));
    }
    if (s.adaptive && !s.d_base_rec.ptr) {
        FS_TRY(S.ctx->pool.ensure(s.d_base_rec, sizeof(int32_t) * s.base.rec.size(), st));
        FS_HIP_CHECK(hipMemcpyAsync(s.d_base_rec.ptr, s.base.rec.data(),
                                    sizeof(int32_t) * s.base.rec.size(), hipMemcpyHostToDevice, st));
    }
    if (grid_too) {
        const size_t bg = align_up(sizeof(double) * nb, 256);
        FS_TRY(S.ctx->pool.ensure(s.d_grid, 2 * bg + sizeof(int32_t) * nb, st));
        s.gx = s.d_grid.as<double>(0);
        s.gp = s.d_grid.as<double>(bg);
        s.rows_d = s.d_grid.as<int32_t>(2 * bg);
        FS_HIP_CHECK(hipMemcpyAsync(s.gx, s.base.xs.data(), sizeof(double) * nb,
                                    hipMemcpyHostToDevice, st));
        FS_HIP_CHECK(hipMemcpyAsync(s.gp, s.base.ps.data(), sizeof(double) * nb,
                                    hipMemcpyHostToDevice, st));
        FS_HIP_CHECK(hipMemcpyAsync(s.rows_d, s.rows.data(), sizeof(int32_t) * nb,
                                    hipMemcpyHostToDevice, st));
    }
    FS_TRY(S.ctx->pool.ensure(s.d_fnew, sizeof(double) * std::max(n, nb), st));
    FS_TRY(S.ctx->pool.ensure(s.d_old, rm::mesh_bytes(P), st));
    FS_TRY(S.ctx->pool.ensure(s.d_work, rm::grid_bytes(nb, s.np), st));
    rm::Job& j = s.job;
    j = rm::Job{};
    rm::carve_mesh(j, s.d_old.ptr, P);
    rm::carve_grid(j, s.d_work.ptr, nb, s.np);
    j.rec = s.dev_tree ? s.amr_out.rec : s.d_rec.as<int32_t>();
    j.npanels = P;
    j.x = s.gx;
    j.p = s.gp;
    j.rows = s.rows_d;
    j.nx = s.nx;
    j.np = s.np;
    j.Lx = D.Lx();
    j.periodic = D.prm.bcs == 0;
    j.f_beyond = s.f_beyond;
    j.f = s.d_fnew.as<double>();
    return FS_OK;
}

// (Re)builds the concatenated arrays for the current decks. A species whose
// state is resident (off >= 0 and keep_resident) is copied device to device
// from the previous arrays; every other species is uploaded from its host
// mesh (es included when es_host is given). Plans the batched fields and
// drops the captured graphs.
int layout(Sim& S, bool keep_resident) {
    hipStream_t st = S.stream();
    drop_graphs(S);
    S.sp.clear();
    S.any_adaptive = S.any_treecode = S.any_drive = false;
    int N = 0;
    for (auto& D : S.decks) {
        D->off = N;
        for (auto& s : D->sp) {
            S.sp.push_back(s.get());
            N += s->n();
        }
        D->n = N - D->off;
        S.any_adaptive = S.any_adaptive || D->adaptive;
        S.any_treecode = S.any_treecode || uses_treecode(D->prm);
        S.any_drive = S.any_drive || D->prm.use_external_field > 0;
    }
    S.N = N;
    const int nsp = (int)S.sp.size();
    const size_t bN = align_up(sizeof(double) * std::max(N, 1), 256);
    DevBuf state;
    FS_TRY(S.ctx->pool.ensure(state, 5 * bN, st));
    double* xs = state.as<double>(0);
    double* ps = state.as<double>(bN);
    double* qw = state.as<double>(2 * bN);
    double* es = state.as<double>(3 * bN);
    double* fsv = state.as<double>(4 * bN);
    int o = 0;
    std::vector<std::vector<double>> host_qw(nsp);
    for (int k = 0; k < nsp; ++k) {
        Species& s = *S.sp[k];
        const int n = s.n();
        if (s.from_amr) {
            // refined on the device: points from its tree (q_ws below)
            const double* src[3] = {s.amr_out.x, s.amr_out.p, s.amr_out.f};
            double* dst[3] = {xs, ps, fsv};
            for (int a = 0; a < 3; ++a)
                FS_HIP_CHECK(hipMemcpyAsync(dst[a] + o, src[a], sizeof(double) * n,
                                            hipMemcpyDeviceToDevice, st));
        } else if (keep_resident && s.off >= 0) {
            const double* src[5] = {S.xs, S.ps, S.qw, S.es, S.fs_};
            double* dst[5] = {xs, ps, qw, es, fsv};
            for (int a = 0; a < 5; ++a)
                FS_HIP_CHECK(hipMemcpyAsync(dst[a] + o, src[a] + s.off, sizeof(double) * n,
                                            hipMemcpyDeviceToDevice, st));
        } else {
            std::vector<double>& q = host_qw[k];
            q.resize(n);
            for (int i = 0; i < n; ++i) q[i] = s.base_w[i] * s.tree.fs[i];
            FS_HIP_CHECK(hipMemcpyAsync(xs + o, s.tree.xs.data(), sizeof(double) * n,
                                        hipMemcpyHostToDevice, st));
            FS_HIP_CHECK(hipMemcpyAsync(ps + o, s.tree.ps.data(), sizeof(double) * n,
                                        hipMemcpyHostToDevice, st));
            FS_HIP_CHECK(hipMemcpyAsync(fsv + o, s.tree.fs.data(), sizeof(double) * n,
                                        hipMemcpyHostToDevice, st));
            FS_HIP_CHECK(hipMemcpyAsync(qw + o, q.data(), sizeof(double) * n,
                                        hipMemcpyHostToDevice, st));
        }
        o += n;
    }
    // the previous arrays must outlive the copies from them
    FS_HIP_CHECK(hipStreamSynchronize(st));
    std::swap(S.state, state);
    S.ctx->pool.put(state);
    S.xs = xs;
    S.ps = ps;
    S.qw = qw;
    S.es = es;
    S.fs_ = fsv;
    o = 0;
    for (int k = 0; k < nsp; ++k) {
        Species& s = *S.sp[k];
        s.off = o;
        if (s.from_amr && s.n() > 0) {
            qw_kernel<<<blocks(s.n()), kBlock, 0, st>>>(s.n(), s.amr_out.bw, fsv + o, qw + o);
            FS_HIP_CHECK(hipGetLastError());
        }
        s.from_amr = false;
        o += s.n();
    }
    FS_TRY(S.ctx->pool.ensure(S.rk, 7 * bN, st));
    S.xk = S.rk.as<double>(0);
    S.p2 = S.rk.as<double>(bN);
    S.p3 = S.rk.as<double>(2 * bN);
    S.p4 = S.rk.as<double>(3 * bN);
    S.e2 = S.rk.as<double>(4 * bN);
    S.e3 = S.rk.as<double>(5 * bN);
    S.e4 = S.rk.as<double>(6 * bN);
    // per-species constants and ranges; the per-point species index on the
    // device
    std::vector<SpeciesInfo> info(std::max(nsp, 1));
    std::vector<int2> range(std::max(nsp, 1), make_int2(0, 0));
    for (int k = 0; k < nsp; ++k) {
        const Species& s = *S.sp[k];
        const double dt = S.decks[s.deck]->prm.dt;
        info[k] = SpeciesInfo{s.qm, 0.5 * dt, dt, dt / 6.0};
        range[k] = make_int2(s.off, s.n());
    }
    const size_t bsid = align_up(sizeof(int32_t) * std::max(N, 1), 256);
    const size_t binfo = align_up(sizeof(SpeciesInfo) * info.size(), 256);
    FS_TRY(S.ctx->pool.ensure(S.pinfo, bsid + binfo + sizeof(int2) * range.size(), st));
    S.sid = S.pinfo.as<int32_t>(0);
    S.sinfo = S.pinfo.as<SpeciesInfo>(bsid);
    int2* d_range = S.pinfo.as<int2>(bsid + binfo);
    FS_HIP_CHECK(hipMemcpyAsync(S.sinfo, info.data(), sizeof(SpeciesInfo) * info.size(),
                                hipMemcpyHostToDevice, st));
    FS_HIP_CHECK(hipMemcpyAsync(d_range, range.data(), sizeof(int2) * range.size(),
                                hipMemcpyHostToDevice, st));
    if (N > 0 && nsp > 0) {
        sid_kernel<<<blocks(N), kBlock, 0, st>>>(N, d_range, nsp, S.sid);
        FS_HIP_CHECK(hipGetLastError());
    }
    // per-species device copies of what a remesh reads (and the host trees'
    // reduced-x maps)
    for (auto& D : S.decks)
        for (auto& sp : D->sp) FS_TRY(upload_species(S, *D, *sp, sp->d_grid.ptr == nullptr || !keep_resident));
    // reduced map: each deck's species' unique x concatenated (duplicates
    // across species kept, AMRSimulation_field.cpp:9-13), each species' map
    // placed on the device at its offsets
    int ubase = 0;
    std::vector<int> ub(std::max(nsp, 1), 0);
    {
        int k = 0;
        for (auto& D : S.decks) {
            D->uoff = ubase;
            for (auto& sp : D->sp) {
                ub[k++] = ubase;
                ubase += sp->nunique();
            }
            D->u = ubase - D->uoff;
        }
    }
    S.U = ubase;
    const size_t bU = align_up(sizeof(double) * std::max(S.U, 1), 256);
    const size_t bo = align_up(sizeof(int32_t) * (S.U + 1), 256);
    const size_t bm = align_up(sizeof(int32_t) * std::max(N, 1), 256);
    FS_TRY(S.ctx->pool.ensure(S.red, 3 * bU + bo + 3 * bm, st));
    S.rx = S.red.as<double>(0);
    S.rw = S.red.as<double>(bU);
    S.re = S.red.as<double>(2 * bU);
    S.r_off = S.red.as<int32_t>(3 * bU);
    S.r_mem = S.red.as<int32_t>(3 * bU + bo);
    S.r_uof = S.red.as<int32_t>(3 * bU + bo + bm);
    S.rsid = S.red.as<int32_t>(3 * bU + bo + 2 * bm);  // U <= N
    S.drive_h.assign(std::max(nsp, 1), DriveDev{});
    FS_TRY(S.drive.ensure(sizeof(DriveDev) * S.drive_h.size(), st));
    S.drive_d = S.drive.as<DriveDev>();
    for (int k = 0; k < nsp; ++k) {
        const Species& s = *S.sp[k];
        const int n = s.n(), U = s.nunique();
        const int m = std::max(n, U);
        if (m == 0) continue;
        const int32_t *roff, *members, *uof;
        const double* ux;
        if (s.dev_tree) {
            const amr::Output& a = s.amr_out;
            roff = a.roff;
            members = a.members;
            uof = a.uof;
            ux = a.ux;
        } else {
            const size_t cu = align_up(sizeof(double) * U, 256);
            const size_t co = align_up(sizeof(int32_t) * (U + 1), 256);
            const size_t cm = align_up(sizeof(int32_t) * n, 256);
            ux = s.d_red.as<double>(0);
            roff = s.d_red.as<int32_t>(cu);
            members = s.d_red.as<int32_t>(cu + co);
            uof = s.d_red.as<int32_t>(cu + co + cm);
        }
        place_map_kernel<<<blocks(m), kBlock, 0, st>>>(n, U, s.off, ub[k], k, roff, members, uof,
                                                       ux, S.r_off, S.r_mem, S.r_uof, S.rx,
                                                       S.rsid);
        FS_HIP_CHECK(hipGetLastError());
    }
    set_int_kernel<<<1, 1, 0, st>>>(S.r_off + S.U, N);
    FS_HIP_CHECK(hipGetLastError());
    FS_TRY(S.status.ensure(sizeof(int) * std::max(nsp, 1), st));
    S.d_status = S.status.as<int>();
    FS_TRY(S.onbuf.ensure(sizeof(int) * std::max(nsp, 1), st));
    S.d_on = S.onbuf.as<int>();
    S.on_dev.clear();
    FS_TRY(set_active(S, S.deck_on));
    zero_ints<<<blocks(std::max(nsp, 1)), kBlock, 0, st>>>(S.d_status, std::max(nsp, 1));
    FS_HIP_CHECK(hipGetLastError());
    // job descriptors: the old mesh of species k is its own slice of the state
    std::vector<rm::Job> jobs(std::max(nsp, 1));
    std::vector<ResetJob> resets(std::max(nsp, 1));
    S.launch = rm::Launch{};
    S.launch.njobs = nsp;
    for (int k = 0; k < nsp; ++k) {
        Species& s = *S.sp[k];
        const int o2 = s.off;
        rm::Job& j = s.job;
        j.ox = S.xs + o2;
        j.op = S.ps + o2;
        j.of = S.fs_ + o2;
        j.status = S.d_status + k;
        j.on = S.d_on + k;
        jobs[k] = j;
        // (a refining species' reset is replaced by its refined tree)
        const double* bw = s.dev_tree ? s.amr_out.bw : s.d_bw.as<double>();
        resets[k] = ResetJob{s.nx * s.np, s.gx, s.gp, j.f, bw, S.xs + o2,
                             S.ps + o2, S.fs_ + o2, S.qw + o2, S.d_on + k};
        S.launch.max_panels = std::max(S.launch.max_panels, s.npanels());
        S.launch.max_points = std::max(S.launch.max_points, s.nx * s.np);
        S.launch.max_rows = std::max(S.launch.max_rows, s.np);
        S.launch.max_nx = std::max(S.launch.max_nx, s.nx);
    }
    const size_t bj = align_up(sizeof(rm::Job) * jobs.size(), 256);
    FS_TRY(S.jobs.ensure(bj + sizeof(ResetJob) * resets.size(), st));
    S.d_jobs = S.jobs.as<rm::Job>();
    S.d_reset = S.jobs.as<ResetJob>(bj);
    FS_HIP_CHECK(hipMemcpyAsync(S.d_jobs, jobs.data(), sizeof(rm::Job) * jobs.size(),
                                hipMemcpyHostToDevice, st));
    FS_HIP_CHECK(hipMemcpyAsync(S.d_reset, resets.data(), sizeof(ResetJob) * resets.size(),
                                hipMemcpyHostToDevice, st));
    // batched fields: one segment per deck (treecode decks not evaluated here)
    std::vector<BatchSeg> full, redsg;
    for (size_t d = 0; d < S.decks.size(); ++d) {
        const Deck& D = *S.decks[d];
        BatchSeg a;
        a.off = D.off;
        a.n = D.n;
        a.src = (int)d;
        a.eval = !uses_treecode(D.prm);
        a.open = open_field(D.prm);
        a.L = D.Lx();
        a.eps = D.prm.greens_epsilon;
        full.push_back(a);
        a.off = D.uoff;
        a.n = D.u;
        redsg.push_back(a);
    }
    if (full.empty()) {
        BatchSeg none;
        none.eval = false;
        full.push_back(none);
        redsg.push_back(none);
    }
    S.ffull.set_pool(&S.ctx->pool);
    S.fred.set_pool(&S.ctx->pool);
    FS_TRY(S.ffull.plan(full, N, S.precision, S.ctx->num_cus, S.ctx->tuning, st));
    FS_TRY(S.fred.plan(redsg, S.U, S.precision, S.ctx->num_cus, S.ctx->tuning, st));
    FS_TRY(S.fred.set_shape(1, 8, st));
    // mixed precision: every evaluation measures its own FP32 error first
    // and sets each deck's FP64 band from it
    S.ffull.govern(true);
    S.fred.govern(true);
    // launch shape of the full field (no result depends on it): timed on this
    // batch when its membership or precision changed into a size class the
    // context has not timed yet (total points in half-octaves, decks in
    // octaves), reused otherwise and when an adaptive remesh only resized it
    FS_HIP_CHECK(hipStreamSynchronize(st));
    if (S.tune_pending && N > 0 && S.ctx->tuning.targets_per_thread == 0) {
        const int nd = (int)S.decks.size();
        const uint64_t cls = ((uint64_t)S.precision << 40) |
                             ((uint64_t)(2.0 * std::log2((double)N)) << 8) |
                             (uint64_t)std::ilogb((double)std::max(nd, 1));
        bool known = false;
        {
            std::lock_guard<std::mutex> lk(S.ctx->cache_mu);
            auto it = S.ctx->batch_shape_cache.find(cls);
            if (it != S.ctx->batch_shape_cache.end()) {
                known = true;
                S.tuned_tpt = it->second.first;
                S.tuned_wgs = it->second.second;
            }
        }
        if (known) {
            FS_TRY(S.ffull.set_shape(S.tuned_tpt, S.tuned_wgs, st));
        } else {
            FS_TRY(S.ffull.autotune(S.xs, S.qw, S.e2, st, &S.tuned_tpt, &S.tuned_wgs));
            std::lock_guard<std::mutex> lk(S.ctx->cache_mu);
            S.ctx->batch_shape_cache[cls] = {S.tuned_tpt, S.tuned_wgs};
        }
        S.tune_pending = false;
    } else if (S.ctx->tuning.targets_per_thread == 0) {
        FS_TRY(S.ffull.set_shape(S.tuned_tpt, S.tuned_wgs, st));
    }
    // the host copies above must outlive the async copies
    FS_HIP_CHECK(hipStreamSynchronize(st));
    return FS_OK;
}

// The refining species' trees rebuilt on the device from their base meshes
// and grid f (fs_amr.hip); their points, weights and reduced-x maps join the
// batch arrays at the next layout.
int refine_on_device(Sim& S, const std::vector<Species*>& refine) {
    std::vector<amr::Input> in;
    std::vector<amr::Output> out;
    for (Species* s : refine) {
        const Deck& D = *S.decks[s->deck];
        amr::Input I;
        I.base_rec = s->d_base_rec.as<int32_t>();
        I.base_panels = s->base.size();
        I.mu = s->base.min_unrefined;
        I.gx = s->gx;
        I.gp = s->gp;
        I.gf = s->d_fnew.as<double>();
        I.nb = s->nx * s->np;
        I.old = s->job;
        I.max_height = s->spec.max_height;
        I.periodic = s->spec.periodic;
        I.eps = s->spec.amr_eps;
        I.q = s->q;
        I.simpson = D.prm.quadrature == 1;
        I.w = &s->aw;
        in.push_back(std::move(I));
    }
    FS_TRY(amr::refine(in, out, S.amr_jobs, S.amr_sort, S.stream()));
    FS_TRY(read_status(S));
    for (size_t k = 0; k < refine.size(); ++k) {
        Species& s = *refine[k];
        s.amr_out = out[k];
        s.dev_tree = true;
        s.host_valid = false;
        s.from_amr = true;
    }
    return FS_OK;
}

// Remesh with adaptive refinement: grid interpolation of every species on
// the device; species without refinement are reset to their grid there and
// stay resident; the refining species are refined together on the device
// (amr::refine, every species' pass in the same launches). In the host mode,
// or when a device pass reports a status, each refining species rebuilds its
// tree on a host thread of its own (the trees are independent), evaluating
// each refinement pass' new points with one device launch
// (interpolate_from_mesh; device calls from the threads take turns on the
// stream), and is uploaded again.
int remesh_adaptive(Sim& S) {
    hipStream_t st = S.stream();
    if (S.timing) FS_HIP_CHECK(hipEventRecord(S.ev_r0, st));
    FS_TRY(rm::clear_status(S.d_jobs, S.launch, st));
    FS_TRY(rm::prepare_old_meshes(S.d_jobs, S.launch, st));
    FS_TRY(rm::interp_grids(S.d_jobs, S.launch, st));
    // every species back on its grid (the refining ones are replaced below)
    reset_points_jobs<<<dim3(blocks(S.launch.max_points), (unsigned)S.sp.size()), kBlock, 0, st>>>(
        S.d_reset);
    FS_HIP_CHECK(hipGetLastError());
    if (S.timing) {
        FS_HIP_CHECK(hipEventRecord(S.ev_r1, st));
        S.timed = true;
    }
    FS_TRY(read_status(S));
    std::vector<Species*> refine;
    for (Species* s : S.sp)
        if (s->adaptive && (S.deck_on.empty() || S.deck_on[s->deck])) refine.push_back(s);
    if (S.amr_mode == 0 && !refine.empty()) {
        if (refine_on_device(S, refine) == FS_OK) {
            S.amr_device += 1;
            FS_TRY(layout(S, true));
            return field_reduced(S);
        }
        // a link structure the device passes do not reproduce (or a device
        // error): the host builder from the same base and grid f
        (void)hipGetLastError();
        FS_HIP_CHECK(hipStreamSynchronize(st));
        FS_TRY(clear_status_all(S));
    }
    if (!refine.empty()) S.amr_host += 1;
    // each refining species rebuilds into its work tree, whose buffers stay
    // allocated from remesh to remesh (no fresh pages for ~1e5 panels)
    for (Species* s : refine) {
        s->next = s->base;
        FS_HIP_CHECK(hipMemcpyAsync(s->next.fs.data(), s->d_fnew.ptr,
                                    sizeof(double) * s->nx * s->np, hipMemcpyDeviceToHost, st));
    }
    FS_HIP_CHECK(hipStreamSynchronize(st));
    std::mutex dev_mu;
    std::vector<int> rcs(refine.size(), FS_OK);
    std::vector<std::string> errs(refine.size());
    auto work = [&](size_t q) {
        Species& s = *refine[q];
        DevBuf io;
        int& rc = rcs[q];
        // a point job on this species' old mesh (its records are already built)
        auto eval = [&](const double* x, const double* p, size_t n, double* f) {
            if (rc != FS_OK || n == 0) return;
            std::lock_guard<std::mutex> lk(dev_mu);
            if (hipSetDevice(S.ctx->device) != hipSuccess) {
                rc = fail(FS_E_HIP, "adaptive remesh: set device failed");
                return;
            }
            const size_t b = align_up(sizeof(double) * n, 256);
            const size_t bj = align_up(sizeof(rm::Job), 256);
            if ((rc = io.ensure(3 * b + bj, st)) != FS_OK) return;
            rm::Job j = s.job;
            j.x = io.as<double>(0);
            j.p = io.as<double>(b);
            j.f = io.as<double>(2 * b);
            j.nx = (int)n;
            j.np = 1;
            rm::Job* dj = io.as<rm::Job>(3 * b);
            rm::Launch l;
            l.njobs = 1;
            l.max_points = (int)n;
            if (hipMemcpyAsync((void*)j.x, x, sizeof(double) * n, hipMemcpyHostToDevice, st) != hipSuccess ||
                hipMemcpyAsync((void*)j.p, p, sizeof(double) * n, hipMemcpyHostToDevice, st) != hipSuccess ||
                hipMemcpyAsync(dj, &j, sizeof(rm::Job), hipMemcpyHostToDevice, st) != hipSuccess) {
                rc = fail(FS_E_HIP, "adaptive remesh: upload failed");
                return;
            }
            if ((rc = rm::clear_status(dj, l, st)) != FS_OK) return;
            if ((rc = rm::interp_points(dj, l, st)) != FS_OK) return;
            if (hipMemcpyAsync(f, j.f, sizeof(double) * n, hipMemcpyDeviceToHost, st) != hipSuccess) {
                rc = fail(FS_E_HIP, "adaptive remesh: download failed");
                return;
            }
            int status = 0;
            if (hipMemcpyAsync(&status, j.status, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipStreamSynchronize(st) != hipSuccess) {
                rc = fail(FS_E_HIP, "adaptive remesh: sync failed");
                return;
            }
            rc = rm::status_to_error(status);
        };
        s.next.refine_adaptively(s.spec, eval);
        {
            std::lock_guard<std::mutex> lk(dev_mu);
            (void)hipStreamSynchronize(st);
            io.release();
        }
        if (rc != FS_OK) errs[q] = fs_last_error();
    };
    if (refine.size() == 1) {
        work(0);
    } else {
        std::vector<std::thread> th;
        for (size_t q = 0; q < refine.size(); ++q) th.emplace_back(work, q);
        for (auto& t : th) t.join();
    }
    for (size_t q = 0; q < refine.size(); ++q)
        if (rcs[q] != FS_OK) return fail(rcs[q], "%s", errs[q].c_str());
    FS_HIP_CHECK(hipSetDevice(S.ctx->device));
    for (size_t q = 0; q < refine.size(); ++q) {
        Species& s = *refine[q];
        const Deck& D = *S.decks[s.deck];
        std::swap(s.tree, s.next);
        std::vector<int32_t> leaves;
        s.tree.leaf_weights(s.q, D.prm.quadrature == 1, s.base_w, leaves);
        mesh::reduced_map(s.tree.xs, s.red);
        s.off = -1;  // upload from the host tree
        s.dev_tree = false;
        s.host_valid = true;
        s.from_amr = false;
    }
    FS_TRY(layout(S, true));
    return field_reduced(S);
}

int build_species(const fs_sim_params& P, Species& s) {
    const fs_species_params& p = s.prm;
    const double Lx = P.x_max - P.x_min;
    s.q = p.q;
    s.qm = p.q / p.m;
    s.m = p.m;
    s.kx = 2.0 * M_PI / Lx * p.normalized_wavenumber;
    s.spec.x_min = P.x_min;
    s.spec.x_max = P.x_max;
    s.spec.p_min = p.p_min;
    s.spec.p_max = p.p_max;
    s.spec.initial_height = p.initial_height;
    s.spec.p_height = p.p_height;
    s.spec.max_height = p.max_height;
    s.spec.periodic = P.bcs == 0;
    s.spec.amr_eps.assign(p.amr_epsilons, p.amr_epsilons + std::max(0, std::min(5, p.n_amr_epsilons)));
    s.adaptive = p.do_adaptively_refine != 0 && !s.spec.amr_eps.empty();
    s.nx = (1 << (p.initial_height + 1)) + 1;
    s.np = (1 << (p.initial_height + p.p_height + 1)) + 1;
    try {
        s.base.build_base(s.spec);
    } catch (...) {
        return fail(FS_E_INVALID, "species: initial_height must be >= 1");
    }
    if ((int)s.base.xs.size() != s.nx * s.np)
        return fail(FS_E_INTERNAL, "species: base grid has %zu points, expected %d",
                    s.base.xs.size(), s.nx * s.np);
    s.rows = mesh::grid_rows(s.base.xs, s.base.ps, s.nx, s.np);
    // generate_mesh(f0, adaptive, initial) (AMRStructure_amr.cpp:841-969)
    s.tree = s.base;
    for (size_t i = 0; i < s.tree.xs.size(); ++i)
        s.tree.fs[i] = initial_f(p, s.kx, s.tree.xs[i], s.tree.ps[i]);
    if (s.adaptive) {
        auto f0 = [&](const double* x, const double* pp, size_t n, double* f) {
            for (size_t i = 0; i < n; ++i) f[i] = initial_f(p, s.kx, x[i], pp[i]);
        };
        s.tree.refine_adaptively(s.spec, f0);
    }
    std::vector<int32_t> leaves;
    s.tree.leaf_weights(s.q, P.quadrature == 1, s.base_w, leaves);
    s.f_beyond = *std::min_element(s.tree.fs.begin(), s.tree.fs.end());  // AMRStructure.cpp:124
    s.red = mesh::reduced_map(s.tree.xs);
    return FS_OK;
}

int check_deck(const fs_sim_params* prm, const fs_species_params* species, int n_species) {
    if (!prm || !species) return fail(FS_E_INVALID, "deck: null argument");
    if (n_species < 1 || n_species > 64) return fail(FS_E_INVALID, "deck: 1..64 species");
    if (!(prm->x_max > prm->x_min) || !(prm->dt > 0) || prm->remesh_period < 1)
        return fail(FS_E_INVALID, "deck: bad domain, dt or remesh_period");
    if (prm->bcs != 0 && prm->bcs != 1) return fail(FS_E_INVALID, "deck: bcs is 0 (periodic) or 1 (open)");
    for (int k = 0; k < n_species; ++k) {
        const fs_species_params& p = species[k];
        if (p.initial_height < 1 || p.initial_height > 12 || p.p_height < 0 ||
            p.initial_height + p.p_height > 13 || !(p.p_max > p.p_min) || p.m == 0.0)
            return fail(FS_E_INVALID, "deck: species %d has bad mesh parameters", k);
        // the remesh searches hold one grid row or column per workgroup
        if ((1 << (p.initial_height + p.p_height + 1)) + 1 > rm::kMaxLine)
            return fail(FS_E_INVALID,
                        "deck: species %d has grid lines longer than %d points "
                        "(initial_height + p_height <= 8)", k, rm::kMaxLine);
    }
    return FS_OK;
}

// Builds a deck's species meshes on the host.
int build_deck(const fs_sim_params* prm, const fs_species_params* species, int n_species,
               std::unique_ptr<Deck>& out) {
    FS_TRY(check_deck(prm, species, n_species));
    auto D = std::make_unique<Deck>();
    D->prm = *prm;
    for (int k = 0; k < n_species; ++k) {
        auto s = std::make_unique<Species>();
        s->prm = species[k];
        FS_TRY(build_species(D->prm, *s));
        D->adaptive = D->adaptive || s->adaptive;
        D->sp.push_back(std::move(s));
    }
    out = std::move(D);
    return FS_OK;
}

void renumber(Sim& S) {
    for (size_t d = 0; d < S.decks.size(); ++d)
        for (auto& s : S.decks[d]->sp) s->deck = (int)d;
}

int download_species(Sim& S, int k, double* xs, double* ps, double* fs, double* qw, double* es) {
    hipStream_t st = S.stream();
    const int o = S.sp[k]->off, n = S.sp[k]->n();
    double* dst[5] = {xs, ps, fs, qw, es};
    const double* src[5] = {S.xs, S.ps, S.fs_, S.qw, S.es};
    for (int a = 0; a < 5; ++a)
        if (dst[a])
            FS_HIP_CHECK(hipMemcpyAsync(dst[a], src[a] + o, sizeof(double) * n,
                                        hipMemcpyDeviceToHost, st));
    FS_HIP_CHECK(hipStreamSynchronize(st));
    return FS_OK;
}

// The device part of one step (everything but the adaptive host work).
// Step modes: no deck remeshes after the RK4 step (the full field at the
// moved points), every deck does (remesh and the reduced-x field), or some do
// (decks of different remesh_period): the full field for every deck, then
// the remesh and the reduced field of the remeshing decks' species, chosen
// by the device flags (set_active) so one graph serves every such step.
enum StepMode { kNoRemesh = 0, kAllRemesh = 1, kSomeRemesh = 2 };

int step_body(Sim& S, int mode) {
    FS_TRY(rk4(S));
    if (mode == kAllRemesh) return remesh_static(S);
    FS_TRY(field_full(S, S.es, S.xs, 1, false));
    if (mode == kSomeRemesh) return remesh_static(S);
    return FS_OK;
}

int capture(Sim& S, int mode, hipGraphExec_t& exec) {
    hipStream_t st = S.stream();
    FS_HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    const int rc = step_body(S, mode);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(st, &g);
    if (rc != FS_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess)
        return fail(FS_E_HIP, "step capture failed: %s", hipGetErrorString(e));
    const hipError_t ei = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
        exec = nullptr;
        return fail(FS_E_HIP, "graph instantiate failed: %s", hipGetErrorString(ei));
    }
    return FS_OK;
}

}  // namespace

// ---- batch operations for the serving loop --------------------------------

int build_deck_host(const fs_deck* deck, Deck** out) {
    if (!deck || !out) return fail(FS_E_INVALID, "deck: null argument");
    *out = nullptr;
    std::unique_ptr<Deck> D;
    FS_TRY(build_deck(deck->sim, deck->species, deck->n_species, D));
    *out = D.release();
    return FS_OK;
}

int load_deck_files(Deck& D, const char* sim_dir, const char* const* names, int iter) {
    if (!names) return fail(FS_E_INVALID, "restart: null species names");
    if (iter < 0) return fail(FS_E_INVALID, "restart: negative step");
    if (iter % D.prm.remesh_period != 0)
        return fail(FS_E_INVALID,
                    "restart: step %d is not a remesh step (remesh_period %d): the weights of "
                    "moved points are not recoverable from the files",
                    iter, D.prm.remesh_period);
    for (size_t k = 0; k < D.sp.size(); ++k) {
        Species& s = *D.sp[k];
        if (!names[k]) return fail(FS_E_INVALID, "restart: null name of species %zu", k);
        std::vector<double> xs, ps, fv;
        std::vector<int32_t> rec;
        FS_TRY(io::read_species(sim_dir, names[k], iter, xs, ps, fv, rec));
        if (!s.adaptive && (xs.size() != s.base.xs.size() || rec.size() != s.base.rec.size()))
            return fail(FS_E_INVALID,
                        "restart: species %s has %zu points and %zu panels, its deck's mesh "
                        "%zu and %d",
                        names[k], xs.size(), rec.size() / mesh::kInts, s.base.xs.size(),
                        s.base.size());
        // the points must sit on their mesh (a state written after a remesh):
        // a static species on its base grid exactly, an adaptive one on every
        // panel's 3x3 lattice (shared points carry their neighbour's rounding)
        if (!s.adaptive) {
            if (xs != s.base.xs || ps != s.base.ps)
                return fail(FS_E_INVALID,
                            "restart: species %s's points are not on its grid (not written "
                            "after a remesh)", names[k]);
        } else {
            const double tx = 1e-12 * (D.prm.x_max - D.prm.x_min);
            const double tp = 1e-12 * (s.prm.p_max - s.prm.p_min);
            for (size_t r = 0; r < rec.size(); r += mesh::kInts) {
                const int32_t* q = &rec[r + mesh::PT];
                for (int c = 0; c < 3; ++c)
                    if (std::fabs(xs[q[3 * c]] - xs[q[3 * c + 1]]) > tx ||
                        std::fabs(xs[q[3 * c]] - xs[q[3 * c + 2]]) > tx ||
                        std::fabs(ps[q[c]] - ps[q[c + 3]]) > tp ||
                        std::fabs(ps[q[c]] - ps[q[c + 6]]) > tp)
                        return fail(FS_E_INVALID,
                                    "restart: species %s's points are off its panels' lattice "
                                    "(not written after a remesh)", names[k]);
            }
        }
        s.tree.xs = std::move(xs);
        s.tree.ps = std::move(ps);
        s.tree.fs = std::move(fv);
        s.tree.rec = std::move(rec);
        std::vector<int32_t> leaves;
        s.tree.leaf_weights(s.q, D.prm.quadrature == 1, s.base_w, leaves);
        s.red = mesh::reduced_map(s.tree.xs);
        s.off = -1;
    }
    D.iter = iter;
    D.t = iter * D.prm.dt;
    return FS_OK;
}

int build_deck_restart_host(const fs_deck* deck, const char* sim_dir, const char* const* names,
                            int iter, Deck** out) {
    if (!out) return fail(FS_E_INVALID, "restart: null argument");
    *out = nullptr;
    Deck* D = nullptr;
    FS_TRY(build_deck_host(deck, &D));
    const int rc = load_deck_files(*D, sim_dir, names, iter);
    if (rc != FS_OK) {
        free_deck(D);
        return rc;
    }
    *out = D;
    return FS_OK;
}

void free_deck(Deck* d) {
    if (!d) return;
    for (auto& s : d->sp) s->release();
    delete d;
}

int64_t deck_points(const Deck* d) {
    int64_t n = 0;
    for (const auto& s : d->sp) n += s->n();
    return n;
}

size_t deck_footprint(const Deck* d, int precision) {
    // per point: state 5 + RK4 7 doubles, reduced map and species index,
    // field partials (up to 16 splits) and, in mixed precision, the sorted
    // copies; per panel: record, coefficient, walk and descent records; in
    // FP64, the pair-once partials of the deck's full field (about n^2 / 512
    // bytes, fs_sym.h)
    size_t b = 0;
    if (precision == FS_PRECISION_F64 && !uses_treecode(d->prm))
        b += sym::batch_bytes(deck_points(d), FieldTuning{});
    for (const auto& s : d->sp) {
        const size_t n = (size_t)s->n(), P = (size_t)s->npanels(), g = (size_t)(s->nx * s->np);
        b += n * (12 * 8 + 3 * 4 + 16 * 8 + (precision == FS_PRECISION_MIXED ? 96 : 0));
        b += P * (80 + 96 + 104 + 184);
        b += g * (4 * 8 + 4) + rm::grid_bytes((int)g, s->np);
        // device refinement: tree, pass state and weights/map work per point
        // and panel, the base records
        if (s->adaptive) b += n * 200 + P * 170 + (size_t)s->base.size() * 80;
    }
    return b + (size_t(1) << 20);
}

int create_empty(fs_ctx* ctx, int precision, fs_sim** out) {
    if (!ctx || !out) return fail(FS_E_INVALID, "null argument");
    if (precision != FS_PRECISION_F64 && precision != FS_PRECISION_MIXED)
        return fail(FS_E_INVALID, "unknown precision %d", precision);
    FS_HIP_CHECK(hipSetDevice(ctx->device));
    auto S = std::make_unique<fs_sim>();
    S->ctx = ctx;
    S->precision = precision;
    FS_TRY(layout(*S, true));
    *out = S.release();
    return FS_OK;
}

int attach_decks(fs_sim* S, std::vector<Deck*>& decks, std::vector<int>* indices) {
    if (decks.empty()) return FS_OK;
    auto drop_all = [&]() {
        for (Deck*& e : decks) free_deck(e), e = nullptr;
        decks.clear();
    };
    // every deck is checked before any joins, so a refused call leaves the
    // batch as it was
    for (Deck* d : decks) {
        if (!d) {
            drop_all();
            return fail(FS_E_INVALID, "attach: null deck");
        }
        if (S->precision == FS_PRECISION_MIXED && open_field(d->prm)) {
            drop_all();
            return fail(FS_E_INVALID, "open-boundary decks run in FP64 only");
        }
    }
    if (hipSetDevice(S->ctx->device) != hipSuccess) {
        drop_all();
        return fail(FS_E_HIP, "attach: hipSetDevice failed");
    }
    const size_t before = S->decks.size();
    for (Deck*& d : decks) {
        S->decks.emplace_back(d);
        d = nullptr;
    }
    decks.clear();
    renumber(*S);
    S->tune_pending = true;
    int rc = layout(*S, true);
    // the joining decks' initial field (AMRSimulation.cpp:84-89); the resident
    // decks keep theirs (mid-period ones hold the full field at moved points)
    if (rc == FS_OK) {
        S->deck_on.assign(S->decks.size(), 0);
        for (size_t k = before; k < S->decks.size(); ++k) S->deck_on[k] = 1;
        rc = set_active(*S, S->deck_on);
    }
    if (rc == FS_OK) rc = set_drive(*S, false);
    if (rc == FS_OK) rc = field_reduced(*S);
    S->deck_on.clear();
    if (rc == FS_OK && hipStreamSynchronize(S->ctx->stream) != hipSuccess)
        rc = fail(FS_E_HIP, "attach: stream synchronization failed");
    if (rc != FS_OK) {
        // the joining decks leave again; the resident ones keep their state
        // (layout copies it device to device from wherever it lives now)
        const std::string why = fs_last_error();
        for (size_t k = before; k < S->decks.size(); ++k)
            for (auto& s : S->decks[k]->sp) s->release(S->ctx->pool);
        S->decks.resize(before);
        renumber(*S);
        S->tune_pending = true;
        (void)layout(*S, true);
        return fail(rc, "%s", why.c_str());
    }
    if (indices)
        for (size_t k = before; k < S->decks.size(); ++k) indices->push_back((int)k);
    return FS_OK;
}

int detach_decks(fs_sim* S, const std::vector<int>& idx,
                 std::vector<std::vector<SpeciesState>>& out) {
    out.assign(idx.size(), {});
    FS_HIP_CHECK(hipSetDevice(S->ctx->device));
    std::vector<char> drop(S->decks.size(), 0);
    for (size_t q = 0; q < idx.size(); ++q) {
        const int d = idx[q];
        if (d < 0 || d >= (int)S->decks.size()) return fail(FS_E_INVALID, "bad deck index");
        drop[d] = 1;
        int first = 0;
        for (int e = 0; e < d; ++e) first += (int)S->decks[e]->sp.size();
        for (size_t k = 0; k < S->decks[d]->sp.size(); ++k) {
            const int g = first + (int)k;
            SpeciesState st;
            const int n = S->sp[g]->n();
            st.xs.resize(n);
            st.ps.resize(n);
            st.fs.resize(n);
            st.qw.resize(n);
            st.es.resize(n);
            FS_TRY(download_species(*S, g, st.xs.data(), st.ps.data(), st.fs.data(),
                                    st.qw.data(), st.es.data()));
            FS_TRY(fs_sim_diagnostics(S, g, st.diag));
            FS_TRY(host_tree(*S, *S->sp[g]));
            st.rec = S->sp[g]->tree.rec;
            st.iter = S->decks[d]->iter;
            out[q].push_back(std::move(st));
        }
    }
    std::vector<std::unique_ptr<Deck>> keep;
    for (size_t d = 0; d < S->decks.size(); ++d) {
        if (drop[d]) {
            for (auto& s : S->decks[d]->sp) s->release(S->ctx->pool);
        } else {
            keep.push_back(std::move(S->decks[d]));
        }
    }
    S->decks = std::move(keep);
    renumber(*S);
    S->tune_pending = true;
    return layout(*S, true);
}

}  // namespace sim
}  // namespace fs

using namespace fs;

extern "C" {

int fs_batch_create(fs_ctx* ctx, const fs_deck* decks, int n_decks, int precision,
                    fs_sim** out) {
    if (!ctx || !decks || !out) return fail(FS_E_INVALID, "fs_batch_create: null argument");
    *out = nullptr;
    if (n_decks < 1) return fail(FS_E_INVALID, "fs_batch_create: no decks");
    if (precision != FS_PRECISION_F64 && precision != FS_PRECISION_MIXED)
        return fail(FS_E_INVALID, "fs_batch_create: unknown precision %d", precision);
    for (int d = 0; d < n_decks; ++d)
        if (!decks[d].sim) return fail(FS_E_INVALID, "fs_batch_create: null deck");
    FS_HIP_CHECK(hipSetDevice(ctx->device));
    auto S = std::make_unique<fs_sim>();
    S->ctx = ctx;
    S->precision = precision;
    for (int d = 0; d < n_decks; ++d) {
        std::unique_ptr<sim::Deck> D;
        FS_TRY(sim::build_deck(decks[d].sim, decks[d].species, decks[d].n_species, D));
        S->decks.push_back(std::move(D));
    }
    sim::renumber(*S);
    FS_TRY(sim::layout(*S, false));
    FS_TRY(sim::set_drive(*S, false));
    FS_TRY(sim::field_reduced(*S));  // AMRSimulation.cpp:84-89
    FS_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    *out = S.release();
    return FS_OK;
}

int fs_sim_create(fs_ctx* ctx, const fs_sim_params* prm, const fs_species_params* species,
                  int n_species, fs_sim** out) {
    if (!ctx || !prm || !species || !out) return fail(FS_E_INVALID, "fs_sim_create: null argument");
    fs_deck d{prm, species, n_species};
    return fs_batch_create(ctx, &d, 1, FS_PRECISION_F64, out);
}

int fs_sim_add_deck(fs_sim* S, const fs_deck* deck, int* index) {
    if (!S || !deck) return fail(FS_E_INVALID, "fs_sim_add_deck: null argument");
    if (!deck->sim) return fail(FS_E_INVALID, "fs_sim_add_deck: null deck");
    FS_HIP_CHECK(hipSetDevice(S->ctx->device));
    std::unique_ptr<sim::Deck> D;
    FS_TRY(sim::build_deck(deck->sim, deck->species, deck->n_species, D));
    // resident decks keep their state; the field of the reduced points is
    // recomputed for all (bitwise the same for the resident ones); on failure
    // the batch is left as it was
    std::vector<sim::Deck*> one{D.release()};
    std::vector<int> idx;
    FS_TRY(sim::attach_decks(S, one, &idx));
    if (index) *index = idx.empty() ? -1 : idx[0];
    return FS_OK;
}

int fs_deck_footprint(const fs_deck* deck, int precision, int64_t* bytes, int64_t* n_points) {
    if (!deck || !bytes) return fail(FS_E_INVALID, "fs_deck_footprint: null argument");
    if (precision != FS_PRECISION_F64 && precision != FS_PRECISION_MIXED)
        return fail(FS_E_INVALID, "unknown precision %d", precision);
    sim::Deck* D = nullptr;
    FS_TRY(sim::build_deck_host(deck, &D));
    *bytes = (int64_t)sim::deck_footprint(D, precision);
    if (n_points) *n_points = sim::deck_points(D);
    sim::free_deck(D);
    return FS_OK;
}

int fs_sim_remove_deck(fs_sim* S, int deck) {
    if (!S || deck < 0 || deck >= (int)S->decks.size())
        return fail(FS_E_INVALID, "fs_sim_remove_deck: bad deck index");
    FS_HIP_CHECK(hipSetDevice(S->ctx->device));
    FS_HIP_CHECK(hipStreamSynchronize(S->ctx->stream));
    for (auto& s : S->decks[deck]->sp) s->release(S->ctx->pool);
    S->decks.erase(S->decks.begin() + deck);
    sim::renumber(*S);
    S->tune_pending = true;
    return sim::layout(*S, true);
}

int fs_sim_num_decks(fs_sim* S, int* n) {
    if (!S || !n) return fail(FS_E_INVALID, "null argument");
    *n = (int)S->decks.size();
    return FS_OK;
}

int fs_sim_deck_info(fs_sim* S, int deck, int* first_species, int* n_species, int* iter,
                     int64_t* n_points) {
    if (!S || deck < 0 || deck >= (int)S->decks.size())
        return fail(FS_E_INVALID, "bad deck index");
    const sim::Deck& D = *S->decks[deck];
    int first = 0;
    for (int d = 0; d < deck; ++d) first += (int)S->decks[d]->sp.size();
    if (first_species) *first_species = first;
    if (n_species) *n_species = (int)D.sp.size();
    if (iter) *iter = D.iter;
    if (n_points) *n_points = D.n;
    return FS_OK;
}

int fs_sim_set_options(fs_sim* S, int precision, int use_graph) {
    if (!S) return fail(FS_E_INVALID, "null simulation");
    if (precision != FS_PRECISION_F64 && precision != FS_PRECISION_MIXED)
        return fail(FS_E_INVALID, "unknown precision %d", precision);
    FS_HIP_CHECK(hipSetDevice(S->ctx->device));
    S->use_graph = use_graph != 0;
    if (precision != S->precision) {
        const int before = S->precision;
        S->precision = precision;
        S->tune_pending = true;
        const int rc = sim::layout(*S, true);
        if (rc != FS_OK) {  // e.g. open-boundary decks asked to run mixed
            const std::string why = fs_last_error();
            S->precision = before;
            S->tune_pending = true;
            (void)sim::layout(*S, true);
            return fail(rc, "%s", why.c_str());
        }
    } else {
        sim::drop_graphs(*S);
    }
    return FS_OK;
}

int fs_sim_set_refinement(fs_sim* S, int mode) {
    if (!S) return fail(FS_E_INVALID, "null simulation");
    if (mode != 0 && mode != 1) return fail(FS_E_INVALID, "refinement mode is 0 (device) or 1 (host)");
    S->amr_mode = mode;
    return FS_OK;
}

int fs_sim_refinement_counts(fs_sim* S, int64_t* on_device, int64_t* on_host) {
    if (!S) return fail(FS_E_INVALID, "null simulation");
    if (on_device) *on_device = S->amr_device;
    if (on_host) *on_host = S->amr_host;
    return FS_OK;
}

int fs_sim_step_counts(fs_sim* S, int64_t* replayed, int64_t* eager) {
    if (!S) return fail(FS_E_INVALID, "null simulation");
    if (replayed) *replayed = S->steps_replayed;
    if (eager) *eager = S->steps_eager;
    return FS_OK;
}

int fs_sim_set_timing(fs_sim* S, int on) {
    if (!S) return fail(FS_E_INVALID, "null simulation");
    FS_HIP_CHECK(hipSetDevice(S->ctx->device));
    if (on && !S->ev_r0) {
        FS_HIP_CHECK(hipEventCreate(&S->ev_r0));
        FS_HIP_CHECK(hipEventCreate(&S->ev_r1));
    }
    S->timing = on != 0;
    S->timed = false;
    return FS_OK;
}

int fs_sim_remesh_ms(fs_sim* S, double* ms) {
    if (!S || !ms) return fail(FS_E_INVALID, "null argument");
    if (!S->timing || !S->timed)
        return fail(FS_E_INVALID, "fs_sim_remesh_ms: no timed remesh (fs_sim_set_timing, then a "
                                  "remeshing step)");
    FS_HIP_CHECK(hipEventSynchronize(S->ev_r1));
    float t = 0.0f;
    FS_HIP_CHECK(hipEventElapsedTime(&t, S->ev_r0, S->ev_r1));
    *ms = t;
    return FS_OK;
}

int fs_sim_destroy(fs_sim* S) {
    if (!S) return FS_OK;
    (void)hipStreamSynchronize(S->ctx->stream);
    sim::drop_graphs(*S);
    for (auto& D : S->decks)
        for (auto& s : D->sp) s->release(S->ctx->pool);
    S->state.release();
    S->rk.release();
    S->red.release();
    S->status.release();
    S->jobs.release();
    S->pinfo.release();
    S->ffull.release();
    S->fred.release();
    S->tree_work.release();
    S->drive.release();
    S->onbuf.release();
    if (S->ev_r0) (void)hipEventDestroy(S->ev_r0);
    if (S->ev_r1) (void)hipEventDestroy(S->ev_r1);
    S->diag_work.release();
    S->amr_jobs.release();
    S->amr_sort.release();
    delete S;
    return FS_OK;
}

int fs_sim_step(fs_sim* S) {
    if (!S) return fail(FS_E_INVALID, "null simulation");
    FS_HIP_CHECK(hipSetDevice(S->ctx->device));
    // which decks remesh after this step (each by its own remesh_period and
    // iteration count, AMRSimulation_step.cpp:74-90)
    std::vector<char> on(S->decks.size(), 0);
    size_t nrem = 0;
    for (size_t d = 0; d < S->decks.size(); ++d) {
        const sim::Deck& D = *S->decks[d];
        on[d] = ((D.iter + 1) % D.prm.remesh_period) == 0 ? 1 : 0;
        nrem += on[d];
    }
    const int mode = nrem == 0 ? sim::kNoRemesh
                     : nrem == S->decks.size() ? sim::kAllRemesh : sim::kSomeRemesh;
    // an adaptive deck remeshing now changes the arrays' sizes (its tree is
    // rebuilt): that step runs eagerly; every other step (no remesh, or only
    // static decks remeshing) replays a captured graph
    bool adaptive_now = false;
    for (size_t d = 0; d < S->decks.size(); ++d)
        adaptive_now = adaptive_now || (on[d] && S->decks[d]->adaptive);
    S->deck_on = mode == sim::kSomeRemesh ? on : std::vector<char>{};
    FS_TRY(sim::set_active(*S, S->deck_on));
    const bool eager = adaptive_now || S->any_treecode || !S->use_graph || S->timing;
    S->timed = false;
    FS_TRY(sim::set_drive(*S, true));
    if (eager) {
        FS_TRY(sim::rk4(*S));
        if (mode != sim::kAllRemesh) FS_TRY(sim::field_full(*S, S->es, S->xs, 1, false));
        if (mode != sim::kNoRemesh) {
            if (adaptive_now) {
                FS_TRY(sim::remesh_adaptive(*S));
            } else {
                FS_TRY(sim::remesh_static(*S));
            }
        }
        S->steps_eager += 1;
    } else {
        hipGraphExec_t& g = mode == sim::kAllRemesh ? S->g_remesh
                            : mode == sim::kNoRemesh ? S->g_plain : S->g_mixed;
        if (!g) FS_TRY(sim::capture(*S, mode, g));
        FS_HIP_CHECK(hipGraphLaunch(g, S->ctx->stream));
        S->steps_replayed += 1;
    }
    S->iter += 1;
    for (auto& D : S->decks) {
        D->iter += 1;
        D->t += D->prm.dt;
    }
    S->deck_on.clear();
    if (mode != sim::kNoRemesh) return sim::read_status(*S);
    FS_HIP_CHECK(hipStreamSynchronize(S->ctx->stream));
    return FS_OK;
}

int fs_mesh_build(const fs_sim_params* prm, const fs_species_params* species,
                  int64_t* n_points, int64_t* n_panels, int32_t* panels, double* xs, double* ps,
                  double* fs, double* q_ws) {
    if (!prm || !species) return fail(FS_E_INVALID, "fs_mesh_build: null argument");
    if (!(prm->x_max > prm->x_min)) return fail(FS_E_INVALID, "fs_mesh_build: x_max <= x_min");
    const fs_species_params& p = *species;
    if (p.initial_height < 1 || p.initial_height > 12 || p.p_height < 0 ||
        p.initial_height + p.p_height > 13 || !(p.p_max > p.p_min) || p.m == 0.0)
        return fail(FS_E_INVALID, "fs_mesh_build: bad mesh parameters");
    sim::Species s;
    s.prm = p;
    FS_TRY(sim::build_species(*prm, s));
    if (n_points) *n_points = s.n();
    if (n_panels) *n_panels = s.npanels();
    if (panels) std::copy(s.tree.rec.begin(), s.tree.rec.end(), panels);
    if (xs) std::copy(s.tree.xs.begin(), s.tree.xs.end(), xs);
    if (ps) std::copy(s.tree.ps.begin(), s.tree.ps.end(), ps);
    if (fs) std::copy(s.tree.fs.begin(), s.tree.fs.end(), fs);
    if (q_ws)
        for (int i = 0; i < s.n(); ++i) q_ws[i] = s.base_w[i] * s.tree.fs[i];
    return FS_OK;
}

int fs_sim_iter(fs_sim* S, int* iter) {
    if (!S || !iter) return fail(FS_E_INVALID, "null argument");
    *iter = S->iter;
    return FS_OK;
}

int fs_sim_species_counts(fs_sim* S, int k, int64_t* n_points, int64_t* n_panels) {
    if (!S || k < 0 || k >= (int)S->sp.size()) return fail(FS_E_INVALID, "bad species index");
    if (n_points) *n_points = S->sp[k]->n();
    if (n_panels) *n_panels = S->sp[k]->npanels();
    return FS_OK;
}

int fs_sim_get_points(fs_sim* S, int k, double* xs, double* ps, double* fs, double* q_ws,
                      double* es) {
    if (!S || k < 0 || k >= (int)S->sp.size()) return fail(FS_E_INVALID, "bad species index");
    return sim::download_species(*S, k, xs, ps, fs, q_ws, es);
}

int fs_sim_field_governor(fs_sim* S, int deck, double* max_e, double* errs, double* bands) {
    if (!S || deck < 0 || deck >= (int)S->decks.size()) return fail(FS_E_INVALID, "bad deck index");
    std::vector<double> g;
    FS_TRY(S->ffull.get_gov(g, S->stream()));
    double v[fs::kGovVals] = {0};
    if ((size_t)(deck + 1) * fs::kGovVals <= g.size())
        std::memcpy(v, &g[(size_t)deck * fs::kGovVals], sizeof(v));
    if (max_e) *max_e = v[0];
    if (errs) std::memcpy(errs, v + 1, sizeof(double) * (fs::kGovVals - 1));
    if (bands) {
        const sim::Deck& D = *S->decks[deck];
        for (int q = 0; q < fs::kGovVals - 1; ++q)
            bands[q] = fs::governor_band(D.prm.greens_epsilon / D.Lx(), q);
    }
    return FS_OK;
}

int fs_sim_get_panels(fs_sim* S, int k, int32_t* panels) {
    if (!S || !panels || k < 0 || k >= (int)S->sp.size())
        return fail(FS_E_INVALID, "bad species index");
    FS_HIP_CHECK(hipSetDevice(S->ctx->device));
    FS_TRY(sim::host_tree(*S, *S->sp[k]));
    const auto& r = S->sp[k]->tree.rec;
    std::copy(r.begin(), r.end(), panels);
    return FS_OK;
}

int fs_batch_create_empty(fs_ctx* ctx, int precision, fs_sim** out) {
    return sim::create_empty(ctx, precision, out);
}

int fs_sim_write_output(fs_sim* S, int deck, const char* sim_dir, const char* const* names) {
    if (!S || deck < 0 || deck >= (int)S->decks.size() || !names)
        return fail(FS_E_INVALID, "fs_sim_write_output: bad argument");
    FS_HIP_CHECK(hipSetDevice(S->ctx->device));
    const sim::Deck& D = *S->decks[deck];
    int first = 0;
    for (int e = 0; e < deck; ++e) first += (int)S->decks[e]->sp.size();
    for (size_t k = 0; k < D.sp.size(); ++k) {
        if (!names[k]) return fail(FS_E_INVALID, "fs_sim_write_output: null species name");
        const int g = first + (int)k;
        const int n = S->sp[g]->n();
        std::vector<double> xs(n), ps(n), fv(n), qw(n), es(n);
        FS_TRY(sim::download_species(*S, g, xs.data(), ps.data(), fv.data(), qw.data(),
                                     es.data()));
        FS_TRY(sim::host_tree(*S, *S->sp[g]));
        const auto& rec = S->sp[g]->tree.rec;
        FS_TRY(io::write_species(sim_dir, names[k], D.iter, n, xs.data(), ps.data(), fv.data(),
                                 qw.data(), es.data(), (int64_t)(rec.size() / msh::kInts),
                                 rec.data()));
    }
    // the deck's reduced set and its field (the last reduced evaluation)
    std::vector<double> rx(D.u), re(D.u);
    if (D.u > 0) {
        FS_HIP_CHECK(hipMemcpyAsync(rx.data(), S->rx + D.uoff, sizeof(double) * D.u,
                                    hipMemcpyDeviceToHost, S->stream()));
        FS_HIP_CHECK(hipMemcpyAsync(re.data(), S->re + D.uoff, sizeof(double) * D.u,
                                    hipMemcpyDeviceToHost, S->stream()));
        FS_HIP_CHECK(hipStreamSynchronize(S->stream()));
    }
    return io::write_reduced(sim_dir, D.iter, D.u, rx.data(), re.data());
}

int fs_sim_restart_deck(fs_sim* S, const fs_deck* deck, const char* sim_dir,
                        const char* const* names, int iter, int* index) {
    if (!S || !deck || !deck->sim) return fail(FS_E_INVALID, "fs_sim_restart_deck: null argument");
    sim::Deck* D = nullptr;
    FS_TRY(sim::build_deck_restart_host(deck, sim_dir, names, iter, &D));
    std::vector<sim::Deck*> one{D};
    std::vector<int> idx;
    FS_TRY(sim::attach_decks(S, one, &idx));
    if (index) *index = idx.empty() ? -1 : idx[0];
    return FS_OK;
}

int fs_sim_diagnostics_full(fs_sim* S, int k, double* out, int n_out) {
    if (!S || !out || k < 0 || k >= (int)S->sp.size())
        return fail(FS_E_INVALID, "bad species index");
    if (n_out < FS_DIAG_COUNT) return fail(FS_E_INVALID, "need %d outputs", FS_DIAG_COUNT);
    const sim::Species& s = *S->sp[k];
    const sim::Deck& D = *S->decks[s.deck];
    FS_HIP_CHECK(hipSetDevice(S->ctx->device));
    const int n = s.n(), o = s.off;
    FS_TRY(S->diag_work.ensure(diag::work_bytes(n), S->stream()));
    return diag::species(S->xs + o, S->ps + o, S->fs_ + o, S->qw + o, S->es + o, n, s.q, s.qm,
                         D.Lx(), S->diag_work.ptr, out, S->stream());
}

int fs_sim_diagnostics(fs_sim* S, int k, double out[4]) {
    if (!out) return fail(FS_E_INVALID, "null output");
    double full[FS_DIAG_COUNT];
    FS_TRY(fs_sim_diagnostics_full(S, k, full, FS_DIAG_COUNT));
    for (int q = 0; q < 4; ++q) out[q] = full[q];
    return FS_OK;
}

}  // extern "C"
