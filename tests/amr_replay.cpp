// Host replay of the device refinement passes (multi_species_amd/csrc/
// fs_amr_pass.h) against the sequential host builder (fs_mesh.cpp, itself
// pinned against the oracle's restatement of refine_panels).
//
// Each launch of a pass runs its per-thread steps one after another in a
// chosen thread order (ascending, descending or shuffled: a step must not
// depend on another thread of the same launch), with every buffer filled
// with garbage first (a step must not read what no step wrote). Random
// meshes: initial height 1..5, p height 0..2, up to four levels of
// refinement, periodic or open, one to five refinement criteria, four
// shapes of f; trapezoid or Simpson weights. Compared bit for bit: panel
// records, points, f, quadrature weights, the reduced-x map.
//
// Usage: amr_replay <trials> <order 0|1|2> <seed>; exit status 0 when every
// trial matches.
#pragma clang fp contract(off)

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <numeric>
#include <random>
#include <vector>

#include "../multi_species_amd/csrc/fs_amr_pass.h"
#include "../multi_species_amd/csrc/fs_mesh.h"

using namespace fs;
using namespace fs::amr;
using namespace fs::amr::pass;

namespace {

struct HostOps {
    void or_(int* a, int v) const { *a |= v; }
    int add(int* a, int v) const {
        const int o = *a;
        *a += v;
        return o;
    }
};

int g_order = 0;
std::mt19937 g_shuffle(7);

template <class Fn>
void launch(int a, int b, Fn fn) {
    std::vector<int> ks;
    for (int k = a; k < b; ++k) ks.push_back(k);
    if (g_order == 1) std::reverse(ks.begin(), ks.end());
    if (g_order == 2) std::shuffle(ks.begin(), ks.end(), g_shuffle);
    for (int k : ks) fn(k);
}

using F = std::function<double(double, double)>;

struct Tree {
    std::vector<int32_t> rec;
    std::vector<double> x, p, f, w, ux;
    std::vector<int32_t> offsets, members, uof;
    int status = 0;
};

const int kGarbage = 0x5a5a5a5a;

Tree replay(const mesh::PanelTree& base, const mesh::Spec& s, double q, bool simpson, const F& fn) {
    const double nan = std::nan("");
    std::vector<int32_t> rec(base.rec.size(), kGarbage), chain(base.size(), 7),
        plan((size_t)base.size() * kPlanInts, 3);
    std::vector<int2> cnt(base.size(), make_int2(1, 9)), off(base.size(), make_int2(5, 5));
    std::vector<double> x(base.xs.size(), nan), p(base.xs.size(), nan), f(base.xs.size(), nan);
    std::copy(base.rec.begin(), base.rec.end(), rec.begin());
    const int nb = (int)base.xs.size();
    for (int i = 0; i < nb; ++i) {
        x[i] = base.xs[i];
        p[i] = base.ps[i];
        f[i] = fn(x[i], p[i]);
    }
    Counters C{base.size(), nb, 0, 0, 0, 0, base.min_unrefined, 0};
    remesh::Job pj{};
    Job J{};
    J.mu = base.min_unrefined;
    J.max_height = s.max_height;
    J.periodic = s.periodic;
    J.n_eps = (int)s.amr_eps.size();
    for (size_t e = 0; e < s.amr_eps.size(); ++e) J.eps[e] = s.amr_eps[e];
    J.q = q;
    J.simpson = simpson;
    J.ctr = &C;
    J.pj = &pj;
    auto views = [&]() {
        J.rec = rec.data();
        J.x = x.data();
        J.p = p.data();
        J.f = f.data();
        J.chain = chain.data();
        J.cnt = cnt.data();
        J.off = off.data();
        J.plan = plan.data();
        J.cap_p = (int)chain.size();
        J.cap_v = (int)x.size();
    };
    views();
    HostOps ops;
    auto plan_pass = [&]() {
        launch(J.mu, C.P, [&](int k) { C.flagged += test_one(J, k, ops); });
        launch(J.mu, C.P, [&](int k) { chain_clear_one(J, k); });
        launch(J.mu, C.P, [&](int k) { chain_one(J, k, ops); });
        launch(J.mu, C.P, [&](int k) { count_one(J, k); });
        int2 run = make_int2(0, 0);
        for (int k = 0; k < C.P - J.mu; ++k) {
            off[k] = run;
            run = add2(run, cnt[k]);
        }
        C.nR = run.x;
        C.newV = run.y;
    };
    Tree T;
    plan_pass();
    while (C.nR > 0 && !C.status) {
        // capacity exactly for this pass (as the device driver sizes it)
        const size_t np_ = (size_t)C.P + 4 * (size_t)C.nR, nv_ = (size_t)C.V + C.newV;
        if (np_ > chain.size()) {
            rec.resize(np_ * kI, kGarbage);
            chain.resize(np_, 7);
            plan.resize(np_ * kPlanInts, 3);
            cnt.resize(np_, make_int2(1, 9));
            off.resize(np_, make_int2(5, 5));
        }
        if (nv_ > x.size()) {
            x.resize(nv_, nan);
            p.resize(nv_, nan);
            f.resize(nv_, nan);
        }
        views();
        launch(J.mu, C.P, [&](int k) { build_one(J, k, ops); });
        launch(J.mu, C.P, [&](int k) { link_one(J, k, ops); });
        commit_one(J);
        for (int i = 0; i < pj.nx; ++i) const_cast<double*>(pj.f)[i] = fn(pj.x[i], pj.p[i]);
        plan_pass();
    }
    // weights
    const int V = C.V, P = C.P;
    std::vector<int> nslot(V, 3);
    std::vector<uint64_t> spath((size_t)V * kSlots, 0x5a5a5a5a5a5a5a5aull);
    std::vector<double> sval((size_t)V * kSlots, nan), bw(V, nan);
    J.nslot = nslot.data();
    J.spath = spath.data();
    J.sval = sval.data();
    J.bw = bw.data();
    launch(0, V, [&](int k) { leaf_init_one(J, k); });
    launch(0, P, [&](int k) { leaf_terms_one(J, k, ops); });
    launch(0, V, [&](int k) { weight_one(J, k); });
    // reduced-x map: keys, a stable sort, run heads, their scan, the map
    std::vector<uint64_t> key(V), skey(V);
    std::vector<int32_t> perm(V), members(V, kGarbage), uof(V, kGarbage), roff(V + 1, kGarbage);
    std::vector<int2> vcnt(V, make_int2(1, 9)), voff(V, make_int2(5, 5));
    std::vector<double> ux(V, nan);
    J.key = key.data();
    J.skey = skey.data();
    J.perm = perm.data();
    J.vcnt = vcnt.data();
    J.voff = voff.data();
    J.ux = ux.data();
    J.roff = roff.data();
    J.members = members.data();
    J.uof = uof.data();
    launch(0, V, [&](int k) { key_one(J, k); });
    std::iota(perm.begin(), perm.end(), 0);
    std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return key[a] < key[b]; });
    for (int k = 0; k < V; ++k) skey[k] = key[perm[k]];
    launch(0, V, [&](int k) { head_one(J, k); });
    int run = 0;
    for (int k = 0; k < V; ++k) {
        voff[k] = make_int2(run, 0);
        run += vcnt[k].x;
    }
    C.U = run;
    launch(0, V, [&](int k) { map_one(J, k); });

    T.rec.assign(rec.begin(), rec.begin() + (size_t)P * kI);
    T.x.assign(x.begin(), x.begin() + V);
    T.p.assign(p.begin(), p.begin() + V);
    T.f.assign(f.begin(), f.begin() + V);
    T.w = bw;
    T.ux.assign(ux.begin(), ux.begin() + C.U);
    T.offsets.assign(roff.begin(), roff.begin() + C.U + 1);
    T.members = members;
    T.uof = uof;
    T.status = C.status;
    return T;
}

bool same(const std::vector<double>& a, const std::vector<double>& b) {
    if (a.size() != b.size()) return false;
    for (size_t i = 0; i < a.size(); ++i)
        if (!(a[i] == b[i]) || std::signbit(a[i]) != std::signbit(b[i])) return false;
    return true;
}

}  // namespace

int main(int argc, char** argv) {
    const int trials = argc > 1 ? atoi(argv[1]) : 100;
    g_order = argc > 2 ? atoi(argv[2]) : 0;
    std::mt19937 rng(argc > 3 ? (unsigned)atoi(argv[3]) : 12345u);
    int bad = 0, refined = 0;
    for (int t = 0; t < trials; ++t) {
        std::uniform_int_distribution<int> hd(1, 5), pd(0, 2), ed(1, 4), kd(0, 3);
        auto uni = [&](double a, double b) { return std::uniform_real_distribution<double>(a, b)(rng); };
        mesh::Spec s;
        s.x_min = 0;
        s.x_max = 4 * M_PI;
        s.p_min = -6;
        s.p_max = 6;
        s.initial_height = hd(rng);
        s.p_height = std::min(pd(rng), 6 - s.initial_height);
        s.max_height = s.initial_height + s.p_height + ed(rng);
        s.periodic = (rng() % 4) != 0;
        const int kind = kd(rng);
        const double amp = uni(0.05, 0.8), vth = uni(0.3, 1.5), xc = uni(0, 4 * M_PI);
        const F fn = [=](double x, double p) {
            const double g = std::exp(-p * p / (2 * vth * vth)) / vth;
            if (kind == 0) return g * (1 + amp * std::cos(0.5 * x));
            if (kind == 1)
                return g * (1 + amp * std::cos(1.5 * x)) + 0.2 * std::exp(-(p - 3) * (p - 3) * 4);
            if (kind == 2) return std::exp(-((x - xc) * (x - xc) + p * p) * 0.8) + 0.01;
            return 0.5 * (std::exp(-(p - 2.4) * (p - 2.4) / 0.5) +
                          std::exp(-(p + 2.4) * (p + 2.4) / 0.5)) *
                   (1 + amp * std::cos(0.5 * x));
        };
        const int ne = 1 + (int)(rng() % 5);
        const double e0[5] = {0.02, 0.3, 2.0, 0.05, 0.05};
        for (int e = 0; e < ne; ++e) s.amr_eps.push_back(e0[e] * uni(0.3, 3.0));
        const bool simpson = rng() % 2;
        const double q = (rng() % 2) ? -1.0 : 0.5;

        mesh::PanelTree base;
        base.build_base(s);
        mesh::PanelTree h = base;
        for (size_t i = 0; i < h.xs.size(); ++i) h.fs[i] = fn(h.xs[i], h.ps[i]);
        h.refine_adaptively(s, [&](const double* x, const double* p, size_t n, double* f) {
            for (size_t i = 0; i < n; ++i) f[i] = fn(x[i], p[i]);
        });
        std::vector<double> hw;
        std::vector<int32_t> leaves;
        h.leaf_weights(q, simpson, hw, leaves);
        const mesh::ReducedMap hm = mesh::reduced_map(h.xs);

        const Tree r = replay(base, s, q, simpson, fn);
        refined += h.size() > base.size() ? 1 : 0;
        const bool ok = r.status == 0 && r.rec == h.rec && same(r.x, h.xs) && same(r.p, h.ps) &&
                        same(r.f, h.fs) && same(r.w, hw) && same(r.ux, hm.ux) &&
                        r.offsets == hm.offsets && r.members == hm.members && r.uof == hm.unique_of;
        if (!ok) {
            ++bad;
            std::printf("trial %d: h %d p %d max %d periodic %d kind %d criteria %d status %d: "
                        "%zu panels vs %d, %zu points vs %zu\n",
                        t, s.initial_height, s.p_height, s.max_height, (int)s.periodic, kind, ne,
                        r.status, r.rec.size() / kI, h.size(), r.x.size(), h.xs.size());
        }
    }
    std::printf("%d of %d trials differ (%d refined beyond the base)\n", bad, trials, refined);
    return bad != 0 || refined < trials / 2;
}
