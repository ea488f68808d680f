// Test infrastructure only: FP64 CPU restatement of the FARRSIGHT step loop
// (src/simulation_files/AMRSimulation_step.cpp, _field.cpp, _remesh.cpp,
// _gather_scatter.cpp) and the C API of the oracle. File output is not part
// of it (the reduced-x output directories are taken to exist, so
// evaluate_field_uniform_grid never takes its early return,
// AMRSimulation_field.cpp:91-94).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <vector>

#include "fs_oracle_internal.h"

namespace orc {

struct Sim {
    orc_sim_params prm{};
    std::vector<std::unique_ptr<Species>> species;
    std::vector<size_t> start, end;
    std::vector<double> xs, ps, q_ws, es;
    std::vector<double> qms, ms;
    int iter = 0;
    double t = 0;
    bool need_gather = true, need_scatter = false;

    double Lx() const { return prm.x_max - prm.x_min; }

    // (*calculate_e) from make_field_return_ptr (AMRSimulation_load.cpp:49-96):
    // periodic decks the direct sum or the treecode; open decks the
    // open-boundary direct sum, or the periodic direct sum when use_treecode
    // is set (:61-73).
    void field_only(std::vector<double>& e, const std::vector<double>& x,
                    const std::vector<double>& w) {
        std::vector<double> xc(x), xc2(x);
        const bool open = prm.bcs != 0;
        if (prm.use_treecode > 0 && !open) {
            if (!treecode(e.data(), xc2.data(), w.data(), (int64_t)xc.size(), Lx(),
                          prm.greens_epsilon, prm.mac, prm.degree, prm.max_source))
                throw std::runtime_error("treecode: more than max_source coincident points");
        } else if (open && prm.use_treecode <= 0) {
            direct_sum_open(e.data(), xc.data(), (int64_t)e.size(), xc2.data(), w.data(),
                            (int64_t)xc.size(), prm.greens_epsilon);
        } else {
            direct_sum(e.data(), xc.data(), (int64_t)e.size(), xc2.data(), w.data(),
                       (int64_t)xc.size(), Lx(), prm.greens_epsilon);
        }
    }

    // (*calculate_e_external) at time tt (AMRSimulation_field.cpp:70-72,
    // 169-171; make_external_field, AMRSimulation_load.cpp:99-136)
    void drive(std::vector<double>& e, const std::vector<double>& x, double tt) {
        if (prm.use_external_field > 0)
            external_add(e.data(), x.data(), (int64_t)e.size(), tt, prm.external_type, prm.Am,
                         prm.k, prm.omega);
    }

    // AMRSimulation_field.cpp:164-173
    void field(std::vector<double>& e, const std::vector<double>& x,
               const std::vector<double>& w, double tt) {
        field_only(e, x, w);
        drive(e, x, tt);
    }

    // AMRSimulation_field.cpp:5-162
    void field_uniform_grid() {
        std::vector<double> rx, rw;
        for (auto& s : species) {
            s->get_reduced_xs_ws();
            rx.insert(rx.end(), s->reduced_xs.begin(), s->reduced_xs.end());
            rw.insert(rw.end(), s->reduced_ws.begin(), s->reduced_ws.end());
        }
        std::vector<double> re(rw.size());
        field(re, rx, rw, t);
        size_t off = 0;
        for (auto& s : species) {
            s->get_reduced_es(re.data() + off);
            off += s->reduced_xs.size();
        }
        need_gather = true;
    }

    // AMRSimulation_gather_scatter.cpp:3-32
    void gather() {
        xs.clear();
        ps.clear();
        q_ws.clear();
        es.clear();
        start.clear();
        end.clear();
        size_t n = 0;
        for (auto& s : species) {
            start.push_back(n);
            n += s->xs.size();
            end.push_back(n);
        }
        for (auto& s : species) {
            xs.insert(xs.end(), s->xs.begin(), s->xs.end());
            ps.insert(ps.end(), s->ps.begin(), s->ps.end());
            q_ws.insert(q_ws.end(), s->q_ws.begin(), s->q_ws.end());
            es.insert(es.end(), s->es.begin(), s->es.end());
        }
        need_gather = false;
    }

    // AMRSimulation_gather_scatter.cpp:34-51
    void scatter(bool send_e) {
        size_t off = 0;
        for (auto& s : species) {
            for (size_t i = 0; i < s->xs.size(); ++i) {
                s->xs[i] = xs[off + i];
                s->ps[i] = ps[off + i];
                s->q_ws[i] = q_ws[off + i];
                if (send_e) s->es[i] = es[off + i];
            }
            off += s->xs.size();
        }
        need_scatter = false;
    }

    // AMRSimulation_step.cpp:122-263 (get_4th_e = false)
    void rk4() {
        const double dt = prm.dt;
        const size_t N = xs.size();
        std::vector<double> xk = xs;
        std::vector<double> p1 = ps, p2(N), p3(N), p4(N);
        std::vector<double> ef1 = es, ef2(N), ef3(N), ef4(N);
        for (size_t sp = 0; sp < species.size(); ++sp)
            for (size_t i = start[sp]; i < end[sp]; ++i) {
                ef1[i] *= qms[sp];
                p2[i] = ps[i] + 0.5 * dt * ef1[i];
                xk[i] = xs[i] + 0.5 * dt * p1[i];
            }
        const double t2 = t + dt * 0.5, t3 = t2, t4 = t + dt;   // :170, 201, 224
        field(ef2, xk, q_ws, t2);
        for (size_t sp = 0; sp < species.size(); ++sp)
            for (size_t i = start[sp]; i < end[sp]; ++i) {
                ef2[i] *= qms[sp];
                p3[i] = ps[i] + 0.5 * dt * ef2[i];
                xk[i] = xs[i] + 0.5 * dt * p2[i];
            }
        field(ef3, xk, q_ws, t3);
        for (size_t sp = 0; sp < species.size(); ++sp)
            for (size_t i = start[sp]; i < end[sp]; ++i) {
                ef3[i] *= qms[sp];
                p4[i] = ps[i] + dt * ef3[i];
                xk[i] = xs[i] + dt * p3[i];
            }
        field(ef4, xk, q_ws, t4);
        for (size_t sp = 0; sp < species.size(); ++sp)
            for (size_t i = start[sp]; i < end[sp]; ++i) ef4[i] *= qms[sp];
        for (size_t i = 0; i < N; ++i) {
            xs[i] += dt / 6.0 * (p1[i] + 2 * p2[i] + 2 * p3[i] + p4[i]);
            ps[i] += dt / 6.0 * (ef1[i] + 2 * ef2[i] + 2 * ef3[i] + ef4[i]);
        }
        need_scatter = true;
    }

    // AMRSimulation_step.cpp:5-99 (file output omitted)
    void step() {
        if (need_gather) gather();
        rk4();
        iter += 1;
        t += prm.dt;
        if (iter % prm.remesh_period == 0) {
            if (need_scatter) scatter(false);           // AMRSimulation_remesh.cpp:4-8
            for (auto& s : species) s->remesh();
            field_uniform_grid();
            need_gather = true;
        } else {
            field(es, xs, q_ws, t);
        }
    }

    // The species' current E at its own points.
    void species_es(int i, double* out) {
        if (need_gather) {
            std::copy(species[i]->es.begin(), species[i]->es.end(), out);
        } else {
            std::copy(es.begin() + start[i], es.begin() + end[i], out);
        }
    }
};

}  // namespace orc

using orc::Sim;
using orc::Species;

struct orc_species : Species {};
struct orc_sim : Sim {};

extern "C" {

orc_species* orc_species_create(const orc_sim_params* sim, const orc_species_params* sp) {
    try {
        auto* s = new orc_species();
        s->init(*sim, *sp);
        s->old_panels = s->panels;
        s->old_xs = s->xs;
        s->old_ps = s->ps;
        s->old_fs = s->fs;
        return s;
    } catch (...) {
        return nullptr;
    }
}
void orc_species_destroy(orc_species* s) { delete s; }
int64_t orc_species_num_points(const orc_species* s) { return (int64_t)s->xs.size(); }
int64_t orc_species_num_panels(const orc_species* s) { return (int64_t)s->panels.size(); }
void orc_species_get_points(const orc_species* s, double* xs, double* ps, double* fs,
                            double* q_ws) {
    const size_t n = s->xs.size();
    if (xs) std::copy(s->xs.begin(), s->xs.end(), xs);
    if (ps) std::copy(s->ps.begin(), s->ps.end(), ps);
    if (fs) std::copy(s->fs.begin(), s->fs.end(), fs);
    if (q_ws) {
        if (s->q_ws.size() == n)
            std::copy(s->q_ws.begin(), s->q_ws.end(), q_ws);
        else
            std::fill(q_ws, q_ws + n, 0.0);
    }
}
int orc_species_set_points(orc_species* s, const double* xs, const double* ps) {
    std::copy(xs, xs + s->xs.size(), s->xs.begin());
    std::copy(ps, ps + s->ps.size(), s->ps.begin());
    return 0;
}
void orc_species_get_panels(const orc_species* s, int32_t* rec) {
    for (size_t i = 0; i < s->panels.size(); ++i) {
        const orc::Panel& P = s->panels[i];
        int32_t* r = rec + i * ORC_PANEL_INTS;
        r[0] = P.panel_ind;
        for (int k = 0; k < 9; ++k) r[1 + k] = P.pt[k];
        r[10] = P.level;
        r[11] = P.parent;
        r[12] = P.which_child;
        r[13] = P.left;
        r[14] = P.top;
        r[15] = P.right;
        r[16] = P.bottom;
        r[17] = (P.is_left_bdry ? 1 : 0) | (P.is_right_bdry ? 2 : 0) |
                (P.needs_refinement ? 4 : 0) | (P.refined_xp ? 8 : 0) |
                (P.refined_p ? 16 : 0);
        r[18] = P.child_start;
        r[19] = 0;
    }
}
int64_t orc_species_num_leaves(const orc_species* s) { return (int64_t)s->leaf_inds.size(); }
void orc_species_get_leaves(const orc_species* s, int32_t* out) {
    std::copy(s->leaf_inds.begin(), s->leaf_inds.end(), out);
}
double orc_species_f_beyond_boundary(const orc_species* s) { return s->f_beyond_boundary; }
int orc_species_remesh(orc_species* s) {
    try {
        s->remesh();
    } catch (...) {
        return -1;
    }
    return 0;
}
int orc_species_copy_to_old(orc_species* s) {
    s->old_panels = s->panels;
    s->old_xs = s->xs;
    s->old_ps = s->ps;
    s->old_fs = s->fs;
    return 0;
}
int orc_species_interp_from_mesh(orc_species* s, const double* x, const double* p,
                                 int64_t n, double* f_out) {
    try {
        for (int64_t i = 0; i < n; ++i) f_out[i] = s->interpolate_from_mesh(x[i], p[i]);
    } catch (...) {
        return -1;
    }
    return 0;
}
int orc_species_set_fs(orc_species* s, const double* fs) {
    std::copy(fs, fs + s->fs.size(), s->fs.begin());
    return 0;
}
int orc_species_descend_points(orc_species* s, const double* x, const double* p, int64_t n,
                               int32_t* leaf_out) {
    try {
        for (int64_t i = 0; i < n; ++i) {
            // the first half of Species::interpolate_from_mesh
            std::vector<double> xv(1, x[i]), pv(1, p[i]), sv(1, x[i]);
            if (s->periodic) s->shift_xs(sv, xv, pv);
            double sx = sv[0];
            bool beyond = false;
            int leaf = s->descend(sx, p[i], beyond, 0);
            leaf_out[i] = beyond ? 0 : leaf;
        }
    } catch (...) {
        return -1;
    }
    return 0;
}
int orc_species_interp_to_grid(orc_species* s, const double* xs, const double* ps, int nx,
                               int np, double* f_out, int32_t* leaf_sorted,
                               int32_t* sort_perm) {
    const size_t n = (size_t)nx * (size_t)np;
    std::vector<double> x(xs, xs + n), p(ps, ps + n), f;
    std::vector<int> leaf, perm;
    try {
        s->interpolate_to_initial_xps(f, x, p, nx, np, &leaf, &perm);
    } catch (...) {
        return -1;
    }
    std::copy(f.begin(), f.end(), f_out);
    if (leaf_sorted) std::copy(leaf.begin(), leaf.end(), leaf_sorted);
    if (sort_perm) std::copy(perm.begin(), perm.end(), sort_perm);
    return 0;
}

orc_sim* orc_sim_create(const orc_sim_params* prm, const orc_species_params* sp, int nsp) {
    try {
        auto* s = new orc_sim();
        s->prm = *prm;
        for (int i = 0; i < nsp; ++i) {
            auto sp_i = std::make_unique<Species>();
            sp_i->init(*prm, sp[i]);
            s->qms.push_back(sp_i->qm);
            s->ms.push_back(sp_i->q / sp_i->qm);
            s->species.push_back(std::move(sp_i));
        }
        s->iter = 0;
        s->t = 0;
        s->field_uniform_grid();  // AMRSimulation.cpp:84-89
        return s;
    } catch (...) {
        return nullptr;
    }
}
void orc_sim_destroy(orc_sim* s) { delete s; }
int orc_sim_step(orc_sim* s) {
    try {
        s->step();
    } catch (...) {
        return -1;
    }
    return 0;
}
int orc_sim_num_species(const orc_sim* s) { return (int)s->species.size(); }
orc_species* orc_sim_species(orc_sim* s, int i) {
    if (s->need_scatter) s->scatter(false);
    return static_cast<orc_species*>(s->species[i].get());
}
void orc_sim_get_es(orc_sim* s, int i, double* es) { s->species_es(i, es); }
int orc_sim_iter(const orc_sim* s) { return s->iter; }

// standard_interface/diagnostic_plot.py:116-143 for one species, in point
// order (np.dot and np.sum order aside).
void orc_diagnostics(int64_t n_, const double* xs, const double* ps, const double* fs,
                     const double* q_ws, const double* e, double q, double qm, double L,
                     double* out) {
    const size_t n = (size_t)n_;
    double charge = 0, mom = 0, kin = 0, ent = 0, l1 = 0, l2 = 0, neg = 0, absw = 0;
    double fmin = n ? fs[0] : 0.0, fmax = n ? fs[0] : 0.0;
    int64_t nneg = 0;
    for (size_t k = 0; k < n; ++k) {
        charge += q_ws[k];
        mom += ps[k] * q_ws[k];
        kin += ps[k] * ps[k] * q_ws[k];
        if (fs[k] > 0) ent += fs[k] * std::log10(fs[k]) * q_ws[k];
        l1 += std::fabs(q_ws[k]);
        l2 += q_ws[k] * fs[k];
        if (fs[k] < 0) {
            neg += q_ws[k];
            ++nneg;
        }
        absw += std::fabs(q_ws[k]);
        if (fs[k] < fmin) fmin = fs[k];
        if (fs[k] > fmax) fmax = fs[k];
    }
    out[0] = charge;
    out[1] = 1.0 / q * mom;
    out[2] = 1.0 / 2.0 / qm * kin;
    // electric energy over the unique x in [0, Lx): first occurrence of each x
    std::vector<size_t> idx(n);
    for (size_t k = 0; k < n; ++k) idx[k] = k;
    std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return xs[a] < xs[b]; });
    std::vector<double> ux, ue;
    for (size_t k = 0; k < n; ++k) {
        const double x = xs[idx[k]];
        if (k > 0 && x == xs[idx[k - 1]]) continue;  // np.unique keeps the first
        if (x >= 0 && x < L) {
            ux.push_back(x);
            ue.push_back(e[idx[k]]);
        }
    }
    const size_t m = ux.size();
    double pot = 0;
    if (m >= 3) {
        for (size_t k = 0; k < m; ++k) {
            double dxk;
            if (k == 0)
                dxk = .5 * (ux[1] + L - ux[m - 1]);
            else if (k == m - 1)
                dxk = .5 * (ux[0] + L - ux[m - 2]);
            else
                dxk = .5 * (ux[k + 1] - ux[k - 1]);
            pot += dxk * ue[k] * ue[k];
        }
        pot *= 0.5;
    }
    out[3] = pot;
    out[4] = 1.0 / q * ent;
    out[5] = 1.0 / std::fabs(q) * l1;
    out[6] = 1.0 / q * l2;
    out[7] = fmin;
    out[8] = fmax;
    out[9] = 1.0 / q * neg;
    out[10] = (double)nneg;
    out[11] = n ? (double)nneg / (double)n : 0.0;
    out[12] = absw > 0 ? std::fabs(neg) / absw : 0.0;
    out[13] = (double)n;
}

void orc_sim_diagnostics_full(orc_sim* s, int i, double* out) {
    if (s->need_scatter) s->scatter(false);
    Species& sp = *s->species[i];
    const size_t n = sp.xs.size();
    std::vector<double> e(n);
    s->species_es(i, e.data());
    orc_diagnostics((int64_t)n, sp.xs.data(), sp.ps.data(), sp.fs.data(), sp.q_ws.data(),
                    e.data(), sp.q, sp.qm, s->Lx(), out);
}

void orc_sim_diagnostics(orc_sim* s, int i, double out[4]) {
    double full[ORC_DIAG_COUNT];
    orc_sim_diagnostics_full(s, i, full);
    for (int k = 0; k < 4; ++k) out[k] = full[k];
}

}  // extern "C"
