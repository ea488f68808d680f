// Test infrastructure only: FP64 CPU restatement of one species' phase-space
// panel mesh (src/structure_files/AMRStructure_amr.cpp), its remesh
// interpolation (src/structure_files/AMRStructure_interp.cpp) and the
// reduced-x utilities (src/structure_files/AMRStructure.cpp:147-196).
//
// The panel and point insertion order, neighbour bookkeeping, quadrature
// accumulation order and the leaf-search tie rules follow the reference line
// by line, because they decide the bit-exact tree and the point each new f
// value is interpolated from. Three places where the reference has undefined
// behaviour are given their evident intent:
//  * refine_panels / refine_panels_refine_v read the parent panel through a
//    pointer after panels.push_back may have reallocated the vector
//    (AMRStructure_amr.cpp:474-480, :796-816); here the parent's fields are
//    read before the children are appended.
//  * test_panel's p-derivative criterion leaves abs_dfdps[2..5]
//    uninitialised (AMRStructure_amr.cpp:1044-1053); here they are 0, so
//    only the two values its loop leaves behind take part.
//  * The 9x9 solve uses this file's own partial-pivot LU, not Eigen's
//    PartialPivLU (AMRStructure_interp.cpp:935): equal up to rounding.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <numeric>
#include <stdexcept>

#include "fs_oracle_internal.h"

namespace orc {

// initial_distributions.cpp:12-152 via AMRSimulation_load.cpp:149-221. Type
// 11 (bump-on-tail) has no counterpart in the reference.
double f0_eval(const orc_species_params& sp, double kx, double x, double p) {
    const double pth = sp.pth, pstr = sp.pstr, amp = sp.amp;
    auto maxw = [](double v, double th, double c) {
        return 1.0 / std::sqrt(2.0 * M_PI) / th *
               std::exp(-(v - c) * (v - c) / 2 / th / th);
    };
    switch (sp.ics_type) {
        case 3:  // F0_strong_two_stream: F0_LD(pth, k, amp) * p^2
            return 1.0 / std::sqrt(2.0 * M_PI) / pth *
                   std::exp(-(p - 0.0) * (p - 0.0) / 2 / pth / pth) *
                   (1 + amp * std::cos(kx * x)) * p * p;
        case 4:  // F0_colder_two_stream
            return .5 * (maxw(p - pstr, pth, 0.0) + maxw(p + pstr, pth, 0.0)) *
                   (1 + amp * std::cos(kx * x));
        case 9:   // F0_ion_acoustic_electron
        case 10:  // F0_ion_acoustic_ion (same formula, mass unused)
            return 1.0 / std::sqrt(2.0 * M_PI) / pth *
                   std::exp(-(p - pstr) * (p - pstr) / 2 / pth / pth) *
                   (1 + amp * std::cos(kx * x));
        case 11: {  // bump-on-tail (new)
            const double nb = sp.beam_frac;
            return ((1.0 - nb) * maxw(p, pth, pstr) +
                    nb * maxw(p, sp.beam_vth, sp.beam_v)) *
                   (1 + amp * std::cos(kx * x));
        }
        default:  // 1, 2 and unknown: F0_LD
            return 1.0 / std::sqrt(2.0 * M_PI) / pth *
                   std::exp(-(p - pstr) * (p - pstr) / 2 / pth / pth) *
                   (1 + amp * std::cos(kx * x));
    }
}

void Species::init(const orc_sim_params& sim, const orc_species_params& sp) {
    params = sp;
    q = sp.q;
    qm = sp.q / sp.m;
    x_min = sim.x_min;
    x_max = sim.x_max;
    p_min = sp.p_min;
    p_max = sp.p_max;
    periodic = (sim.bcs == 0);
    simpson = (sim.quadrature == 1);
    initial_height = sp.initial_height;
    p_height = sp.p_height;
    max_height = sp.max_height;
    adaptive = sp.do_adaptively_refine != 0 && sp.n_amr_epsilons > 0;
    amr_eps.assign(sp.amr_epsilons, sp.amr_epsilons + std::max(0, sp.n_amr_epsilons));
    Lx = x_max - x_min;  // AMRStructure.cpp:112-117
    Lp = p_max - p_min;
    npanels_x = int(std::pow(2, initial_height));
    npanels_p = int(std::pow(2, initial_height + p_height));
    initial_dx = Lx / npanels_x;
    initial_dp = Lp / npanels_p;
    kx = 2.0 * M_PI / (sim.x_max - sim.x_min) * sp.normalized_wavenumber;
    generate_mesh([this](double x, double p) { return f0_eval(params, kx, x, p); },
                  adaptive, true);
    f_beyond_boundary = *std::min_element(fs.begin(), fs.end());
}

// AMRStructure_amr.cpp:127-230
void Species::create_prerefined_mesh() {
    if (initial_height < 1) throw std::invalid_argument("height must be greater than 1");
    const double dx = (x_max - x_min) / 4, dp = (p_max - p_min) / 4;
    double gx[5], gp[5];
    for (int i = 0; i < 5; ++i) {
        gx[i] = x_min + i * dx;
        gp[i] = p_min + i * dp;
    }
    panels.clear();
    xs.clear();
    ps.clear();
    fs.clear();
    auto add = [&](double x, double p) {
        xs.push_back(x);
        ps.push_back(p);
    };
    for (int i = 0; i < 5; i += 2)
        for (int j = 0; j < 5; j += 2) add(gx[i], gp[j]);       // points 0..8
    for (int i = 0; i < 2; ++i) {                                // 9..22
        add(gx[2 * i], gp[1]);
        add(gx[2 * i], gp[3]);
        for (int j = 0; j < 5; ++j) add(gx[1 + 2 * i], gp[j]);
    }
    for (int j = 1; j < 5; j += 2) add(gx[4], gp[j]);           // 23, 24
    fs.assign(25, 1.0);

    auto mk = [](int ind, int lev, int par, int wc, int l, int t, int r, int b) {
        Panel pn;
        pn.panel_ind = ind;
        pn.level = lev;
        pn.parent = par;
        pn.which_child = wc;
        pn.left = l;
        pn.top = t;
        pn.right = r;
        pn.bottom = b;
        return pn;
    };
    Panel root;  // Panel::Panel() (Panel.cpp:15-24)
    if (periodic) {
        root.is_left_bdry = root.is_right_bdry = true;
        panels.push_back(root);
        panels.push_back(mk(1, 1, 0, 0, 3, 2, 3, -2));
        panels[1].is_left_bdry = true;
        panels.push_back(mk(2, 1, 0, 1, 4, -2, 4, 1));
        panels[2].is_left_bdry = true;
        panels.push_back(mk(3, 1, 0, 2, 1, 4, 1, -2));
        panels[3].is_right_bdry = true;
        panels.push_back(mk(4, 1, 0, 3, 2, -2, 2, 3));
        panels[4].is_right_bdry = true;
    } else {
        Panel r0 = mk(0, 0, -1, -1, -2, -2, -2, -2);
        panels.push_back(r0);
        panels.push_back(mk(1, 1, 0, 0, -2, 2, 3, -2));
        panels.push_back(mk(2, 1, 0, 1, -2, -2, 4, 1));
        panels.push_back(mk(3, 1, 0, 2, 1, 4, -2, -2));
        panels.push_back(mk(4, 1, 0, 3, 2, -2, -2, 3));
    }
    const int pts[4][9] = {{0, 9, 1, 11, 12, 13, 3, 16, 4},
                           {1, 10, 2, 13, 14, 15, 4, 17, 5},
                           {3, 16, 4, 18, 19, 20, 6, 23, 7},
                           {4, 17, 5, 20, 21, 22, 7, 24, 8}};
    for (int c = 0; c < 4; ++c) {
        std::copy(pts[c], pts[c] + 9, panels[1 + c].pt);
        panels[1 + c].needs_refinement = true;
    }
    panels[0].refined_xp = true;  // set_child_inds_start(1) (Panel.cpp:104-108)
    panels[0].needs_refinement = false;
    panels[0].child_start = 1;
    minimum_unrefined_index = 1;

    auto one = [](double, double) { return 1.0; };
    for (int level = 1; level < initial_height; ++level) {
        const int before = (int)panels.size();
        for (size_t k = minimum_unrefined_index; k < panels.size(); ++k)
            panels[k].needs_refinement = true;
        refine_panels(one);
        minimum_unrefined_index = before;
    }
    for (int level = 0; level < p_height; ++level) {
        const int before = (int)panels.size();
        for (size_t k = minimum_unrefined_index; k < panels.size(); ++k)
            panels[k].needs_refinement = true;
        refine_panels_refine_v(one);
        minimum_unrefined_index = before;
    }
}

// AMRStructure_amr.cpp:506-839 (refinement in x and p: four children).
void Species::refine_panels(const std::function<double(double, double)>& f) {
    std::vector<double> nxs, nps;
    int nv = (int)xs.size();
    const int before = (int)panels.size();
    auto push2 = [&](double xa, double pa, double xb, double pb) {
        nxs.push_back(xa);
        nxs.push_back(xb);
        nps.push_back(pa);
        nps.push_back(pb);
    };
    for (int jj = minimum_unrefined_index; jj < before; ++jj) {
        if (!panels[jj].needs_refinement) continue;
        const Panel P = panels[jj];  // parent fields as they are before refinement
        const double px0 = xs[P.pt[0]], pp0 = ps[P.pt[0]];
        const double dx = xs[P.pt[3]] - px0, dp = ps[P.pt[1]] - pp0;
        const double sdx = 0.5 * dx, sdp = 0.5 * dp;
        const int base = (int)panels.size();
        double sx[5], sp[5];
        for (int i = 0; i < 5; ++i) {
            sx[i] = px0 + sdx * i;
            sp[i] = pp0 + sdp * i;
        }
        int p9, p10, p11, p15, p18, p22, p23, p24;
        int c0L = -1, c1L = -1, c1T = -1, c0B = -1, c2B = -1, c2R = -1, c3R = -1,
            c3T = -1;
        auto flag_parent_nbr = [&](int parent_nbr) {
            Panel& q = panels[parent_nbr];
            if (!q.refined_xp) {
                q.needs_refinement = true;
                need_further_refinement = true;
            }
        };
        // left edge (:566-613)
        if (P.left == -2) {
            c0L = c1L = -2;
            p9 = nv++;
            p10 = nv++;
            push2(sx[0], sp[1], sx[0], sp[3]);
        } else if (P.left == -1) {
            flag_parent_nbr(panels[P.parent].left);
            p9 = nv;
            p10 = p9 + 1;
            nv += 2;
            push2(sx[0], sp[1], sx[0], sp[3]);
        } else {
            const Panel& PL = panels[P.left];
            if (!PL.refined_xp) {
                p9 = nv++;
                p10 = nv++;
                push2(sx[0], sp[1], sx[0], sp[3]);
            } else {
                c0L = PL.child_start + 2;
                c1L = PL.child_start + 3;
                panels[c0L].right = base;
                panels[c1L].right = base + 1;
                if (P.is_left_bdry && periodic) {
                    p9 = nv++;
                    p10 = nv++;
                    push2(sx[0], sp[1], sx[0], sp[3]);
                } else {
                    p9 = panels[c0L].pt[7];
                    p10 = panels[c1L].pt[7];
                }
            }
        }
        // bottom edge (:614-656)
        if (P.bottom == -2) {
            c0B = c2B = -2;
            p11 = nv++;
            p18 = nv++;
            push2(sx[1], sp[0], sx[3], sp[0]);
        } else if (P.bottom == -1) {
            flag_parent_nbr(panels[P.parent].bottom);
            p11 = nv++;
            p18 = nv++;
            push2(sx[1], sp[0], sx[3], sp[0]);
        } else {
            const Panel& PB = panels[P.bottom];
            if (!PB.refined_xp) {
                p11 = nv++;
                p18 = nv++;
                push2(sx[1], sp[0], sx[3], sp[0]);
            } else {
                c0B = PB.child_start + 1;
                c2B = PB.child_start + 3;
                panels[c0B].top = base;
                panels[c2B].top = base + 2;
                p11 = panels[c0B].pt[5];
                p18 = panels[c2B].pt[5];
            }
        }
        // top edge (:658-698)
        if (P.top == -2) {
            c1T = c3T = -2;
            p15 = nv++;
            p22 = nv++;
            push2(sx[1], sp[4], sx[3], sp[4]);
        } else if (P.top == -1) {
            flag_parent_nbr(panels[P.parent].top);
            p15 = nv++;
            p22 = nv++;
            push2(sx[1], sp[4], sx[3], sp[4]);
        } else {
            const Panel& PT = panels[P.top];
            if (!PT.refined_xp) {
                p15 = nv++;
                p22 = nv++;
                push2(sx[1], sp[4], sx[3], sp[4]);
            } else {
                c1T = PT.child_start;
                c3T = PT.child_start + 2;
                panels[c1T].bottom = base + 1;
                panels[c3T].bottom = base + 3;
                p15 = panels[c1T].pt[3];
                p22 = panels[c3T].pt[3];
            }
        }
        // right edge (:700-752)
        if (P.right == -2) {
            c2R = c3R = -2;
            p23 = nv++;
            p24 = nv++;
            push2(sx[4], sp[1], sx[4], sp[3]);
        } else if (P.right == -1) {
            flag_parent_nbr(panels[P.parent].right);
            p23 = nv++;
            p24 = nv++;
            push2(sx[4], sp[1], sx[4], sp[3]);
        } else {
            const Panel& PR = panels[P.right];
            if (!PR.refined_xp) {
                p23 = nv++;
                p24 = nv++;
                push2(sx[4], sp[1], sx[4], sp[3]);
                if (P.right == jj) {
                    c2R = base;
                    c0L = base + 2;
                    c3R = base + 1;
                    c1L = base + 3;
                }
            } else {
                c2R = PR.child_start;
                c3R = PR.child_start + 1;
                panels[c2R].left = base + 2;
                panels[c3R].left = base + 3;
                if (P.is_right_bdry && periodic) {
                    p23 = nv++;
                    p24 = nv++;
                    push2(sx[4], sp[1], sx[4], sp[3]);
                } else {
                    p23 = panels[c2R].pt[1];
                    p24 = panels[c3R].pt[1];
                }
            }
        }
        // interior points (:754-770)
        const int p12 = nv, p16 = p12 + 3, p19 = p16 + 2;
        auto push1 = [&](double xa, double pa) {
            nxs.push_back(xa);
            nps.push_back(pa);
        };
        for (int i = 0; i < 3; ++i) push1(sx[1], sp[1 + i]);
        for (int i = 0; i < 2; ++i) push1(sx[2], sp[1 + 2 * i]);
        for (int i = 0; i < 3; ++i) push1(sx[3], sp[1 + i]);
        nv += 8;

        panels[jj].refined_xp = true;  // set_child_inds_start(base)
        panels[jj].needs_refinement = false;
        panels[jj].child_start = base;
        const int lev = P.level + 1, pind = P.panel_ind;
        auto child = [&](int ind, int wc, std::initializer_list<int> pts, int l,
                         int t, int r, int b, bool lb, bool rb) {
            Panel c;
            c.panel_ind = ind;
            c.level = lev;
            c.parent = pind;
            c.which_child = wc;
            std::copy(pts.begin(), pts.end(), c.pt);
            c.left = l;
            c.top = t;
            c.right = r;
            c.bottom = b;
            c.is_left_bdry = lb;
            c.is_right_bdry = rb;
            c.child_start = -1;
            panels.push_back(c);
        };
        const int* q = P.pt;
        child(base, 0, {q[0], p9, q[1], p11, p12, p12 + 1, q[3], p16, q[4]}, c0L,
              base + 1, base + 2, c0B, P.is_left_bdry, false);
        child(base + 1, 1, {q[1], p10, q[2], p12 + 1, p12 + 2, p15, q[4], p16 + 1, q[5]},
              c1L, c1T, base + 3, base, P.is_left_bdry, false);
        child(base + 2, 2, {q[3], p16, q[4], p18, p19, p19 + 1, q[6], p23, q[7]}, base,
              base + 3, c2R, c2B, false, P.is_right_bdry);
        child(base + 3, 3, {q[4], p16 + 1, q[5], p19 + 1, p19 + 2, p22, q[7], p24, q[8]},
              base + 1, c3T, c3R, base + 2, false, P.is_right_bdry);
    }
    // new f values in insertion order (:821-830)
    std::vector<double> nfs(nxs.size());
    for (size_t i = 0; i < nxs.size(); ++i) nfs[i] = f(nxs[i], nps[i]);
    for (size_t i = 0; i < nxs.size(); ++i) {
        xs.push_back(nxs[i]);
        ps.push_back(nps[i]);
        fs.push_back(nfs[i]);
    }
}

// AMRStructure_amr.cpp:233-504 (refinement in p only: two children).
void Species::refine_panels_refine_v(const std::function<double(double, double)>& f) {
    std::vector<double> nxs, nps;
    int nv = (int)xs.size();
    const int before = (int)panels.size();
    auto push2 = [&](double xa, double pa, double xb, double pb) {
        nxs.push_back(xa);
        nxs.push_back(xb);
        nps.push_back(pa);
        nps.push_back(pb);
    };
    for (int jj = minimum_unrefined_index; jj < before; ++jj) {
        if (!panels[jj].needs_refinement) continue;
        const Panel P = panels[jj];
        const double px0 = xs[P.pt[0]], pp0 = ps[P.pt[0]];
        const double dx = xs[P.pt[3]] - px0, dp = ps[P.pt[1]] - pp0;
        const double sdp = 0.5 * dp, sdx = 0.5 * dx;
        const int base = (int)panels.size();
        double sx[5], sp[5];
        for (int i = 0; i < 5; ++i) {
            sp[i] = pp0 + sdp * i;
            sx[i] = px0 + sdx * i;
        }
        int p9, p10, p13, p14;
        int c0B = -1, c0L = base, c0R = base, c1L = base + 1, c1T = -1, c1R = base + 1;
        auto refined = [](const Panel& q) { return q.refined_xp || q.refined_p; };
        // left (:294-346)
        if (P.left == -2) {
            c0L = c1L = -2;
            p9 = nv++;
            p10 = nv++;
            push2(sx[0], sp[1], sx[0], sp[3]);
        } else if (P.left == -1) {
            Panel& q = panels[panels[P.parent].left];
            if (!refined(q)) {
                q.needs_refinement = true;
                need_further_refinement = true;
            }
            p9 = nv;
            p10 = p9 + 1;
            nv += 2;
            push2(sx[0], sp[1], sx[0], sp[3]);
        } else {
            const Panel& PL = panels[P.left];
            if (!refined(PL)) {
                p9 = nv++;
                p10 = nv++;
                push2(sx[0], sp[1], sx[0], sp[3]);
            } else {
                if (PL.refined_xp) {
                    c0L = PL.child_start + 2;
                    c1L = PL.child_start + 3;
                } else {
                    c0L = PL.child_start;
                    c1L = PL.child_start + 1;
                }
                panels[c0L].right = base;
                panels[c1L].right = base + 1;
                if (P.is_left_bdry && periodic) {
                    p9 = nv++;
                    p10 = nv++;
                    push2(sx[0], sp[1], sx[0], sp[3]);
                } else {
                    p9 = panels[c0L].pt[7];
                    p10 = panels[c1L].pt[7];
                }
            }
        }
        // bottom (:348-369)
        if (P.bottom == -2) {
            c0B = -2;
        } else if (P.bottom == -1) {
            std::fprintf(stderr, "not allowed to refine in v if panel doesn't have bottom neighbor!\n");
        } else {
            Panel& PB = panels[P.bottom];
            if (PB.refined_xp) {
                std::fprintf(stderr, "Shouldn't be allowed to call refine in v if bottom neighbor is refined in x and v!\n");
            } else if (!PB.refined_p) {
                c0B = P.bottom;
                PB.top = base;
            } else {
                c0B = PB.child_start + 1;
                panels[c0B].top = base;
            }
        }
        // top (:371-392)
        if (P.top == -2) {
            c1T = -2;
        } else if (P.top == -1) {
            std::fprintf(stderr, "not allowed to refine in v if panel doesn't have top neighbor!\n");
        } else {
            const Panel& PT = panels[P.top];
            if (PT.refined_xp) {
                std::fprintf(stderr, "Shouldn't be allowed to call refine in v if bottom neighbor is refined in x and v!\n");
            } else if (!PT.refined_p) {
                c1T = -1;
            } else {
                c1T = PT.child_start;
                panels[c1T].bottom = base + 1;
            }
        }
        // right (:394-439)
        if (P.right == -2) {
            c0R = c1R = -2;
            p13 = nv++;
            p14 = nv++;
            push2(sx[4], sp[1], sx[4], sp[3]);
        } else if (P.right == -1) {
            Panel& q = panels[panels[P.parent].right];
            if (!refined(q)) {
                q.needs_refinement = true;
                need_further_refinement = true;
            }
            p13 = nv++;
            p14 = nv++;
            push2(sx[4], sp[1], sx[4], sp[3]);
        } else {
            const Panel& PR = panels[P.right];
            if (!refined(PR)) {
                p13 = nv++;
                p14 = nv++;
                push2(sx[4], sp[1], sx[4], sp[3]);
            } else {
                c0R = PR.child_start;
                c1R = PR.child_start + 1;
                panels[c0R].left = base;
                panels[c1R].left = base + 1;
                if (P.is_right_bdry && periodic) {
                    p13 = nv++;
                    p14 = nv++;
                    push2(sx[4], sp[1], sx[4], sp[3]);
                } else {
                    p13 = panels[c0R].pt[1];
                    p14 = panels[c1R].pt[1];
                }
            }
        }
        // interior (:441-447)
        const int p11 = nv;
        push2(sx[2], sp[1], sx[2], sp[3]);
        nv += 2;

        panels[jj].refined_p = true;  // set_child_inds_start(base, true)
        panels[jj].needs_refinement = false;
        panels[jj].child_start = base;
        const int lev = P.level + 1, pind = P.panel_ind;
        const int* q = P.pt;
        Panel c0, c1;
        c0.panel_ind = base;
        c1.panel_ind = base + 1;
        c0.level = c1.level = lev;
        c0.parent = c1.parent = pind;
        c0.which_child = 0;
        c1.which_child = 1;
        const int a0[9] = {q[0], p9, q[1], q[3], p11, q[4], q[6], p13, q[7]};
        const int a1[9] = {q[1], p10, q[2], q[4], p11 + 1, q[5], q[7], p14, q[8]};
        std::copy(a0, a0 + 9, c0.pt);
        std::copy(a1, a1 + 9, c1.pt);
        c0.left = c0L;
        c0.top = base + 1;
        c0.right = c0R;
        c0.bottom = c0B;
        c1.left = c1L;
        c1.top = c1T;
        c1.right = c1R;
        c1.bottom = base;
        c0.is_left_bdry = c1.is_left_bdry = P.is_left_bdry;
        c0.is_right_bdry = c1.is_right_bdry = P.is_right_bdry;
        c0.child_start = c1.child_start = -1;
        panels.push_back(c0);
        panels.push_back(c1);
    }
    std::vector<double> nfs(nxs.size());
    for (size_t i = 0; i < nxs.size(); ++i) nfs[i] = f(nxs[i], nps[i]);
    for (size_t i = 0; i < nxs.size(); ++i) {
        xs.push_back(nxs[i]);
        ps.push_back(nps[i]);
        fs.push_back(nfs[i]);
    }
}

// AMRStructure_amr.cpp:841-969
void Species::generate_mesh(const std::function<double(double, double)>& f,
                            bool adaptive_now, bool initial) {
    create_prerefined_mesh();
    if (initial) {
        for (size_t i = 0; i < xs.size(); ++i) fs[i] = f0_eval(params, kx, xs[i], ps[i]);
    } else {
        const int nx = 2 * npanels_x + 1, np = 2 * npanels_p + 1;
        interpolate_to_initial_xps(fs, xs, ps, nx, np, nullptr, nullptr);
    }
    if (adaptive_now) {
        for (size_t i = minimum_unrefined_index; i < panels.size(); ++i) test_panel((int)i);
        while (need_further_refinement) {
            need_further_refinement = false;
            refine_panels(f);
            for (size_t i = minimum_unrefined_index; i < panels.size(); ++i)
                if (!panels[i].refined_xp) test_panel((int)i);
        }
    }
    set_leaves_weights();
}

// AMRStructure_amr.cpp:971-1081
void Species::test_panel(int pi) {
    Panel& P = panels[pi];
    double f9[9];
    for (int i = 0; i < 9; ++i) f9[i] = fs[P.pt[i]];
    bool crit = false;
    const size_t ne = amr_eps.size();
    if (ne > 0) {
        double mx = f9[0], mn = f9[0];
        for (int i = 1; i < 9; ++i) {
            if (mx < f9[i]) mx = f9[i];
            if (mn > f9[i]) mn = f9[i];
        }
        crit = crit || (mx - mn > amr_eps[0]);
        if (ne > 1) crit = crit || (mx - mn) / f9[4] > amr_eps[1];
        if (ne > 2) crit = crit || mx / mn > amr_eps[2];
    }
    if (ne > 3) {
        const double dx = xs[P.pt[3]] - xs[P.pt[0]];
        double m = 0;
        for (int i = 0; i < 6; ++i) {
            const double a = std::fabs((f9[3 + i] - f9[i]) / dx);
            if (i == 0 || m < a) m = a;
        }
        crit = crit || (m > amr_eps[3]);
    }
    if (ne > 4) {
        const double dp = ps[P.pt[1]] - ps[P.pt[0]];
        double a[6] = {0, 0, 0, 0, 0, 0};
        for (int jj = 0; jj < 3; ++jj)
            for (int ii = 0; ii < 2; ++ii)
                a[ii] = std::fabs((f9[3 * jj + ii + 1] - f9[3 * jj + ii]) / dp);
        double m = a[0];
        for (int i = 1; i < 6; ++i)
            if (m < a[i]) m = a[i];
        crit = crit || (m > amr_eps[4]);
    }
    if (P.level < max_height && crit) {
        P.needs_refinement = true;
        need_further_refinement = true;
    }
}

// AMRStructure_amr.cpp:1084-1151: per-leaf quadrature weights accumulated in
// depth-first leaf order, then multiplied by f.
void Species::set_leaves_weights() {
    leaf_inds.clear();
    q_ws.assign(xs.size(), 0.0);
    std::function<void(int)> rec = [&](int pi) {
        const Panel& P = panels[pi];
        if (P.refined_p) {
            for (int i = 0; i < 2; ++i) rec(P.child_start + i);
        } else if (P.refined_xp) {
            for (int i = 0; i < 4; ++i) rec(P.child_start + i);
        } else {
            leaf_inds.push_back(pi);
            const double dx = xs[P.pt[3]] - xs[P.pt[0]];
            const double dp = ps[P.pt[1]] - ps[P.pt[0]];
            if (simpson) {
                const double w = q * dx * dp / 9.0;
                const double k[9] = {1.0, 4.0, 1.0, 4.0, 16.0, 4.0, 1.0, 4.0, 1.0};
                for (int i = 0; i < 9; ++i) q_ws[P.pt[i]] += w * k[i];
            } else {
                const double w = q * dx * dp / 4.0;
                const double k[9] = {1.0, 2.0, 1.0, 2.0, 4.0, 2.0, 1.0, 2.0, 1.0};
                for (int i = 0; i < 9; ++i) q_ws[P.pt[i]] += w * k[i];
            }
        }
    };
    rec(0);
    for (size_t i = 0; i < xs.size(); ++i) q_ws[i] *= fs[i];
}

// AMRStructure_amr.cpp:1156-1202 (sqrt_f and do_unshear are false,
// AMRStructure.hpp:100-104).
void Species::remesh() {
    old_panels = panels;
    old_xs = xs;
    old_ps = ps;
    old_fs = fs;
    generate_mesh([this](double x, double p) { return interpolate_from_mesh(x, p); },
                  adaptive, false);
}

// ---------------------------------------------------------------- interp

// AMRStructure_interp.cpp:429-491: move each x by whole periods until it is
// between the advected left and right edges of the root panel.
void Species::shift_xs(std::vector<double>& out, const std::vector<double>& x,
                       const std::vector<double>& p) const {
    const double x_bl = old_xs[0], p_bl = old_ps[0];
    const double x_tl = old_xs[2], p_tl = old_ps[2];
    const double x_br = old_xs[6], p_br = old_ps[6];
    const double x_tr = old_xs[8], p_tr = old_ps[8];
    out.resize(x.size());
    for (size_t i = 0; i < x.size(); ++i) {
        const double pi = p[i];
        double xt = x[i];
        bool left_ok = (x_tl - x_bl) * (pi - p_bl) <= (p_tl - p_bl) * (xt - x_bl);
        int cnt = 0;
        while (!left_ok) {
            xt += Lx;
            left_ok = (x_tl - x_bl) * (pi - p_bl) <= (p_tl - p_bl) * (xt - x_bl);
            if (++cnt > 10) throw std::runtime_error("too many shifts!");
        }
        bool right_ok = (x_tr - x_br) * (pi - p_br) > (p_tr - p_br) * (xt - (x_br));
        cnt = 0;
        while (!right_ok) {
            xt -= Lx;
            right_ok = (x_tr - x_br) * (pi - p_br) > (p_tr - p_br) * (xt - (x_br));
            if (++cnt > 10) throw std::runtime_error("too many shifts!");
        }
        out[i] = xt;
    }
}

// find_leaf_containing_xp_recursively (AMRStructure_interp.cpp:232-426),
// written as a loop; x is shifted in place like the reference's reference
// parameter.
int Species::descend(double& x, double p, bool& beyond, int pi) const {
    for (;;) {
        const Panel& P = old_panels[pi];
        const int cs = P.child_start;
        const double* X = old_xs.data();
        const double* Pp = old_ps.data();
        if (!(P.refined_xp || P.refined_p)) {
            const double x_bl = X[P.pt[0]], p_bl = Pp[P.pt[0]];
            const double x_tl = X[P.pt[2]], p_tl = Pp[P.pt[2]];
            const double x_br = X[P.pt[6]], p_br = Pp[P.pt[6]];
            const double x_tr = X[P.pt[8]], p_tr = Pp[P.pt[8]];
            const bool r = (x_tr - x_br) * (p - p_br) > (p_tr - p_br) * (x - (x_br));
            const bool l = (x_tl - x_bl) * (p - p_bl) <= (p_tl - p_bl) * (x - x_bl);
            const bool t = (x_tr - x_tl) * (p - p_tl) < (p_tr - p_tl) * (x - x_tl);
            const bool b = (x_br - x_bl) * (p - p_bl) >= (p_br - p_bl) * (x - x_bl);
            if ((!l && P.left == -2) || (!r && P.right == -2) ||
                (!t && P.top == -2) || (!b && P.bottom == -2))
                beyond = true;
            return pi;
        }
        const double x_bl = X[P.pt[0]], p_bl = Pp[P.pt[0]];
        const double x_ml = X[P.pt[1]], p_ml = Pp[P.pt[1]];
        const double x_tl = X[P.pt[2]], p_tl = Pp[P.pt[2]];
        const double x_bm = X[P.pt[3]], p_bm = Pp[P.pt[3]];
        const double x_mm = X[P.pt[4]], p_mm = Pp[P.pt[4]];
        const double x_tm = X[P.pt[5]], p_tm = Pp[P.pt[5]];
        const double x_br = X[P.pt[6]], p_br = Pp[P.pt[6]];
        const double x_mr = X[P.pt[7]], p_mr = Pp[P.pt[7]];
        const double x_tr = X[P.pt[8]], p_tr = Pp[P.pt[8]];
        const bool b1 = (x_mm - x_ml) * (p - p_ml) >= (p_mm - p_ml) * (x - x_ml);
        const bool r1 = (x_tm - x_mm) * (p - p_mm) >= (p_tm - p_mm) * (x - x_mm);
        const bool b3 = (x_mr - x_mm) * (p - p_mm) >= (p_mr - p_mm) * (x - x_mm);
        int sub;
        if (b1 && r1) {
            const bool t1 = (x_tm - x_tl) * (p - p_tl) <= (p_tm - p_tl) * (x - x_tl);
            const Panel& c1 = old_panels[cs + 1];
            if (t1 || c1.top < 0) {
                const bool l1 = (x_tl - x_ml) * (p - p_ml) <= (p_tl - p_ml) * (x - x_ml);
                if (l1 || c1.left < 0) {
                    sub = cs + 1;
                } else {
                    sub = c1.left;
                    if (P.is_left_bdry && periodic) x += Lx;
                }
            } else {
                sub = c1.top;
            }
        } else if (b3 && !r1) {
            const bool t3 = (x_tr - x_tm) * (p - p_tm) <= (p_tr - p_tm) * (x - x_tm);
            const Panel& c3 = P.refined_p ? old_panels[cs + 1] : old_panels[cs + 3];
            if (t3 || c3.top < 0) {
                const bool r3 = (x_tr - x_mr) * (p - p_mr) <= (p_tr - p_mr) * (x - x_mr);
                if (!r3 || c3.right < 0) {
                    sub = P.refined_p ? cs + 1 : cs + 3;
                } else {
                    sub = c3.right;
                    if (P.is_right_bdry && periodic) x -= Lx;
                }
            } else {
                sub = c3.top;
            }
        } else {
            const bool r0 = (x_mm - x_bm) * (p - p_bm) >= (p_mm - p_bm) * (x - x_bm);
            if (r0) {
                const bool b0 = (x_bm - x_bl) * (p - p_bl) >= (p_bm - p_bl) * (x - x_bl);
                const Panel& c0 = old_panels[cs];
                if (b0 || c0.bottom < 0) {
                    const bool l0 = (x_ml - x_bl) * (p - p_bl) <= (p_ml - p_bl) * (x - x_bl);
                    if (l0 || c0.left < 0) {
                        sub = cs;
                    } else {
                        sub = c0.left;
                        if (P.is_left_bdry && periodic) x += Lx;
                    }
                } else {
                    sub = c0.bottom;
                }
            } else {
                const bool b2 = (x_br - x_bm) * (p - p_bm) >= (p_br - p_bm) * (x - x_bm);
                const Panel& c2 = P.refined_p ? old_panels[cs] : old_panels[cs + 2];
                if (b2 || c2.bottom < 0) {
                    const bool r2 = (x_mr - x_br) * (p - p_br) <= (p_mr - p_br) * (x - x_br);
                    if (!r2 || c2.right < 0) {
                        sub = P.refined_p ? cs : cs + 2;
                    } else {
                        sub = c2.right;
                        if (P.is_right_bdry && periodic) x -= Lx;
                    }
                } else {
                    sub = c2.bottom;
                }
            }
        }
        pi = sub;
    }
}

// find_leaf_containing_point_from_neighbor (AMRStructure_interp.cpp:12-230)
// as a loop: test the current leaf's four advected edges, step to the
// neighbour across the first failed edge (right, top, bottom, left), stop at
// a leaf already visited. Every visited level that failed an edge with no
// neighbour (-2) marks the point as beyond the boundary.
int Species::walk(double& tx, double tp, bool& beyond, int leaf) const {
    std::vector<int> history(1, leaf);
    int cur = leaf;
    for (;;) {
        if (cur == 0) return 0;
        const Panel& P = old_panels[cur];
        const double x_bl = old_xs[P.pt[0]], p_bl = old_ps[P.pt[0]];
        const double x_tl = old_xs[P.pt[2]], p_tl = old_ps[P.pt[2]];
        const double x_mid = old_xs[P.pt[4]];
        const double x_br = old_xs[P.pt[6]], p_br = old_ps[P.pt[6]];
        const double x_tr = old_xs[P.pt[8]], p_tr = old_ps[P.pt[8]];
        if (tx - x_mid >= Lx / 2 && periodic) tx -= Lx;
        if (tx - x_mid < -Lx / 2 && periodic) tx += Lx;
        const bool r = (x_tr - x_br) * (tp - p_br) > (p_tr - p_br) * (tx - (x_br));
        const bool l = (x_tl - x_bl) * (tp - p_bl) <= (p_tl - p_bl) * (tx - x_bl);
        const bool t = (x_tr - x_tl) * (tp - p_tl) < (p_tr - p_tl) * (tx - x_tl);
        const bool b = (x_br - x_bl) * (tp - p_bl) >= (p_br - p_bl) * (tx - x_bl);
        int nxt = cur;
        if (!r && P.right != -2) {
            if (P.is_right_bdry && periodic) tx -= Lx;
            if (P.right == -1) {
                nxt = old_panels[P.parent].right;
            } else {
                const Panel& q = old_panels[P.right];
                nxt = q.refined_xp ? q.child_start : P.right;
            }
        } else if (!t && P.top != -2) {
            if (P.top == -1) {
                nxt = old_panels[P.parent].top;
            } else {
                const Panel& q = old_panels[P.top];
                nxt = q.refined_xp ? q.child_start : P.top;
            }
        } else if (!b && P.bottom != -2) {
            if (P.bottom == -1) {
                nxt = old_panels[P.parent].bottom;
            } else {
                const Panel& q = old_panels[P.bottom];
                nxt = q.refined_xp ? q.child_start + 1 : P.bottom;
            }
        } else if (!l && P.left != -2) {
            if (P.is_left_bdry && periodic) tx += Lx;
            if (P.left == -1) {
                nxt = old_panels[P.parent].left;
            } else {
                const Panel& q = old_panels[P.left];
                nxt = q.refined_xp ? q.child_start + 2 : P.left;
            }
        }
        if ((!r && P.right == -2) || (!l && P.left == -2) || (!t && P.top == -2) ||
            (!b && P.bottom == -2))
            beyond = true;
        if (std::find(history.begin(), history.end(), nxt) != history.end()) return nxt;
        history.push_back(nxt);
        cur = nxt;
    }
}

// AMRStructure_interp.cpp:493-793
void Species::interpolate_to_initial_xps(std::vector<double>& f_out,
                                         const std::vector<double>& x,
                                         const std::vector<double>& p, int nx,
                                         int np, std::vector<int>* leaf_sorted,
                                         std::vector<int>* perm) const {
    const size_t n = x.size();
    std::vector<double> sx(n);
    if (periodic)
        shift_xs(sx, x, p);
    else
        sx = x;
    std::vector<int> order(n);
    for (size_t i = 0; i < n; ++i) order[i] = (int)i;
    const double thr = initial_dp / 10.0;
    std::sort(order.begin(), order.end(), [&](int a, int b) {
        if (std::fabs(p[a] - p[b]) >= thr) return p[a] < p[b];
        return sx[a] < sx[b];
    });
    std::vector<double> sxs(n), sps(n), sfs(n);
    for (size_t i = 0; i < n; ++i) {
        sxs[i] = sx[order[i]];
        sps[i] = p[order[i]];
    }
    std::vector<int> leaf_of(n, 0);
    bool beyond = false;
    int leaf = descend(sxs[0], sps[0], beyond, 0);
    std::vector<int> first(np);
    for (int r = 0; r < np; ++r) {  // first column (:632-650)
        beyond = false;
        const int k = r * nx;
        leaf = walk(sxs[k], sps[k], beyond, leaf);
        first[r] = leaf;
        leaf_of[k] = beyond ? 0 : leaf;
    }
    for (int r = 0; r < np; ++r) {  // rows (:696-722)
        int k = r * nx;
        int lc = first[r];
        for (int c = 1; c < nx; ++c) {
            beyond = false;
            ++k;
            lc = walk(sxs[k], sps[k], beyond, lc);
            leaf_of[k] = beyond ? 0 : lc;
        }
    }
    std::vector<std::vector<int>> by_leaf(old_panels.size());
    for (size_t i = 0; i < n; ++i) by_leaf[leaf_of[i]].push_back((int)i);
    for (size_t pi = 0; pi < old_panels.size(); ++pi)
        if (!by_leaf[pi].empty())
            interpolate_from_panel_to_points(sfs, sxs, sps, by_leaf[pi], (int)pi);
    f_out.resize(n);
    for (size_t i = 0; i < n; ++i) f_out[order[i]] = sfs[i];
    if (leaf_sorted) *leaf_sorted = leaf_of;
    if (perm) *perm = order;
}

// Forward elimination with partial pivoting and back substitution.
void solve9(const double Ain[81], const double bin[9], double c[9]) {
    double A[81], b[9];
    std::copy(Ain, Ain + 81, A);
    std::copy(bin, bin + 9, b);
    int perm[9];
    for (int i = 0; i < 9; ++i) perm[i] = i;
    for (int k = 0; k < 9; ++k) {
        int piv = k;
        double best = std::fabs(A[k * 9 + k]);
        for (int i = k + 1; i < 9; ++i) {
            const double a = std::fabs(A[i * 9 + k]);
            if (a > best) {
                best = a;
                piv = i;
            }
        }
        if (piv != k) {
            for (int j = 0; j < 9; ++j) std::swap(A[k * 9 + j], A[piv * 9 + j]);
            std::swap(perm[k], perm[piv]);
        }
        const double d = A[k * 9 + k];
        if (d != 0.0)
            for (int i = k + 1; i < 9; ++i) A[i * 9 + k] /= d;
        for (int i = k + 1; i < 9; ++i) {
            const double l = A[i * 9 + k];
            for (int j = k + 1; j < 9; ++j) A[i * 9 + j] -= l * A[k * 9 + j];
        }
    }
    double y[9];
    for (int i = 0; i < 9; ++i) {
        double s = bin[perm[i]];
        for (int j = 0; j < i; ++j) s -= A[i * 9 + j] * y[j];
        y[i] = s;
    }
    for (int i = 8; i >= 0; --i) {
        double s = y[i];
        for (int j = i + 1; j < 9; ++j) s -= A[i * 9 + j] * c[j];
        c[i] = s / A[i * 9 + i];
    }
    (void)b;
}

namespace {
// The monomial matrix of AMRStructure_interp.cpp:923-933 and its solve.
void panel_coefficients(const Species& S, int pi, double c[9], double& x4, double& p4) {
    const Panel& P = S.old_panels[pi];
    double px[9], pp[9], pf[9];
    for (int i = 0; i < 9; ++i) {
        px[i] = S.old_xs[P.pt[i]];
        pp[i] = S.old_ps[P.pt[i]];
        pf[i] = S.old_fs[P.pt[i]];
    }
    double A[81];
    for (int i = 0; i < 9; ++i) {
        const double dx = px[i] - px[4], dp = pp[i] - pp[4];
        double* r = &A[i * 9];
        r[0] = 1;
        r[1] = dx;
        r[2] = dx * dp;
        r[3] = dp;
        r[4] = dx * dx;
        r[5] = dx * dx * dp;
        r[6] = dx * dx * dp * dp;
        r[7] = dx * dp * dp;
        r[8] = dp * dp;
    }
    solve9(A, pf, c);
    x4 = px[4];
    p4 = pp[4];
}
}  // namespace

// AMRStructure_interp.cpp:795-975 (batched evaluation, Dx * c row by row).
void Species::interpolate_from_panel_to_points(std::vector<double>& values,
                                               const std::vector<double>& x,
                                               const std::vector<double>& p,
                                               const std::vector<int>& idx,
                                               int pi) const {
    if (pi == 0) {
        for (int k : idx) values[k] = f_beyond_boundary;
        return;
    }
    double c[9], x4, p4;
    panel_coefficients(*this, pi, c, x4, p4);
    for (int k : idx) {
        const double dx = x[k] - x4, dp = p[k] - p4;
        const double dxsq = dx * dx, dpsq = dp * dp;
        const double d[9] = {1, dx, dx * dp, dp, dxsq, dxsq * dp, dxsq * dpsq, dx * dpsq, dpsq};
        double v = 0.0;
        for (int j = 0; j < 9; ++j) v += d[j] * c[j];
        values[k] = v;
    }
}

// AMRStructure_interp.cpp:977-1092 (single point)
double Species::interpolate_from_panel(double x, double p, int pi) const {
    if (pi == 0) return f_beyond_boundary;
    double c[9], x4, p4;
    panel_coefficients(*this, pi, c, x4, p4);
    const double dx = x - x4, dp = p - p4;
    return c[0] + c[1] * dx + c[2] * dx * dp + c[3] * dp + c[4] * dx * dx +
           c[5] * dx * dx * dp + c[6] * dx * dx * dp * dp + c[7] * dx * dp * dp +
           c[8] * dp * dp;
}

// AMRStructure_interp.cpp:1094-1143
double Species::interpolate_from_mesh(double x, double p) const {
    std::vector<double> xv(1, x), pv(1, p), sv(1, x);
    if (periodic) shift_xs(sv, xv, pv);
    double sx = sv[0];
    bool beyond = false;
    int leaf = descend(sx, p, beyond, 0);
    if (beyond) leaf = 0;
    return interpolate_from_panel(sx, p, leaf);
}

// AMRStructure.cpp:147-181
void Species::get_reduced_xs_ws() {
    const size_t n = xs.size();
    std::vector<size_t> inds(n);
    std::iota(inds.begin(), inds.end(), 0);
    std::sort(inds.begin(), inds.end(), [&](size_t a, size_t b) { return xs[a] < xs[b]; });
    double val = xs[inds[0]];
    reduced_xs.assign(1, val);
    inv_inds_reduced_xs.assign(n, 0);
    size_t u = 0;
    for (size_t i = 1; i < n; ++i) {
        const double v = xs[inds[i]];
        if (v > val) {
            val = v;
            ++u;
            reduced_xs.push_back(v);
        }
        inv_inds_reduced_xs[inds[i]] = u;
    }
    reduced_ws.assign(reduced_xs.size(), 0.0);
    for (size_t i = 0; i < n; ++i) reduced_ws[inv_inds_reduced_xs[i]] += q_ws[i];
}

// AMRStructure.cpp:183-196
void Species::get_reduced_es(const double* reduced_es) {
    es.assign(xs.size(), 0.0);
    for (size_t i = 0; i < inv_inds_reduced_xs.size(); ++i)
        es[i] = reduced_es[inv_inds_reduced_xs[i]];
}

}  // namespace orc
