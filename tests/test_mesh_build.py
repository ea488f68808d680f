"""The library's host-side panel-tree builder (fs_mesh_build) against the
oracle's restatement of AMRStructure_amr.cpp: panel records, point
coordinates, f0 values and quadrature weights bit for bit, on uniform,
p-refined, Simpson and adaptively refined meshes; plus closed-form counts."""
import numpy as np
import pytest

import multi_species_amd as ms
import oracle

L0 = 4 * np.pi


def decks():
    out = []
    for h, ph in [(1, 0), (2, 0), (3, 1), (4, 1), (5, 1), (3, 2), (4, 0)]:
        out.append((f"uniform_h{h}_p{ph}", dict(initial_height=h, p_height=ph), {}))
    out.append(("simpson_h4", dict(initial_height=4, p_height=1), dict(quadrature=1)))
    out.append(("amr_df", dict(initial_height=4, p_height=1, do_adaptively_refine=True,
                               amr_epsilons=[1e-3], max_height=6, amp=0.3, pth=0.6), {}))
    out.append(("amr_two_crit", dict(initial_height=3, p_height=1, do_adaptively_refine=True,
                                     amr_epsilons=[2e-3, 0.5], max_height=7, amp=0.3, pth=0.6),
                {}))
    out.append(("amr_all_five", dict(initial_height=3, p_height=1, do_adaptively_refine=True,
                                     amr_epsilons=[0.05, 5.0, 1e9, 0.05, 0.05], max_height=6,
                                     amp=0.2, pth=0.8), {}))
    out.append(("two_stream_amr", dict(initial_height=4, p_height=1, do_adaptively_refine=True,
                                       amr_epsilons=[1e-3], max_height=7, ics_type=4, pth=0.5,
                                       pstr=2.4, amp=1e-3, normalized_wavenumber=1.0), {}))
    out.append(("bump_on_tail", dict(initial_height=4, p_height=1, ics_type=11, pth=1.0,
                                     beam_frac=0.1, beam_v=4.5, beam_vth=0.5, amp=0.04), {}))
    out.append(("strong_two_stream", dict(initial_height=3, p_height=1, ics_type=3, amp=0.05), {}))
    return out


@pytest.mark.parametrize("name,sp,top", decks(), ids=[d[0] for d in decks()])
def test_tree_matches_oracle_bitwise(name, sp, top):
    base = dict(q=-1.0, m=1.0, p_min=-6.0, p_max=6.0, ics_type=1, amp=0.01, pth=1.0)
    base.update(sp)
    deck = oracle.Deck(x_min=0.0, x_max=L0, quadrature=top.get("quadrature", 0),
                       species=[oracle.Species(**base)])
    o = oracle.OracleSpecies(deck, deck.species[0])
    ox, op, of, ow = o.points()
    orec = o.panels()
    rec, xs, ps, fs, qw = ms.mesh_build(ms.sim_params(x_min=0.0, x_max=L0,
                                                      quadrature=top.get("quadrature", 0)),
                                        ms.species_params(**base))
    assert rec.shape == orec.shape
    assert np.array_equal(rec, orec)
    for a, b in [(xs, ox), (ps, op), (fs, of), (qw, ow)]:
        assert np.array_equal(a, b)


@pytest.mark.parametrize("h,ph", [(1, 0), (3, 1), (5, 1), (4, 2)])
def test_closed_form_counts(h, ph):
    rec, xs, *_ = ms.mesh_build(ms.sim_params(x_min=0.0, x_max=L0),
                                ms.species_params(initial_height=h, p_height=ph, p_min=-6,
                                                  p_max=6))
    assert xs.size == (2 ** (h + 1) + 1) * (2 ** (h + ph + 1) + 1)
    npanels = 1 + sum(4 ** l for l in range(1, h + 1)) + sum(4 ** h * 2 ** l for l in range(1, ph + 1))
    assert rec.shape[0] == npanels


def test_quadrature_weights_integrate_area():
    for quad in (0, 1):
        rec, xs, ps, fs, qw = ms.mesh_build(
            ms.sim_params(x_min=0.0, x_max=L0, quadrature=quad),
            ms.species_params(initial_height=4, p_height=1, p_min=-6, p_max=6, q=2.0, amp=0.0,
                              pth=1e9))
        # f is (nearly) constant = 1/(sqrt(2 pi) pth): weights integrate the box area
        w = qw / fs
        assert abs(w.sum() - 2.0 * L0 * 12.0) < 1e-10 * 2.0 * L0 * 12.0


def test_neighbour_symmetry_and_shared_edges():
    rec, xs, ps, *_ = ms.mesh_build(ms.sim_params(x_min=0.0, x_max=L0),
                                    ms.species_params(initial_height=4, p_height=1, p_min=-6,
                                                      p_max=6))
    leaf = (rec[:, 17] & (8 | 16)) == 0
    for i in np.nonzero(leaf)[0]:
        r = rec[i, 15]
        assert r >= 0 and rec[r, 13] == i
        t = rec[i, 14]
        if t >= 0:
            assert rec[t, 16] == i
            assert list(rec[i, [3, 6, 9]]) == list(rec[t, [1, 4, 7]])
        if not (rec[i, 17] & 2):  # not on the periodic seam
            assert list(rec[i, [7, 8, 9]]) == list(rec[r, [1, 2, 3]])
