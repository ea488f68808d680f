"""The reference's on-disk state (AMRStructure_io.cpp:3-101,
AMRStructure_restart.cpp:33-98), written by the library's host writer and
read here with numpy as the reference's own diagnostics script reads them
(np.fromfile, standard_interface/diagnostic_plot.py:106-113): raw doubles per
point, 9 int32 per leaf in depth-first order, and the full tree as an int64
count then 80-byte records laid out as the reference's PanelDumpRecord. No GPU
needed. Mistakes caught: a field swapped or missing in the record, padding
misplaced, leaves in the wrong order, a flag bit mapped to the wrong byte."""
import os

import numpy as np
import pytest

import multi_species_amd as ms

DUMP = np.dtype([("panel_ind", "<i4"), ("level", "<i4"), ("parent", "<i4"), ("which", "<i4"),
                 ("pts", "<i4", 9), ("left", "<i4"), ("top", "<i4"), ("right", "<i4"),
                 ("bottom", "<i4"), ("child", "<i4"), ("lb", "i1"), ("rb", "i1"),
                 ("rxp", "i1"), ("rp", "i1"), ("nr", "i1"), ("pad", "i1", 3)])


def deck(adaptive):
    sim = ms.sim_params(x_min=0.0, x_max=10 * np.pi, quadrature=0, greens_epsilon=0.05, dt=0.05)
    sp = ms.species_params(q=-1.0, m=1.0, p_min=-6.0, p_max=6.0, initial_height=3, p_height=1,
                           max_height=6 if adaptive else 4, do_adaptively_refine=adaptive,
                           amr_epsilons=[2e-3] if adaptive else [], ics_type=4,
                           normalized_wavenumber=1.0, amp=1e-3, pth=0.5, pstr=2.4)
    return sim, sp


def dfs_leaves(rec):
    out, stack = [], [0]
    while stack:
        p = stack.pop()
        fl = rec[p, 17]
        nch = 2 if fl & 16 else 4 if fl & 8 else 0
        if nch == 0:
            out.append(p)
        else:
            stack.extend(rec[p, 18] + c for c in range(nch - 1, -1, -1))
    return np.array(out)


@pytest.mark.parametrize("adaptive", [False, True])
def test_files_have_the_reference_layout(tmp_path, adaptive):
    assert DUMP.itemsize == 80
    sim, sp = deck(adaptive)
    rec, xs, ps, fs, qw = ms.mesh_build(sim, sp)
    es = np.sin(xs)
    d = str(tmp_path)
    ms.write_species_output(d, "electrons", 7, xs, ps, fs, rec, q_ws=qw, es=es)
    base = os.path.join(d, "simulation_output", "electrons")
    for sub, arr in (("xs", xs), ("ps", ps), ("fs", fs), ("qws", qw), ("es", es)):
        got = np.fromfile(os.path.join(base, sub, f"{sub}_7"))
        assert np.array_equal(got, arr), sub
    lp = np.fromfile(os.path.join(base, "panels", "leaf_point_inds_7"), dtype="<i4").reshape(-1, 9)
    assert np.array_equal(lp, rec[dfs_leaves(rec), 1:10])
    raw = open(os.path.join(base, "panels", "full_tree_7"), "rb").read()
    n = np.frombuffer(raw[:8], dtype="<i8")[0]
    assert n == rec.shape[0] and len(raw) == 8 + 80 * n
    t = np.frombuffer(raw[8:], dtype=DUMP)
    assert np.array_equal(t["panel_ind"], rec[:, 0])
    assert np.array_equal(t["pts"], rec[:, 1:10])
    for f, c in (("level", 10), ("parent", 11), ("which", 12), ("left", 13), ("top", 14),
                 ("right", 15), ("bottom", 16), ("child", 18)):
        assert np.array_equal(t[f], rec[:, c]), f
    for f, bit in (("lb", 1), ("rb", 2), ("nr", 4), ("rxp", 8), ("rp", 16)):
        assert np.array_equal(t[f], ((rec[:, 17] & bit) != 0).astype(np.int8)), f
    assert not t["pad"].any()
    # the reference's diagnostics read these files: charge from qws
    assert np.sum(np.fromfile(os.path.join(base, "qws", "qws_7"))) == pytest.approx(np.sum(qw))
    # and the library reads them back exactly
    x2, p2, f2, r2 = ms.read_species_output(d, "electrons", 7)
    assert np.array_equal(x2, xs) and np.array_equal(p2, ps) and np.array_equal(f2, fs)
    assert np.array_equal(r2, rec)


def test_damaged_files_are_refused(tmp_path):
    sim, sp = deck(False)
    rec, xs, ps, fs, qw = ms.mesh_build(sim, sp)
    d = str(tmp_path)
    ms.write_species_output(d, "e", 1, xs, ps, fs, rec)
    tree = os.path.join(d, "simulation_output", "e", "panels", "full_tree_1")
    raw = open(tree, "rb").read()
    open(tree, "wb").write(raw[:-80])  # one record short
    with pytest.raises(ms.FarrsightError, match="bad header or size"):
        ms.read_species_output(d, "e", 1)
    open(tree, "wb").write(raw)
    open(os.path.join(d, "simulation_output", "e", "ps", "ps_1"), "wb").write(ps[:-1].tobytes())
    with pytest.raises(ms.FarrsightError, match="mismatched"):
        ms.read_species_output(d, "e", 1)
    with pytest.raises(ms.FarrsightError):
        ms.read_species_output(d, "e", 2)  # no such step
