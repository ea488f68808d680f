"""The oracle's treecode tree (orc_treecode_tree, a restatement of
src/FieldStructure.cpp:506-702) checked against what a bisection treecode must
satisfy, independently of how it is built:

* leaves partition the tree order in contiguous ranges of fewer than
  max_source points, in increasing x (a wrong split side or a dropped point
  breaks it);
* inside a leaf the points keep input order (the stable partition; an
  unstable swap breaks it);
* for every leaf its far and near entries cover every source exactly once
  (a wrong recursion, a double count or a missed sibling breaks it);
* near entries are leaves, far entries are not;
* far entries satisfy the MAC (R > (r_b + r_p)/theta with the periodic
  distance), recomputed here from the node geometry derived by bisection of
  the padded box.
The evaluation counts of the same build equal the compiled reference's
(tests/test_oracle_field.py)."""
import numpy as np
import pytest

import oracle

L0 = 4 * np.pi


def _boxes(x, perm, leaves, nodes, L, n):
    """Interval of each node (first, last, depth) by bisecting the padded box
    along the path that contains its points."""
    xs = np.fmod(x, L)[perm]
    lo, hi = xs.min() - 0.1234, xs.max() + 0.001
    out = {}
    for f, l, d in nodes:
        a, b = lo, hi
        v = xs[f]
        for _ in range(d):
            m = 0.5 * (a + b)
            if v <= m:
                b = m
            else:
                a = m
        out[(f, l, d)] = (a, b)
    return out


@pytest.mark.parametrize("seed,n,mac,ms_", [(0, 4000, 0.7, 50), (1, 20000, 0.5, 300),
                                             (2, 3000, 0.9, 5), (3, 777, 0.7, 800)])
def test_tree_invariants(seed, n, mac, ms_):
    rng = np.random.default_rng(seed)
    x = np.concatenate([rng.uniform(-1, L0 + 1, n // 2), rng.normal(6.0, 0.2, n - n // 2)])
    t = oracle.treecode_tree(x, L0, mac, ms_)
    nn, nl, nf, nr = t["stats"]
    lv, perm = t["leaves"], t["perm"]
    assert lv.shape == (nl, 3)
    assert np.array_equal(np.sort(perm), np.arange(n))
    assert lv[0, 0] == 0 and lv[-1, 1] == n - 1 and np.all(lv[1:, 0] == lv[:-1, 1] + 1)
    assert np.all(lv[:, 1] - lv[:, 0] + 1 < ms_) or nl == 1
    xs = np.fmod(x, L0)[perm]
    for a, b in zip(lv[:-1], lv[1:]):
        assert xs[a[0]:a[1] + 1].max() <= xs[b[0]:b[1] + 1].min()
    for a in lv:
        assert np.all(np.diff(perm[a[0]:a[1] + 1]) > 0)
    leafset = {tuple(r) for r in lv}
    allnodes = {tuple(r) for r in t["far"]} | {tuple(r) for r in t["near"]} | leafset
    box = _boxes(x, perm, leaves=lv, nodes=allnodes, L=L0, n=n)
    for i, b in enumerate(lv):
        cover = np.zeros(n, np.int32)
        far = t["far"][t["far_off"][i]:t["far_off"][i + 1]]
        near = t["near"][t["near_off"][i]:t["near_off"][i + 1]]
        for f, l, d in far:
            cover[f:l + 1] += 1
            assert (f, l, d) not in leafset
        for f, l, d in near:
            cover[f:l + 1] += 1
            assert (f, l, d) in leafset
        assert np.all(cover == 1), f"leaf {i}: sources not covered exactly once"
        ba, bb = box[tuple(b)]
        rb = 0.5 * (bb - ba)
        for f, l, d in far:
            pa, pb = box[(f, l, d)]
            r = abs(0.5 * (ba + bb) - 0.5 * (pa + pb))
            R = abs(min(r, L0 - r))  # squared in the reference: points below 0 stretch the box
            assert (rb + 0.5 * (pb - pa)) < mac * R * (1 + 1e-12)


def test_coincident_points_fail():
    x = np.concatenate([np.full(40, 2.5), np.linspace(0.1, 12.0, 200)])
    assert oracle.treecode_tree(x, L0, 0.7, 30) is None
    assert oracle.treecode_tree(x, L0, 0.7, 41) is not None
