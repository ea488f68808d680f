"""GPU treecode (fs_emq_treecode_f64) against the reference's own treecode
(golden vectors of src/FieldStructure.cpp's E_MQ_Treecode) and the oracle.

Bar: within 1e-12 normwise of the reference on the same tree parameters
(the tree, lists and Chebyshev points are the reference's; only rounding
differs), the same kernel-evaluation counts as the oracle, and the
single-leaf identity treecode = sqrt(1 + 4 delta^2) direct sum."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

L0 = 4 * np.pi


def rel(a, b):
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


def test_matches_reference_golden(gpu_ctx, golden):
    n = 0
    for name, c in golden.items():
        if not name.startswith("tc_"):
            continue
        L, eps, mac, deg, ms_ = c["params"]
        e = gpu_ctx.emq_treecode_f64(c["s"], c["w"], L, eps, mac, int(deg), int(ms_))
        print(name, "gpu-vs-ref", rel(e, c["es"]))
        assert rel(e, c["es"]) <= 1e-12, name
        n += 1
    assert n >= 4


@pytest.mark.parametrize("n,mac,deg,ms_", [(5000, 0.7, 8, 200), (20000, 0.5, 6, 500),
                                           (3000, 0.9, 3, 37), (100000, 0.7, 8, 2000)])
def test_matches_oracle_and_counts(gpu_ctx, n, mac, deg, ms_):
    rng = np.random.default_rng(n)
    x = rng.uniform(-1.0, L0 + 1.0, n)
    w = rng.normal(size=n) * 1e-3
    e, (far, near) = gpu_ctx.emq_treecode_f64(x, w, L0, 0.05, mac, deg, ms_, return_evals=True)
    eo = oracle.treecode(x, w, L0, 0.05, mac, deg, ms_)
    st = oracle.treecode_stats()
    assert (far, near) == (st["far_evals"], st["near_evals"])
    assert rel(e, eo) <= 1e-12


def test_single_leaf_identity(gpu_ctx):
    rng = np.random.default_rng(5)
    n = 1500
    x = rng.uniform(0, L0, n)
    w = rng.normal(size=n) * 1e-3
    e = gpu_ctx.emq_treecode_f64(x, w, L0, 0.05, 0.7, 8, n + 1)
    d = oracle.direct_sum_ext(x, x, w, L0, 0.05)
    assert rel(e, np.sqrt(1 + 4 * (0.05 / L0) ** 2) * d) <= 1e-14


def _tree_cases():
    rng = np.random.default_rng(11)
    two = np.concatenate([rng.normal(3.0, 0.05, 6000), rng.normal(9.0, 0.4, 4000)])
    dup = np.repeat(rng.uniform(0, L0, 300), 7)          # 7 coincident points per coordinate
    return [
        ("uniform", rng.uniform(0, L0, 5000), 0.7, 100),
        ("outside_box", rng.uniform(-1.0, L0 + 1.0, 20000), 0.5, 500),
        ("two_clusters", two, 0.7, 200),
        ("duplicates", dup, 0.7, 8),
        ("root_leaf", rng.uniform(0, L0, 99), 0.7, 100),
        ("root_splits", rng.uniform(0, L0, 100), 0.7, 100),
        ("one_point", np.array([1.25]), 0.7, 2),
        ("c3_size", rng.uniform(0, 10 * np.pi, 324000), 0.7, 2000),
        ("small_leaves", rng.uniform(0, L0, 30000), 0.9, 3),
    ]


@pytest.mark.parametrize("case", _tree_cases(), ids=lambda c: c[0])
def test_device_tree_equals_oracle_tree(gpu_ctx, case):
    """The device-built tree (sorted ranges, levels, MAC walk) is the
    reference's: same nodes, leaves, tree order and interaction lists in the
    same order as the oracle's recursive stable-partition build."""
    name, x, mac, ms_ = case
    L = 10 * np.pi if name == "c3_size" else L0
    print(name, x.size, flush=True)
    got = gpu_ctx.treecode_tree(x, L, mac, 8, ms_)
    want = oracle.treecode_tree(x, L, mac, ms_)
    assert got["stats"] == want["stats"], name
    for k in ("perm", "leaves", "far_off", "far", "near_off", "near"):
        assert np.array_equal(got[k], want[k]), f"{name}: {k}"


def test_device_tree_rejects_coincident_points(gpu_ctx):
    """More than max_source points on one coordinate cannot be separated:
    FS_E_RANGE (the reference recurses without end)."""
    x = np.concatenate([np.full(40, 2.5), np.linspace(0.1, 12.0, 200)])
    assert oracle.treecode_tree(x, L0, 0.7, 30) is None
    with pytest.raises(Exception, match="coincident"):
        gpu_ctx.treecode_tree(x, L0, 0.7, 8, 30)
    with pytest.raises(Exception, match="coincident"):
        gpu_ctx.emq_treecode_f64(x, np.ones_like(x), L0, 0.05, 0.7, 8, 30)
