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
