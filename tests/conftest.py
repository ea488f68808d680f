import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")
    config.addinivalue_line("markers", "slow: long CPU-side oracle runs")


def load_golden(name="field_golden.npz"):
    """{case: {key: array}} from a committed fixture file."""
    z = np.load(os.path.join(GOLDEN, name), allow_pickle=False)
    out = {}
    for k in z.files:
        case, key = k.split("__")
        out.setdefault(case, {})[key] = z[k]
    return out


@pytest.fixture(scope="session")
def golden():
    return load_golden()


@pytest.fixture(scope="session")
def gpu_ctx():
    import multi_species_amd as ms

    ctx = ms.Context(0)
    yield ctx
    ctx.close()
