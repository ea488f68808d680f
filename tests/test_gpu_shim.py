"""The ElectricField binding on the GPU: E_MQ_DirectSum_HIP,
E_MQ_DirectSum_openbcs_HIP and E_MQ_Treecode_HIP made and called through the
reference's virtual ElectricField::operator() (the binding compiled against
inc/FieldStructure.hpp, the base class from the reference's compiled
FieldStructure.cpp), against the reference's golden outputs: direct sums
within 1e-12 normwise, the treecode within 1e-12, the open-boundary sum within
1e-12 of max|sum| beyond two roundings of -x; and against the same call on
the reference's own E_MQ_DirectSum object on fresh inputs."""
import numpy as np
import pytest

import shim_util
from multi_species_amd import decks

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not shim_util.available(),
                                 reason="binding or reference build absent")]


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def test_hip_classes_through_the_virtual_call(golden):
    for name, c in golden.items():
        if name.startswith("ds_"):
            L, eps = c["params"]
            f = shim_util.Field(shim_util.KIND_DIRECT_HIP, L, eps)
            es = f(c["t"], c["s"], c["w"])
            f.close()
            print(name, rel(es, c["es"]), flush=True)
            assert rel(es, c["es"]) <= 1e-12, name
        elif name.startswith("tc_"):
            L, eps, mac, deg, msrc = c["params"]
            f = shim_util.Field(shim_util.KIND_TREECODE_HIP, L, eps, mac, int(deg), int(msrc))
            es = f(np.zeros(0), c["s"], c["w"])
            f.close()
            print(name, rel(es, c["es"]), flush=True)
            assert rel(es, c["es"]) <= 1e-12, name


def test_open_class_through_the_virtual_call():
    from conftest import load_golden

    for name, c in load_golden("open_golden.npz").items():
        if name.startswith("op_"):
            f = shim_util.Field(shim_util.KIND_OPEN_HIP, 0.0, float(c["params"][0]))
            es = f(c["t"], c["s"], c["w"])
            f.close()
            bound = 1e-12 * np.max(np.abs(c["es"] + c["t"])) + 4.5e-16 * np.abs(c["t"])
            assert np.all(np.abs(es - c["es"]) <= bound), name


@pytest.mark.parametrize("n", [3000, 20000])
def test_hip_and_reference_objects_agree(n):
    x, w = decks.field_deck(n, seed=n)
    L = 4 * np.pi
    hip = shim_util.Field(shim_util.KIND_DIRECT_HIP, L, 0.05)
    ref = shim_util.Field(shim_util.KIND_DIRECT_REF, L, 0.05)
    e_hip = hip(x, x, w)  # two copies of the positions, as AMRSimulation passes them
    e_ref = ref(x, x, w)
    hip.close()
    ref.close()
    assert rel(e_hip, e_ref) <= 1e-12
