"""The ElectricField binding (multi_species_amd/shim), CPU side: it is built
against the reference's own header, loads beside the reference's compiled
FieldStructure.cpp, and the reference's own classes made and called through
it, via the virtual operator(), reproduce the reference's golden outputs bit
for bit (the call path the HIP classes take on the GPU)."""
import numpy as np
import pytest

import shim_util

pytestmark = pytest.mark.skipif(not shim_util.available(),
                                reason="binding or reference build absent (needs the "
                                       "reference tree at build time)")


def test_reference_direct_sum_through_the_virtual_call(golden):
    for name, c in golden.items():
        if not name.startswith("ds_"):
            continue
        L, eps = c["params"]
        f = shim_util.Field(shim_util.KIND_DIRECT_REF, L, eps)
        es = f(c["t"], c["s"], c["w"])
        f.close()
        assert np.array_equal(es, c["es"]), name


def test_reference_open_sum_through_the_virtual_call():
    from conftest import load_golden

    for name, c in load_golden("open_golden.npz").items():
        if name.startswith("op_"):
            f = shim_util.Field(shim_util.KIND_OPEN_REF, 0.0, float(c["params"][0]))
            es = f(c["t"], c["s"], c["w"])
            f.close()
            assert np.array_equal(es, c["es"]), name


def test_unknown_kind_is_refused():
    with pytest.raises(RuntimeError):
        shim_util.Field(99)
