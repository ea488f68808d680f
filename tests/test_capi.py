"""The C-ABI library loads and exports every symbol include/farrsight_hip.h
declares; host-side argument checks fail cleanly (no GPU needed)."""
import ctypes
import subprocess

import pytest

import multi_species_amd as ms


def test_library_exports_every_declared_symbol():
    declared = ms.declared_symbols()
    assert "fs_emq_periodic_f64" in declared and "fs_ctx_create" in declared
    out = subprocess.run(["nm", "-D", "--defined-only", ms.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = [s for s in declared if s not in exported]
    assert not missing, missing
    lib = ms.lib()
    for s in declared:
        assert hasattr(lib, s)


def test_argument_errors_without_device_work():
    lib = ms.lib()
    assert lib.fs_ctx_create(None, 0) == -1
    assert b"null" in lib.fs_last_error()
    assert lib.fs_emq_periodic_f64(None, None, None, 0, None, None, 0, 1.0, 0.1) == -1
    assert lib.fs_emq_set_tuning(None, 1, 1) == -1
    assert lib.fs_emq_open_f64(None, None, None, 0, None, None, 0, 0.1) == -1
    assert lib.fs_emq_open_f64_dev(None, None, None, 0, None, None, 0, 0.1, None) == -1
    assert lib.fs_external_field_add(None, None, None, 0, 0.0, 0, 0.4, 0.26, 0.37) == -1
    assert lib.fs_external_field_add_dev(None, None, None, 0, 0.0, 0, 0.4, 0.26, 0.37,
                                         None) == -1
    assert b"null" in lib.fs_last_error()
    assert lib.fs_emq_set_mixed_min_pairs(None, 0) == -1


def test_mixed_min_pairs_default_matches_the_header():
    import re
    with open(ms.HEADER) as f:
        m = re.search(r"#define FS_MIXED_MIN_PAIRS_DEFAULT \(INT64_C\(1\) << (\d+)\)", f.read())
    assert m is not None
    assert ms.MIXED_MIN_PAIRS_DEFAULT == 1 << int(m.group(1))


def test_product_refuses_to_run_without_its_library(monkeypatch, tmp_path):
    monkeypatch.setattr(ms, "LIB_PATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(ms, "_lib", None)
    with pytest.raises(ms.FarrsightError):
        ms.lib()
