"""The device refinement passes replayed on the host against the sequential
host builder (tests/amr_replay.cpp): random meshes, every launch's threads
in ascending, descending and shuffled order, buffers filled with garbage;
panel records, points, f, weights and the reduced-x map bit for bit.

A mistake this catches: any pass step that reads what another thread of the
same launch writes, reads memory no step wrote, or numbers panels or points
differently from the sequential pass (each changes records or points on some
of the meshes)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.fixture(scope="module")
def replay_bin(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    out = str(tmp_path_factory.mktemp("amr") / "amr_replay")
    cmd = [HIPCC, "-O2", "-std=c++17", "-x", "hip", "--offload-arch=gfx950",
           "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "amr_replay.cpp"),
           os.path.join(ROOT, "multi_species_amd", "csrc", "fs_mesh.cpp"), "-o", out]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr[-3000:]
    return out


@pytest.mark.parametrize("order,seed", [(0, 12345), (1, 777), (2, 4242)])
def test_passes_equal_the_sequential_builder(replay_bin, order, seed):
    res = subprocess.run([replay_bin, "150", str(order), str(seed)], capture_output=True,
                         text=True, timeout=600)
    assert res.returncode == 0, res.stdout[-3000:] + res.stderr[-2000:]
    assert res.stdout.startswith("0 of 150 trials differ"), res.stdout
