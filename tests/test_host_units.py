"""Host-side logic of the kernels' launch: which address mode a lattice of a
given population stride takes, the cells per thread a requested launch
shape runs with, and the chunk size of a placement window.
tests/host_units.cpp includes the library's headers and checks the limits
(2^28 cells per population for the grouped bases, 4 GiB per population for
per-population bases) at and one past each boundary; it is built with
hipcc's host pass and runs on the CPU."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="needs hipcc")
def test_address_mode_and_window_chunk_limits(tmp_path):
    exe = tmp_path / "host_units"
    cmd = [HIPCC, "-std=c++17", "--cuda-host-only", "-x", "hip", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "host_units.cpp"), "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr + r.stdout
    assert r.stdout.split() == ["ok", "fused_addr_mode", "ok", "fused_cells_per_thread", "ok", "window_chunk"]
