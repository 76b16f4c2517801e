"""The C-ABI from plain C: tests/capi_consumer.c is compiled with gcc as
C99 (-Wall -Wextra -pedantic -Werror) against include/imexlb.h, linked with
libimexlb.so and run.  It needs no GPU: it calls only the entry points that
check configurations, read input.in, lay out ranks and write Tecplot files."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_plain_c_consumer(tmp_path):
    from imexlb_amd import LIB_PATH
    exe = tmp_path / "consumer"
    libdir = os.path.dirname(LIB_PATH)
    cmd = ["gcc", "-std=c99", "-Wall", "-Wextra", "-pedantic", "-Werror", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "capi_consumer.c"), "-o", str(exe), "-L", libdir, "-limexlb",
           f"-Wl,-rpath,{libdir}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    (tmp_path / "input.in").write_text("1.0  0.1  10\n0.1  20  10\n24 20 16\n")
    r = subprocess.run([str(exe), "input.in", "output3.plt"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0, r.stderr + r.stdout
    assert r.stdout.count("ok") == 4
