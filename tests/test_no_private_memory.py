"""Every kernel of the library runs without private (scratch) memory.

Private memory on gfx950 lives in HBM: a register array indexed at run time,
or a register spill, becomes loads and stores that the lattice kernels would
pay on every cell (a runtime index into the 27 populations of a cell once
added 64 bytes of writes per cell to the faithful fused pass). This test
compiles each HIP source for the device only, with the library's own flags,
and reads the kernel descriptors the assembler would emit: every kernel must
have `.amdhsa_private_segment_fixed_size 0` and no dynamic stack.
"""
import os
import re
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

from imexlb_amd import build as B

HIP_SOURCES = [(s, x) for s, x in B.SOURCES if s.endswith(".hip")]


def _device_asm(src, extra, outdir):
    out = os.path.join(outdir, src + ".s")
    cmd = [B.HIPCC] + B.COMMON + extra + ["--cuda-device-only", "-S",
                                         os.path.join(B.CSRC, src), "-o", out]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    assert r.returncode == 0, r.stdout[-4000:]
    with open(out) as fh:
        return src, fh.read()


def _descriptors(asm):
    """(kernel name, private segment bytes, dynamic stack flag) per kernel."""
    out = []
    for m in re.finditer(r"\.amdhsa_kernel (\S+)\n(.*?)\.end_amdhsa_kernel", asm, re.S):
        body = m.group(2)
        priv = int(re.search(r"\.amdhsa_private_segment_fixed_size (\d+)", body).group(1))
        dyn = re.search(r"\.amdhsa_uses_dynamic_stack (\d+)", body)
        out.append((m.group(1), priv, int(dyn.group(1)) if dyn else 0))
    return out


@pytest.mark.skipif(not (os.path.exists(B.HIPCC) or shutil.which(B.HIPCC)), reason="hipcc absent")
def test_kernels_use_no_private_memory(tmp_path):
    with ThreadPoolExecutor(max_workers=min(8, len(HIP_SOURCES))) as ex:
        asms = list(ex.map(lambda sx: _device_asm(sx[0], sx[1], str(tmp_path)), HIP_SOURCES))
    bad = []
    nkern = 0
    for src, asm in asms:
        ds = _descriptors(asm)  # none in a unit that only launches others' kernels
        nkern += len(ds)
        bad += [(src, n, p, dyn) for n, p, dyn in ds if p != 0 or dyn != 0]
    assert not bad, "kernels with private memory (source, kernel, bytes, dynamic stack):\n" + \
        "\n".join(map(str, bad))
    assert nkern > 50
