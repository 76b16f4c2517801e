"""Build libimexlb.so (HIP for gfx950) and the C++ driver, in-tree.

    python -m imexlb_amd.build            # library + driver
    python -m imexlb_amd.build --lib      # library only

Objects go to imexlb_amd/_build/, the library to imexlb_amd/lib/, the
driver to imexlb_amd/bin/.  Every call recompiles.
"""
import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIBDIR = os.path.join(PKG, "lib")
BINDIR = os.path.join(PKG, "bin")
LIB = os.path.join(LIBDIR, "libimexlb.so")
DRIVER = os.path.join(BINDIR, "imexlb")

ARCH = os.environ.get("IMEXLB_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")

COMMON = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
          # packed FP32 instructions stay off (hipcc flag; the host pass warns)
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-I", os.path.join(ROOT, "include"), "-Wall", "-Wno-unused-function"]
# bit-for-bit FP64 with the oracle: no contraction of multiply-add
EXACT = ["-ffp-contract=off"]

SOURCES = [
    ("fused_f64.hip", EXACT),
    ("fused_lp.hip", []),
    ("aux.hip", EXACT),
    ("phased.hip", EXACT),
    ("ensemble.hip", []),
    ("noalloc.hip", EXACT),
    ("capi.cpp", EXACT),
    ("host.cpp", EXACT),
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l
                    and "not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + out)
    return out


def build_lib(jobs=8, verbose=False, out=None, defines=()):
    """Compile every HIP/C++ source and link the library (to `out`, default
    imexlb_amd/lib/libimexlb.so); `defines` are extra -D flags."""
    build = BUILD if out is None else out + ".objs"
    os.makedirs(build, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    cmds = []
    objs = []
    for src, extra in SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(build, src + ".o")
        lang = ["-x", "hip"] if src.endswith(".cpp") else []
        cmds.append([HIPCC] + COMMON + extra + [f"-D{d}" for d in defines] + lang + ["-c", path, "-o", obj])
        objs.append(obj)
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        outs = list(ex.map(_run, cmds))
    if verbose:
        for o in outs:
            if o.strip():
                print(o)
    target = LIB if out is None else out
    tmp = target + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp] + objs)
    os.replace(tmp, target)
    return target


def build_driver():
    os.makedirs(BINDIR, exist_ok=True)
    src = os.path.join(CSRC, "host", "main.cpp")
    if not os.path.exists(src):
        return None
    tmp = DRIVER + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"),
          "-I", os.path.join(CSRC, "host"), src, os.path.join(CSRC, "host", "lbm_facade.cpp"),
          "-o", tmp, "-L", LIBDIR, "-limexlb", f"-Wl,-rpath,$ORIGIN/../lib"])
    os.replace(tmp, DRIVER)
    return DRIVER


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib", action="store_true", help="library only")
    ap.add_argument("-j", type=int, default=8)
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args(argv)
    print(build_lib(a.j, a.v))
    if not a.lib:
        d = build_driver()
        if d:
            print(d)


if __name__ == "__main__":
    sys.exit(main())
