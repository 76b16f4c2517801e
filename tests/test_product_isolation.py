"""The product path never reaches the oracle.

The oracle (oracle/) is the checker of the HIP library. Only the tests, the
smoke check and bench.py's CPU-baseline leg may use it. This file pins that:
- no module of the package imports it, at any depth of its code;
- bench.py imports it only inside `cpu_baseline`;
- the built library neither depends on nor names an oracle object, and
  opens no library at run time (no dlopen).
"""
import ast
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "imexlb_amd")
LIB = os.path.join(PKG, "lib", "libimexlb.so")


def _oracle_imports(tree):
    """(line, enclosing function or None) of every import of the oracle."""
    hits = []

    def visit(node, fn):
        for ch in ast.iter_child_nodes(node):
            name = ch.name if isinstance(ch, (ast.FunctionDef, ast.AsyncFunctionDef)) else fn
            mods = []
            if isinstance(ch, ast.Import):
                mods = [a.name for a in ch.names]
            elif isinstance(ch, ast.ImportFrom):
                mods = [ch.module or ""]
            if any(m == "oracle" or m.startswith("oracle.") for m in mods):
                hits.append((ch.lineno, fn))
            if isinstance(ch, ast.Call) and getattr(ch.func, "id", "") == "__import__":
                if ch.args and isinstance(ch.args[0], ast.Constant) and str(ch.args[0].value).startswith("oracle"):
                    hits.append((ch.lineno, fn))
            visit(ch, name)

    visit(tree, None)
    return hits


def test_package_never_imports_the_oracle():
    found = []
    for d, _, files in os.walk(PKG):
        for f in files:
            if f.endswith(".py"):
                p = os.path.join(d, f)
                with open(p) as fh:
                    src = fh.read()
                found += [(p, ln) for ln, _ in _oracle_imports(ast.parse(src))]
                assert "oracle/_ref" not in src and "importlib" not in src, p
    assert not found, found


def test_bench_imports_the_oracle_only_for_the_cpu_baseline():
    with open(os.path.join(ROOT, "bench.py")) as fh:
        hits = _oracle_imports(ast.parse(fh.read()))
    assert hits, "bench.py's CPU baseline is expected to time the oracle"
    assert all(fn == "cpu_baseline" for _, fn in hits), hits


def test_library_has_no_oracle_dependency():
    if not os.path.exists(LIB):
        pytest.skip("libimexlb.so not built")
    dyn = subprocess.run(["readelf", "-d", "-W", LIB], stdout=subprocess.PIPE, text=True, check=True).stdout
    needed = [l.split("[", 1)[1].rstrip("]") for l in dyn.splitlines() if "(NEEDED)" in l]
    assert needed and not any("oracle" in n or "_ref" in n for n in needed), needed
    syms = subprocess.run(["nm", "-D", "--undefined-only", LIB], stdout=subprocess.PIPE, text=True,
                          check=True).stdout
    assert "dlopen" not in syms and "oracle" not in syms
    with open(LIB, "rb") as fh:
        blob = fh.read()
    assert b"imexlb_oracle" not in blob and b"oracle/_ref" not in blob
