"""CPU tests of the C-ABI library: it loads, exports every declared entry
point, and its host-only pieces (input.in parser, rank grid, configuration
checks) behave as the reference's.  No compute call needs a GPU here."""
import ctypes
import os
import re

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def header_symbols():
    text = open(os.path.join(ROOT, "include", "imexlb.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(imexlb_[a-z_0-9]+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    import imexlb_amd
    from imexlb_amd import _lib
    L = imexlb_amd.lib()
    syms = header_symbols()
    assert len(syms) >= 25
    for s in syms:
        assert hasattr(L, s), s
    # the Python binding types every declared entry point
    bound = {name for name, _, _ in _lib.SIGNATURES}
    assert set(syms) <= bound, set(syms) - bound


def test_system_reads_reference_inputs():
    from imexlb_amd import System
    t = System(os.path.join(HERE, "golden", "input_tgv.in"), "tgv")
    assert (t.rho0, t.R, t.Re, t.u0) == (1.0, 0.1, 1600.0, 0.05)
    assert (t.sx, t.sy, t.sz) == (256, 256, 256)
    assert (t.Time, t.inter) == (100, 10)  # tgv/src/System.cpp:17-18
    assert t.tau == 0.05 * 256 / 1600.0 / (1.0 / 3.0) / 2 / 3.1415926
    assert t.miu == 1.0 * 0.05 * 256 / 2 / 3.1415926 / 1600.0
    t2 = System(os.path.join(HERE, "golden", "input_tgv.in"), "tgv", keep_hardcoded=False)
    assert (t2.Time, t2.inter) == (20, 10)
    c = System(os.path.join(HERE, "golden", "input_cavity.in"), "cavity")
    assert c.tau == 0.1 * 0.1 / 10.0 / (1.0 / 3.0)
    assert (c.Time, c.inter, c.sx) == (100, 10, 100)
    s = System(os.path.join(HERE, "golden", "input_sphere.in"), "sphere")
    assert s.miu == 1.0 * 0.1 * 8.0 / 100.0
    assert s.Time == int((2 * 8.0) ** 2 / s.miu) and s.inter == 1  # sphere/src/System.cpp:21-22


def test_system_rejects_missing_or_short_input(tmp_path):
    from imexlb_amd import ImexlbError, System
    with pytest.raises(ImexlbError):
        System(str(tmp_path / "nope.in"))
    p = tmp_path / "short.in"
    p.write_text("1.0 0.1 100\n0.05 10\n")
    with pytest.raises(ImexlbError):
        System(str(p))


def test_comm_layout_one_rank_periodic_and_walled():
    from imexlb_amd import comm_layout
    a = comm_layout(1, 0, True, 64, 48, 32)
    assert (a["rx"], a["ry"], a["rz"]) == (1, 1, 1)
    assert a["neighbours"] == [0] * 26  # every neighbour is the rank itself
    assert a["l_l"] == [64, 48, 32] and a["lo"] == [0, 0, 0] and a["hi"] == [63, 47, 31]
    b = comm_layout(1, 0, False, 64, 48, 32)
    assert b["neighbours"] == [-1] * 26


@pytest.mark.parametrize("nranks", [2, 4, 6, 8, 12, 27])
def test_comm_layout_grid(nranks):
    from imexlb_amd import comm_layout
    g = 96
    lays = [comm_layout(nranks, r, True, g, g, g) for r in range(nranks)]
    rx, ry, rz = lays[0]["rx"], lays[0]["ry"], lays[0]["rz"]
    assert rx * ry * rz == nranks
    for L in lays:
        me = L["me"]
        assert me == L["px"] + rx * (L["pz"] + rz * L["py"])  # tgv/src/lbm.hpp:69-71
        left, right, down, up, front, back = L["neighbours"][:6]
        # periodic: the right neighbour's left neighbour is me, etc.
        assert lays[right]["neighbours"][0] == me and lays[left]["neighbours"][1] == me
        assert lays[up]["neighbours"][2] == me and lays[down]["neighbours"][3] == me
        assert lays[back]["neighbours"][4] == me and lays[front]["neighbours"][5] == me
    # extents along x tile the global extent
    row = [L for L in lays if L["py"] == 0 and L["pz"] == 0]
    assert sum(L["l_l"][0] for L in row) == g
    assert sorted((L["lo"][0], L["hi"][0]) for L in row)[0][0] == 0


def test_create_rejects_bad_configs_before_touching_a_device():
    import imexlb_amd as im
    from imexlb_amd import ImexlbError
    bad = [dict(sx=1), dict(tau=-0.6), dict(dtype="f32", engine="phased"),
           dict(case="cavity", mode="faithful", engine="fused", storage="aa"),
           dict(case="cavity", mode="faithful", engine="fused", storage="pull", dtype="f32"),
           dict(case="sphere", mode="faithful", engine="fused", storage="aa"),
           dict(case="sphere", mode="config", storage="aa", sphere=(20, 8, 8, 3)),
           dict(case="sphere", mode="config", sphere=(2, 8, 8, 3)), dict(block=100)]
    for kw in bad:
        args = dict(sx=32, sy=16, sz=16, tau=0.1)
        args.update(kw)
        sx, sy, sz, tau = args.pop("sx"), args.pop("sy"), args.pop("sz"), args.pop("tau")
        with pytest.raises(ImexlbError) as e:
            im.LBM(sx, sy, sz, tau, **args)
        assert e.value.status in (-1, -4), (kw, e.value.status)


def test_strerror_covers_every_status():
    import imexlb_amd as im
    L = im.lib()
    for s in range(0, -8, -1):
        assert L.imexlb_strerror(s).decode() not in ("", "unknown status")


def test_tecplot_writer_layout(tmp_path):
    """imexlb_tecplot_write produces the reference's #!TDV112 file
    (tgv/src/lbm.cpp:421-665): 332-byte header, then x y z u v w p."""
    from imexlb_amd import tecplot_write
    rng = np.random.default_rng(5)
    nz, ny, nx = 3, 4, 5
    p, u, v, w = (rng.standard_normal((nz, ny, nx)) for _ in range(4))
    path = tmp_path / "output7.plt"
    tecplot_write(path, 7.0, p, u, v, w)
    raw = path.read_bytes()
    assert len(raw) == 332 + 7 * nx * ny * nz * 8
    assert raw[:8] == b"#!TDV112"
    ints = np.frombuffer(raw[8:80], dtype="<i4")
    assert ints[:4].tolist() == [1, 0, 0, 7]
    assert ints[4:18].tolist() == [120, 0, 121, 0, 122, 0, 117, 0, 118, 0, 119, 0, 112, 0]
    assert np.frombuffer(raw[80:84], "<f4")[0] == 299.0
    assert np.frombuffer(raw[84:120], "<i4").tolist() == [90, 79, 78, 69, 32, 48, 48, 49, 0]
    assert np.frombuffer(raw[120:124], "<i4")[0] == -1
    assert np.frombuffer(raw[124:132], "<f8")[0] == 7.0
    assert np.frombuffer(raw[132:172], "<i4").tolist() == [0, -1, 0, 0, 0, 0, nx, ny, nz, 0]
    assert np.frombuffer(raw[172:180], "<f4").tolist() == [357.0, 299.0]
    assert np.frombuffer(raw[180:220], "<i4").tolist() == [2] * 7 + [0, 0, -1]
    mm = np.frombuffer(raw[220:332], "<f8")
    assert mm[:6].tolist() == [0, 1, 0, 1, 0, 1]
    assert mm[6:].tolist() == [u.min(), u.max(), v.min(), v.max(), w.min(), w.max(), p.min(), p.max()]
    data = np.frombuffer(raw[332:], "<f8").reshape(7, nz, ny, nx)
    X, Y, Z = np.meshgrid(np.arange(nx) / (nx - 1), np.arange(ny) / (ny - 1), np.arange(nz) / (nz - 1), indexing="ij")
    assert np.array_equal(data[0], X.transpose(2, 1, 0))
    assert np.array_equal(data[1], Y.transpose(2, 1, 0))
    assert np.array_equal(data[2], Z.transpose(2, 1, 0))
    for k, a in enumerate((u, v, w, p)):
        assert np.array_equal(data[3 + k], a)
