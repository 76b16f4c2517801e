"""The C++ driver (imexlb_amd/bin/imexlb: the reference's main loop over the
C-ABI) runs each case and agrees with the same steps taken through Python."""
import os
import re
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "imexlb_amd", "bin", "imexlb")


def oracle_diagnostics(tmp_path, case, mode, steps, keep_hardcoded=False):
    """KE and mass of the oracle after `steps` steps of the input.in in
    tmp_path, summed exactly (math.fsum) over the oracle's interior moments,
    plus the sum of |3p| that scales the mass's rounding."""
    import math

    from imexlb_amd import System
    from oracle.oracle import CONFIG, FAITHFUL, LDC, SPHERE, TGV, Oracle
    cid = {"tgv": TGV, "cavity": LDC, "sphere": SPHERE}[case]
    s = System(str(tmp_path / "input.in"), case, keep_hardcoded=keep_hardcoded)
    o = Oracle(cid, s.sx, s.sy, s.sz, tau0=s.tau, rho0=s.rho0, u0=s.u0,
               mode=FAITHFUL if mode == "faithful" else CONFIG)
    o.Initialize()
    o.run(steps)
    p, u, v, w = (a.ravel() for a in o.moments())
    ke = 0.5 * math.fsum(u * u + v * v + w * w)
    mass = math.fsum(3.0 * p)
    return ke, mass, math.fsum(np.abs(3.0 * p))


def assert_diag(got, want):
    """The library's two-pass FP64 reduction against the exactly rounded sum
    of the same fields: within 1e-14 of the sum's scale."""
    (ke, mass), (ke_o, mass_o, mass_scale) = got, want
    assert abs(ke - ke_o) <= 1e-14 * ke_o, (ke, ke_o)
    assert abs(mass - mass_o) <= 1e-14 * mass_scale, (mass, mass_o)


def run(tmp_path, text, args):
    (tmp_path / "input.in").write_text(text)
    r = subprocess.run([DRIVER] + args, cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr + r.stdout
    m = re.search(r"Final KE = (\S+)\s+mass = (\S+)", r.stdout)
    assert m, r.stdout
    assert "Avg time spent in Collision:" in r.stdout and "Avg. Time per time-step:" in r.stdout
    return float(m.group(1)), float(m.group(2)), r.stdout


@pytest.mark.parametrize("storage", ["pull", "aa"])
def test_driver_tgv_matches_python(tmp_path, storage):
    from imexlb_amd import LBM, System
    text = "1.0  0.1  100\n0.05  30  10\n32 32 32\n"
    ke, mass, out = run(tmp_path, text, ["--case", "tgv", "--storage", storage, "--honour-input"])
    assert "Warm-Up" in out and "Reset and Run Main-Loop" in out
    s = System(str(tmp_path / "input.in"), "tgv", keep_hardcoded=False)
    g = LBM(s.sx, s.sy, s.sz, s.tau, s.rho0, s.u0, case="tgv", storage=storage)
    g.Initialize()
    g.step(s.Time)
    ke2, mass2 = g.diagnostics()
    assert abs(ke - ke2) <= 1e-14 * ke2 and abs(mass - mass2) <= 1e-14 * mass2


@pytest.mark.parametrize("mode,engine", [("faithful", "fused"), ("faithful", "phased"), ("config", "fused"),
                                         ("config", "phased")])
def test_driver_cavity_matches_oracle(tmp_path, mode, engine):
    text = "1.0  0.1  10\n0.1  20  10\n24 24 24\n"
    ke, mass, out = run(tmp_path, text, ["--case", "cavity", "--mode", mode, "--engine", engine])
    assert_diag((ke, mass), oracle_diagnostics(tmp_path, "cavity", mode, 20))


@pytest.mark.parametrize("engine", ["fused", "phased"])
def test_driver_sphere_matches_oracle(tmp_path, engine):
    text = "1.0  8.0  100\n0.1  100  10\n72 20 20\n"
    ke, mass, _ = run(tmp_path, text, ["--case", "sphere", "--steps", "20", "--engine", engine])
    assert_diag((ke, mass), oracle_diagnostics(tmp_path, "sphere", "faithful", 20))


def test_driver_fast_path(tmp_path):
    """--fast (imexlb_step over the whole run) ends where the six-call loop
    ends, also with inter = 0, where only the last step writes moments."""
    text = "1.0  0.1  100\n0.05  40  10\n32 32 32\n"
    ke1, m1, _ = run(tmp_path, text, ["--case", "tgv", "--honour-input"])
    ke2, m2, out = run(tmp_path, text, ["--case", "tgv", "--honour-input", "--fast"])
    assert (ke1, m1) == (ke2, m2)
    assert_diag((ke1, m1), oracle_diagnostics(tmp_path, "tgv", "faithful", 40))
    text = "1.0  0.1  100\n0.05  13  0\n24 20 16\n"
    ke3, m3, _ = run(tmp_path, text, ["--case", "tgv", "--honour-input", "--fast", "--storage", "aa"])
    assert_diag((ke3, m3), oracle_diagnostics(tmp_path, "tgv", "faithful", 13))
    text = "1.0  0.1  10\n0.1  13  0\n20 16 12\n"
    ke4, m4, _ = run(tmp_path, text, ["--case", "cavity", "--mode", "faithful", "--honour-input", "--fast"])
    assert_diag((ke4, m4), oracle_diagnostics(tmp_path, "cavity", "faithful", 13))


def test_driver_window_placement(tmp_path):
    """--window places the lattice with the window search (then --tune picks
    the launch shape on it): the run ends where the run without it ends."""
    text = "1.0  0.1  100\n0.05  20  10\n32 32 32\n"
    ke1, m1, _ = run(tmp_path, text, ["--case", "tgv", "--honour-input", "--storage", "aa"])
    ke2, m2, out = run(tmp_path, text, ["--case", "tgv", "--honour-input", "--storage", "aa", "--window", "0.25",
                                        "--tune", str(tmp_path / "tune.jsonl")])
    m = re.search(r"Lattice placement window: kept offset (\d+) of (\d+)", out)
    assert m and int(m.group(2)) == 5 and 0 <= int(m.group(1)) < 5, out
    assert (ke1, m1) == (ke2, m2)
    text = "1.0  0.1  10\n0.1  13  10\n20 16 12\n"
    ke3, m3, _ = run(tmp_path, text, ["--case", "cavity", "--honour-input", "--window", "0.25"])
    assert_diag((ke3, m3), oracle_diagnostics(tmp_path, "cavity", "faithful", 13))


@pytest.mark.parametrize("storage", ["pull", "aa"])
def test_driver_restart_continues_the_run(tmp_path, storage):
    """30 steps, a restart file, then 11 more steps from it, end where one run
    of 41 steps ends (odd totals exercise the AA layout parity)."""
    text = "1.0  0.1  100\n0.05  41  0\n32 16 24\n"
    common = ["--case", "tgv", "--storage", storage, "--honour-input"]
    ke_full, mass_full, _ = run(tmp_path, text, common)
    ck = str(tmp_path / "run.ck")
    run(tmp_path, text, common + ["--steps", "30", "--checkpoint", ck])
    ke, mass, out = run(tmp_path, text, common + ["--steps", "11", "--restart", ck])
    assert "Restarted from" in out and "Warm-Up" not in out
    assert (ke, mass) == (ke_full, mass_full)


@pytest.mark.parametrize("engine", ["fused", "phased"])
def test_driver_restart_faithful_cavity(tmp_path, engine):
    """The reference's own cavity (faithful mode): 12 steps, a restart file
    with the ghost cells (the fused engine's ghost ring, or the phased
    engine's ghost-3 state), then 9 more end where 21 steps end."""
    text = "1.0  0.1  10\n0.1  21  0\n20 16 12\n"
    common = ["--case", "cavity", "--mode", "faithful", "--honour-input", "--engine", engine]
    ke_full, mass_full, _ = run(tmp_path, text, common)
    ck = str(tmp_path / "cav.ck")
    run(tmp_path, text, common + ["--steps", "12", "--checkpoint", ck])
    ke, mass, out = run(tmp_path, text, common + ["--steps", "9", "--restart", ck])
    assert "Restarted from" in out
    assert (ke, mass) == (ke_full, mass_full)


def test_driver_faithful_cavity_fused_engine(tmp_path):
    """The default fused engine runs the reference's own cavity rules on its
    ghost ring, to the final KE and mass of the phased engine.  The
    fields are the same bits (tests/test_gpu_parity.py); the two engines'
    FP64 reductions add them in different orders, so the sums agree to
    rounding."""
    text = "1.0  0.1  10\n0.1  15  0\n20 16 12\n"
    common = ["--case", "cavity", "--mode", "faithful", "--honour-input"]
    ke_p, mass_p, _ = run(tmp_path, text, common + ["--engine", "phased"])
    ke_f, mass_f, _ = run(tmp_path, text, common)
    assert abs(ke_f - ke_p) <= 1e-14 * ke_p and abs(mass_f - mass_p) <= 1e-14 * abs(mass_p)


def test_driver_output_files(tmp_path):
    """--output writes output0.plt (the initial fields) and one file every
    `inter` steps, as the cavity driver does (lid_driven_cavity/main.cpp:26,
    39-44); the last holds the moments of the last step."""
    import numpy as np
    from imexlb_amd import LBM, System
    text = "1.0  0.1  10\n0.1  20  10\n24 20 16\n"
    run(tmp_path, text, ["--case", "cavity", "--mode", "config", "--output"])
    n = 24 * 20 * 16
    for k in range(3):
        raw = (tmp_path / f"output{k}.plt").read_bytes()
        assert len(raw) == 332 + 7 * n * 8 and raw[:8] == b"#!TDV112"
    data0 = np.frombuffer((tmp_path / "output0.plt").read_bytes()[332:], "<f8").reshape(7, 16, 20, 24)
    assert np.all(data0[3:6] == 0.0) and np.all(data0[6] == 1.0 / 3.0)  # rest at p = rho0 / 3
    s = System(str(tmp_path / "input.in"), "cavity")
    g = LBM(s.sx, s.sy, s.sz, s.tau, s.rho0, s.u0, case="cavity", mode="config")
    g.Initialize()
    g.step(20)
    p, u, v, w = g.moments()
    data = np.frombuffer((tmp_path / "output2.plt").read_bytes()[332:], "<f8").reshape(7, 16, 20, 24)
    for k, a in enumerate((u, v, w, p)):
        assert np.array_equal(data[3 + k], a)


@pytest.mark.parametrize("case,steps", [("tgv", 4), ("cavity", 40), ("sphere", 10)])
def test_driver_runs_the_reference_inputs(tmp_path, case, steps):
    """The reference's own input.in files, unchanged (tgv 256^3 Re 1600,
    cavity 100^3, sphere 512x128x128), through the driver's six-call loop on
    the default (fused) engine, with the run shortened to `steps` steps: the
    reference's report lines, and the final KE and mass of the oracle's
    fields after the same steps."""
    text = open(os.path.join(ROOT, "tests", "golden", f"input_{case}.in")).read()
    args = ["--case", case, "--steps", str(steps)]
    print(f"[driver] reference input {case}, {steps} steps", flush=True)
    ke, mass, out = run(tmp_path, text, args)
    for line in ("Avg time spent in Pack:", "Avg time spent in Stream:", "Avg. Total Solver Time:", "Throughput:"):
        assert line in out
    assert_diag((ke, mass), oracle_diagnostics(tmp_path, case, "faithful", steps, keep_hardcoded=True))
