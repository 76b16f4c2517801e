"""GPU parity at the benchmark configurations' own sizes: the Taylor-Green
vortex at 64^3, Re 100, for its full 1000 steps, and the 128^3 lid-driven
cavity (config mode, Re 400, lid 0.1) and the reference's own cavity rules
(faithful mode, phased engine) for their first steps, each against the CPU
oracle, FP64 bit for bit."""
import numpy as np
import pytest

from oracle.oracle import CONFIG, FAITHFUL, LDC, SPHERE, TGV, Oracle

pytestmark = pytest.mark.gpu

PI_R = 3.1415926


def fields_equal(o, g, what):
    for name, a, b in zip("puvw", o.moments(), g.moments()):
        assert np.array_equal(a, b), f"{what}: moment {name} differs by {np.abs(a - b).max():.3e}"


def test_tgv64_re100_full_run_bitexact():
    """64^3 periodic, U0 0.05, Re 100 (tau0 = 3 nu with nu = U0 64 / (2 pi Re)),
    equilibrium start, 1000 steps: pull and AA storage both equal the oracle
    at step 500 and step 1000, moments and post-collision populations."""
    from imexlb_amd import LBM
    n = 64
    tau = 0.05 * n / 100.0 / (1.0 / 3.0) / 2 / PI_R
    o = Oracle(TGV, n, n, n, tau0=tau, u0=0.05)
    o.Initialize()
    gs = {s: LBM(n, n, n, tau, u0=0.05, storage=s) for s in ("pull", "aa")}
    for g in gs.values():
        g.Initialize()
    for k in (500, 500):
        print(f"[tgv64] oracle {k} steps", flush=True)
        o.run(k)
        for s, g in gs.items():
            g.step(k)
            fields_equal(o, g, f"tgv64 {s}")
    # the fused engine's state is the post-collision populations of the step
    o.Collision()
    for g in gs.values():
        assert np.array_equal(o.get_f(), g.get_f())


def test_cavity128_config_first_steps_bitexact():
    """128^3 cavity, Re 400 (tau0 = 3 * 0.1 * 128 / 400), lid 0.1 with the
    momentum term, rest at rho0 = 1: fused AA equals the oracle after 150
    steps."""
    from imexlb_amd import LBM
    n, lid = 128, 0.1
    tau = 3.0 * lid * n / 400.0
    o = Oracle(LDC, n, n, n, tau0=tau, u0=0.0, mode=CONFIG, lid_u=lid)
    o.Initialize()
    g = LBM(n, n, n, tau, u0=0.0, case="cavity", mode="config", storage="aa", lid_u=lid)
    g.Initialize()
    print("[cavity128] oracle 150 steps", flush=True)
    o.run(150)
    g.step(150)
    fields_equal(o, g, "cavity128 config")


def test_cavity128_faithful_phased_first_steps_bitexact():
    """The same cavity with the reference's own wall, lid and halo rules on
    the phased engine: equal to the oracle after 40 steps."""
    from imexlb_amd import LBM
    n = 128
    o = Oracle(LDC, n, n, n, tau0=0.096, u0=0.1, mode=FAITHFUL, lid_u=0.1)
    o.Initialize()
    g = LBM(n, n, n, 0.096, u0=0.1, case="cavity", mode="faithful", engine="phased")
    g.Initialize()
    print("[cavity128 faithful] oracle 40 steps", flush=True)
    o.run(40)
    g.step(40)
    fields_equal(o, g, "cavity128 faithful")
    assert np.array_equal(o.get_f(), g.get_f())


def test_cavity128_faithful_fused_first_steps_bitexact():
    """The reference's own cavity rules on the fused engine at 128^3: equal to
    the oracle after 40 steps, populations included."""
    from imexlb_amd import LBM
    n = 128
    o = Oracle(LDC, n, n, n, tau0=0.096, u0=0.1, mode=FAITHFUL, lid_u=0.1)
    o.Initialize()
    g = LBM(n, n, n, 0.096, u0=0.1, case="cavity", mode="faithful", engine="fused")
    g.Initialize()
    print("[cavity128 faithful fused] oracle 40 steps", flush=True)
    o.run(40)
    g.step(40)
    fields_equal(o, g, "cavity128 faithful fused")
    o.Collision()
    fg = g.get_f()
    assert np.array_equal(o.get_f(), fg)
    # two cells per thread, both tile orders: the same bits as the shape above
    mg = g.moments()
    for shape in [(128, 2, 1, 1), (64, 2, 0, 0)]:
        h = LBM(n, n, n, 0.096, u0=0.1, case="cavity", mode="faithful", engine="fused")
        h.set_launch(*shape)
        assert h.info()["cells_per_thread"] == 2
        h.Initialize()
        h.step(40)
        for a, b in zip(mg, h.moments()):
            assert np.array_equal(a, b), shape
        assert np.array_equal(fg, h.get_f()), shape
        h.close()


def test_reference_sphere_input_faithful_both_engines():
    """The reference's own sphere input (sphere/input.in: 512 x 128 x 128,
    Re 100, u0 0.05, tau0 = 3 u0 2R / Re with R = 511 / 64) with its own
    rules, on the phased and on the fused engine: both equal the oracle after
    4 steps."""
    from oracle.oracle import SPHERE
    from imexlb_amd import LBM
    nx, ny, nz = 512, 128, 128
    r = (nx - 1) / 64.0
    tau = 0.05 * 2 * r / 100.0 / (1.0 / 3.0)
    o = Oracle(SPHERE, nx, ny, nz, tau0=tau, u0=0.05, mode=FAITHFUL)
    o.Initialize()
    print("[sphere input] oracle 4 steps", flush=True)
    o.run(4)
    mo = o.moments()
    for engine in ("phased", "fused"):
        g = LBM(nx, ny, nz, tau, u0=0.05, case="sphere", mode="faithful", engine=engine)
        g.Initialize()
        g.step(4)
        for name, a, b in zip("puvw", mo, g.moments()):
            assert np.array_equal(a, b), f"sphere input {engine}: moment {name} differs"
        g.close()


def test_sphere_channel_first_steps_bitexact():
    """512 x 256 x 256 channel, sphere D = 40 at (128, 128, 128), inlet 0.05,
    Re 200 (tau0 = 3 * 0.05 * 40 / 200): fused pull equals the oracle on the
    fluid cells after 4 steps (solid cells hold p = u = 0 in the fused
    engine)."""
    from oracle.oracle import SPHERE
    from imexlb_amd import LBM
    nx, ny, nz = 512, 256, 256
    tau, sph = 3.0 * 0.05 * 40.0 / 200.0, (128.0, 128.0, 128.0, 20.0)
    o = Oracle(SPHERE, nx, ny, nz, tau0=tau, u0=0.05, mode=CONFIG, sphere=sph)
    o.Initialize()
    g = LBM(nx, ny, nz, tau, u0=0.05, case="sphere", mode="config", storage="pull", sphere=sph)
    g.Initialize()
    print("[sphere channel] oracle 4 steps", flush=True)
    o.run(4)
    g.step(4)
    usr = o.fields_full()[4]
    fluid = usr[3:-3, 3:-3, 3:-3].transpose(2, 1, 0) == 1
    assert 0 < (~fluid).sum() < 40 ** 3
    for name, a, b in zip("puvw", o.moments(), g.moments()):
        assert np.array_equal(a[fluid], b[fluid]), f"sphere channel: moment {name} differs"
        assert np.all(b[~fluid] == 0.0)


def test_tgv512_first_steps_bitexact():
    """The headline configuration, TGV 512^3, U0 0.05, Re 1600, FP64 in AA
    storage: equal to the oracle after 3 steps (both AA layouts), with one
    cell per thread and with the two-adjacent-cell kernel the bench picks."""
    from imexlb_amd import LBM
    n = 512
    tau = 0.05 * n / 1600.0 / (1.0 / 3.0) / 2 / PI_R
    runs = []
    for shape in [None, (128, 3, 1, 1)]:
        print(f"[tgv512] GPU 3 steps, launch {shape}", flush=True)
        g = LBM(n, n, n, tau, u0=0.05, storage="aa")
        if shape is not None:
            g.set_launch(*shape)
        g.Initialize()
        g.step(3)
        runs.append((shape, g.moments()))
        g.close()
    o = Oracle(TGV, n, n, n, tau0=tau, u0=0.05)
    o.Initialize()
    print("[tgv512] oracle 3 steps", flush=True)
    o.run(3)
    mo = o.moments()
    for shape, mg in runs:
        for name, a, b in zip("puvw", mo, mg):
            assert np.array_equal(a, b), f"tgv512 {shape}: moment {name} differs by {np.abs(a - b).max():.3e}"
