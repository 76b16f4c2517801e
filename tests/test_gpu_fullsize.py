"""Full-size checks on the GPU: the benchmark sizes, through size-independent
properties (AA = pull bit for bit, mass conservation) and, where the oracle
finishes in seconds, the oracle itself."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
PI_R = 3.1415926


def test_tgv512_aa_equals_pull_and_conserves_mass():
    from imexlb_amd import LBM
    n = 512
    tau = 0.05 * n / 1600.0 / (1 / 3) / 2 / PI_R
    a = LBM(n, n, n, tau, storage="aa")
    a.Initialize()
    a.step(1)
    _, m1 = a.diagnostics()
    a.step(9)
    ma = a.moments()
    _, m10 = a.diagnostics()
    assert abs(m10 - m1) <= 1e-13 * m1
    a.close()
    b = LBM(n, n, n, tau, storage="pull")
    b.Initialize()
    b.step(10)
    for x, y in zip(ma, b.moments()):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("dtype", ["f32", "f16"])
def test_tgv512_cell_pairs_identical_reduced_precision(dtype):
    """TGV 512^3, AA, FP32/FP16 storage: the two-adjacent-cell kernel (the
    bench's FP16 choice) gives the moments of one cell per thread bit for bit
    after 20 steps, and AA equals pull with pairs too."""
    from imexlb_amd import LBM
    n = 512
    tau = 0.05 * n / 1600.0 / (1 / 3) / 2 / PI_R
    out = []
    for storage, shape in [("aa", (256, 1, 1, 1)), ("aa", (256, 3, 0, 1)), ("pull", (128, 3, 1, 1))]:
        print(f"[tgv512 pairs] {dtype} {storage} {shape}", flush=True)
        g = LBM(n, n, n, tau, u0=0.05, dtype=dtype, storage=storage)
        g.set_launch(*shape)
        g.Initialize()
        g.step(20)
        out.append(g.moments())
        g.close()
    for m in out[1:]:
        for name, x, y in zip("puvw", out[0], m):
            assert np.array_equal(x, y), f"moment {name}"


def test_cavity128_matches_oracle_after_1000_steps():
    """The cavity configuration (128^3, Re 400, lid 0.1, halfway bounce-back
    with the lid term) against the oracle, 1000 steps."""
    from imexlb_amd import LBM
    from oracle.oracle import CONFIG, LDC, Oracle
    n, lid, re = 128, 0.1, 400.0
    tau = 3.0 * lid * n / re
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1
    o = Oracle(LDC, n, n, n, tau0=tau, u0=0.0, mode=CONFIG, lid_u=lid, threads=threads)
    o.Initialize()
    o.run(1000)
    g = LBM(n, n, n, tau, u0=0.0, case="cavity", mode="config", storage="aa", lid_u=lid)
    g.Initialize()
    g.step(1000)
    for name, x, y in zip("puvw", g.moments(), o.moments()):
        assert np.array_equal(x, y), name


def test_sphere_channel_fullsize_runs():
    from imexlb_amd import LBM
    g = LBM(512, 256, 256, 0.03, u0=0.05, case="sphere", mode="config", sphere=(128, 128, 128, 20))
    g.Initialize()
    g.step(50)
    p, u, v, w = g.moments()
    assert np.isfinite(u).all() and np.isfinite(p).all()
    f = g.sphere_force()
    assert f[0] > 0  # drag points downstream
    assert abs(f[1]) < 1e-6 * f[0] and abs(f[2]) < 1e-6 * f[0]


def test_tgv512_reduced_precision_within_stated_error():
    """TGV 512^3 (Re 1600, U0 0.05, AA): after 200 steps the velocity L2 of
    FP16- and FP32-shifted storage against FP64 is within the stated bound
    (FP16 <= 1e-3, FP32 <= 1e-5), and every run keeps its mass."""
    from imexlb_amd import LBM
    n = 512
    tau = 0.05 * n / 1600.0 / (1 / 3) / 2 / PI_R
    vel = {}
    for dt in ("f64", "f32", "f16"):
        g = LBM(n, n, n, tau, dtype=dt, storage="aa")
        g.Initialize()
        _, m0 = g.diagnostics()
        g.step(200)
        _, m1 = g.diagnostics()
        assert abs(m1 - m0) <= (1e-12 if dt == "f64" else 1e-6) * m0
        vel[dt] = g.moments()[1:]
        g.close()

    def l2(a, b):
        return (sum(float(np.sum((x - y) ** 2)) for x, y in zip(a, b)) / sum(float(np.sum(y ** 2)) for y in b)) ** 0.5
    e16, e32 = l2(vel["f16"], vel["f64"]), l2(vel["f32"], vel["f64"])
    print(f"[tgv512] velocity L2 vs FP64 after 200 steps: FP32 {e32:.3e}, FP16 {e16:.3e}", flush=True)
    assert e16 <= 1e-3 and e32 <= 1e-5


def test_sphere_channel_fullsize_reduced_precision():
    """The sphere-channel configuration at its own size (512x256x256, D 40 at
    (128, 128, 128), inflow 0.05, Re 200) in FP32 and FP16 storage with FP32
    collision against FP64 storage after 500 steps: the drag proxy and the
    velocity field.  The FP32/FP16 lattices hold the departure from the
    inflow equilibrium."""
    from imexlb_amd import LBM
    nx, ny, nz = 512, 256, 256
    u_in, d, re = 0.05, 40.0, 200.0
    tau = 3.0 * u_in * d / re
    drag, vel = {}, {}
    for dt in ("f64", "f32", "f16"):
        print(f"[channel] {dt}: 500 steps", flush=True)
        g = LBM(nx, ny, nz, tau, u0=u_in, case="sphere", mode="config", dtype=dt, storage="pull",
                sphere=(128.0, 128.0, 128.0, d / 2))
        g.Initialize()
        g.step(500)
        drag[dt] = g.sphere_force()
        vel[dt] = g.moments()[1:]
        g.close()

    def l2(a, b):
        return (sum(float(np.sum((x - y) ** 2)) for x, y in zip(a, b)) / sum(float(np.sum(y ** 2)) for y in b)) ** 0.5
    err = {dt: (abs(drag[dt][0] - drag["f64"][0]) / abs(drag["f64"][0]), l2(vel[dt], vel["f64"]))
           for dt in ("f32", "f16")}
    print(f"[channel] after 500 steps: FP64 drag {drag['f64'][0]:.6e}; relative drag error / velocity L2 "
          f"FP32 {err['f32'][0]:.3e} / {err['f32'][1]:.3e}, FP16 {err['f16'][0]:.3e} / {err['f16'][1]:.3e}",
          flush=True)
    assert drag["f64"][0] > 0
    assert err["f32"][0] <= 1e-4 and err["f32"][1] <= 1e-4
    assert err["f16"][0] <= 1e-2 and err["f16"][1] <= 1e-3
