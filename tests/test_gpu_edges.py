"""GPU parity at the edges of the input space: ragged and tiny extents (a
partial wavefront, a single block, extents of 2 where x+1 and x-1 are the
same cell under the periodic wrap), every launch shape the C-ABI accepts,
and the error paths of a live handle.  FP64 is compared bit for bit with the
oracle; the launch shape must not change a single bit.
"""
import numpy as np
import pytest

from oracle.oracle import CONFIG, FAITHFUL, LDC, TGV, Oracle

pytestmark = pytest.mark.gpu

PI_R = 3.1415926

RAGGED = [(2, 2, 2), (67, 5, 3), (130, 3, 7), (5, 64, 2), (3, 3, 129)]
SHAPES = [(64, 1, 0, 0), (64, 2, 1, 1), (192, 1, 1, 0), (320, 2, 0, 1), (512, 1, 1, 1), (512, 2, 0, 0)]


def tgv_tau(n, re=100.0, u0=0.05):
    # tgv/src/System.cpp:21
    return u0 * n / re / (1.0 / 3.0) / 2 / PI_R


def run_pair(case, mode, n, tau, u0, storage, nsteps, shape=None, lid_u=0.1):
    from imexlb_amd import LBM
    o = Oracle(case, *n, tau0=tau, u0=u0, mode=mode, lid_u=lid_u)
    o.Initialize()
    g = LBM(*n, tau, u0=u0, case=case, mode=mode, storage=storage, lid_u=lid_u)
    if shape is not None:
        g.set_launch(*shape)
    g.Initialize()
    o.run(nsteps)
    g.step(nsteps)
    return o, g


def assert_bitexact(o, g):
    for name, a, b in zip("puvw", o.moments(), g.moments()):
        assert np.array_equal(a, b), f"moment {name}: max |d| = {np.abs(a - b).max():.3e}"
    o.Collision()
    assert np.array_equal(o.get_f(), g.get_f())


@pytest.mark.parametrize("n", RAGGED)
@pytest.mark.parametrize("storage", ["pull", "aa"])
def test_tgv_ragged_extents_bitexact(n, storage):
    print(f"[case] tgv ragged {n} {storage}", flush=True)
    o, g = run_pair(TGV, FAITHFUL, n, tgv_tau(n[0]), 0.05, storage, 7)
    assert_bitexact(o, g)


@pytest.mark.parametrize("n", [(2, 2, 2), (67, 5, 3), (3, 9, 130)])
@pytest.mark.parametrize("storage", ["pull", "aa"])
def test_cavity_config_ragged_extents_bitexact(n, storage):
    print(f"[case] cavity config ragged {n} {storage}", flush=True)
    o, g = run_pair(LDC, CONFIG, n, 0.096, 0.0, storage, 6)
    assert_bitexact(o, g)


@pytest.mark.parametrize("storage", ["pull", "aa"])
def test_every_launch_shape_gives_identical_bits(storage):
    """Block size, cells per thread, XCD tile order and the non-temporal load
    variant change only the schedule: the populations are bit-identical to
    the oracle for each shape, on an extent that leaves partial blocks."""
    n = (67, 13, 11)
    for shape in SHAPES:
        print(f"[case] launch shape {shape} {storage}", flush=True)
        o, g = run_pair(TGV, FAITHFUL, n, tgv_tau(n[0]), 0.05, storage, 5, shape=shape)
        assert_bitexact(o, g)


PAIR_EXTENTS = [(2, 3, 2), (4, 2, 3), (6, 5, 3), (64, 3, 2), (66, 5, 3), (130, 3, 3), (256, 2, 2)]
PAIR_SHAPES = [(64, 3, 0, 0), (256, 3, 1, 1), (512, 3, 1, 0)]


@pytest.mark.parametrize("storage", ["pull", "aa"])
def test_cell_pairs_bitexact(storage):
    """Two adjacent cells per thread (cells_per_thread 3) on periodic lattices
    with nx even, including the pairs at row ends whose x neighbours wrap
    (nx = 2: every pair): bit for bit the oracle's FP64 run; on odd nx and on
    the cavity the launch falls back to one cell per thread."""
    for n in PAIR_EXTENTS + [(67, 3, 2)]:
        for shape in PAIR_SHAPES:
            print(f"[case] cell pairs {n} {shape} {storage}", flush=True)
            o, g = run_pair(TGV, FAITHFUL, n, tgv_tau(n[0]), 0.05, storage, 5, shape=shape)
            assert_bitexact(o, g)
    print(f"[case] cell pairs cavity config {storage}", flush=True)
    o, g = run_pair(LDC, CONFIG, (6, 4, 5), 0.096, 0.0, storage, 4, shape=(64, 3, 1, 1))
    assert_bitexact(o, g)


@pytest.mark.parametrize("dtype", ["f32", "f16"])
@pytest.mark.parametrize("storage", ["pull", "aa"])
def test_cell_pairs_reduced_precision_identical(dtype, storage):
    """FP32/FP16 storage: the pair kernels store the same bits as one cell
    per thread (the same per-cell arithmetic and conversions)."""
    from imexlb_amd import LBM
    for n in PAIR_EXTENTS:
        print(f"[case] cell pairs {dtype} {n} {storage}", flush=True)
        tau = tgv_tau(n[0])
        runs = []
        for shape in [(256, 1, 1, 0), (128, 3, 1, 1), (64, 3, 0, 0)]:
            g = LBM(*n, tau, u0=0.05, dtype=dtype, storage=storage)
            g.set_launch(*shape)
            g.Initialize()
            g.step(5)
            runs.append((g.get_f(), g.moments()))
            g.close()
        for f, m in runs[1:]:
            assert np.array_equal(f, runs[0][0])
            for a, b in zip(m, runs[0][1]):
                assert np.array_equal(a, b)


@pytest.mark.parametrize("dtype,tol", [("f32", 2e-5), ("f16", 2e-2)])
def test_reduced_precision_ragged(dtype, tol):
    """FP32/FP16 storage on a ragged extent stays within the tolerance of the
    FP64 oracle (relative to the velocity scale u0)."""
    from imexlb_amd import LBM
    n = (67, 5, 9)
    tau = tgv_tau(n[0])
    o = Oracle(TGV, *n, tau0=tau, u0=0.05)
    o.Initialize()
    o.run(20)
    g = LBM(*n, tau, u0=0.05, dtype=dtype, storage="aa")
    g.Initialize()
    g.step(20)
    for a, b in zip(o.moments()[1:], g.moments()[1:]):
        assert np.abs(a - b).max() <= tol * 0.05


def test_live_handle_rejects_bad_calls(tmp_path):
    """Errors on a live handle come back as statuses with a message, and the
    handle keeps working afterwards."""
    from imexlb_amd import LBM, ImexlbError
    g = LBM(16, 8, 8, tgv_tau(16), u0=0.05)
    with pytest.raises(ImexlbError):
        g.step(1)  # before Initialize
    g.Initialize()
    for bad in [(32, 1, 1, 0), (576, 1, 1, 0), (64, 4, 1, 0), (64, 0, 1, 0), (64, 1, 2, 0), (64, 1, 1, 2)]:
        with pytest.raises(ImexlbError):
            g.set_launch(*bad)
    with pytest.raises(ValueError):
        g.set_f(np.zeros((27, 8, 8, 15)))
    with pytest.raises(ImexlbError):
        g.restart(str(tmp_path / "missing.ck"))
    g.step(3)
    o = Oracle(TGV, 16, 8, 8, tau0=tgv_tau(16), u0=0.05)
    o.Initialize()
    o.run(3)
    assert_bitexact(o, g)


@pytest.mark.parametrize("storage", ["pull", "aa"])
def test_placement_search_keeps_results(storage):
    """The lattice placement search swaps in the fastest of several
    allocations and re-initialises: the run that follows is the oracle's, bit
    for bit, and bad arguments come back as errors."""
    from imexlb_amd import LBM, ImexlbError
    n = (64, 24, 20)
    tau = tgv_tau(n[0])
    g = LBM(*n, tau, u0=0.05, storage=storage)
    g.Initialize()
    g.step(3)
    kept, ms = g.tune_placement(4, steps=4)
    assert 0 <= kept < 4 and len(ms) == 4
    assert all(m > 0 for m in ms) and ms[kept] == min(ms)
    assert g.info()["steps"] == 0
    g.step(9)
    o = Oracle(TGV, *n, tau0=tau, u0=0.05)
    o.Initialize()
    o.run(9)
    assert_bitexact(o, g)
    for bad in (0, 17):
        with pytest.raises(ImexlbError):
            g.tune_placement(bad)
    ph = LBM(16, 8, 8, tau, u0=0.05, engine="phased")
    with pytest.raises(ImexlbError):
        ph.tune_placement(2)


@pytest.mark.parametrize("storage", ["pull", "aa"])
def test_placement_window_keeps_results(storage):
    """The window placement search slides the lattice through one address
    range of lattice + slack, keeps the fastest offset and gives back the
    chunks of the range the lattice does not use: the run that follows is the
    oracle's, bit for bit, also after a second window search and after a
    separate-allocation search that starts from a window; the handle then
    holds less than two 2 MiB chunks beyond its lattice; bad arguments come
    back as errors."""
    from imexlb_amd import LBM, ImexlbError
    n = (64, 24, 20)
    tau = tgv_tau(n[0])
    g = LBM(*n, tau, u0=0.05, storage=storage)
    g.Initialize()
    base = g.info()["device_bytes"]
    slack = 8 << 20

    def oracle_after_7():
        # assert_bitexact collides the oracle in place: a fresh one per check
        o = Oracle(TGV, *n, tau0=tau, u0=0.05)
        o.Initialize()
        o.run(7)
        return o
    for rep in range(2):
        print(f"[window] {storage} search {rep}", flush=True)
        kept, ms = g.tune_placement_window(slack_gib=8 / 1024, step_gib=2 / 1024, steps=4)
        assert len(ms) == 5 and 0 <= kept < 5 and all(m > 0 for m in ms) and ms[kept] == min(ms)
        assert g.info()["steps"] == 0
        extra = g.info()["device_bytes"] - base
        assert 0 <= extra < 2 * (2 << 20) < slack, extra
        g.step(7)
        assert_bitexact(oracle_after_7(), g)
    held = g.info()["device_bytes"]
    kept, ms = g.tune_placement(3, steps=4)
    g.Initialize()
    g.step(7)
    assert_bitexact(oracle_after_7(), g)
    assert g.info()["device_bytes"] == (held if kept == 0 else base)
    for slack_gib, step_gib in ((8 / 1024, 3 / 1024), (8 / 1024, 1 / 1024)):
        with pytest.raises(ImexlbError):
            g.tune_placement_window(slack_gib=slack_gib, step_gib=step_gib)
    g.close()


@pytest.mark.parametrize("dtype", ["f32", "f16"])
@pytest.mark.parametrize("storage", ["pull", "aa"])
def test_reduced_precision_bits_do_not_depend_on_launch_shape(dtype, storage):
    """FP32/FP16 storage: every launch shape gives the same bits (the FP32
    collision is written without contraction, so no kernel variant rounds
    differently)."""
    from imexlb_amd import LBM
    n = (192, 5, 6)
    tau = tgv_tau(n[0])
    out = []
    for shape in SHAPES:
        g = LBM(*n, tau, u0=0.05, dtype=dtype, storage=storage)
        g.set_launch(*shape)
        g.Initialize()
        g.step(13)
        out.append((g.get_f(), g.moments()))
    for f, m in out[1:]:
        assert np.array_equal(out[0][0], f)
        for a, b in zip(out[0][1], m):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("dtype", ["f32", "f16"])
def test_reduced_precision_pull_and_aa_agree(dtype):
    """Pull and AA storage run the same collision on the same values: their
    moments agree bit for bit in FP32 and FP16 too."""
    from imexlb_amd import LBM
    n = (64, 9, 7)
    tau = tgv_tau(n[0])
    out = []
    for storage in ("pull", "aa"):
        g = LBM(*n, tau, u0=0.05, dtype=dtype, storage=storage)
        g.Initialize()
        g.step(13)
        out.append(g.moments())
    for a, b in zip(*out):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("storage,dtype", [("aa", "f64"), ("pull", "f64"), ("aa", "f16")])
def test_graph_replay_equals_single_launches(storage, dtype):
    """imexlb_step replays a captured graph for runs of moment-free steps.
    The result equals one launch per step, bit for bit, also after the launch
    shape or the lattice allocation changed under a cached graph."""
    from imexlb_amd import LBM
    n = (64, 10, 9)
    tau = tgv_tau(n[0])
    a = LBM(*n, tau, u0=0.05, storage=storage, dtype=dtype)
    b = LBM(*n, tau, u0=0.05, storage=storage, dtype=dtype)
    a.Initialize()
    b.Initialize()
    a.step(37, 0)      # two graph replays and five single launches
    a.set_launch(128, 2, 0, 1)
    a.step(35, 35)     # a new shape: a new graph; moments on the last step
    for _ in range(72):
        b.step(1, 1)   # one launch per step, moments every step
    for x, y in zip(a.moments(), b.moments()):
        assert np.array_equal(x, y)
    assert np.array_equal(a.get_f(), b.get_f())
    kept, _ = a.tune_placement(3, steps=2)  # new lattices, re-initialised
    b.Initialize()
    a.step(33, 33)
    b.step(33, 33)
    for x, y in zip(a.moments(), b.moments()):
        assert np.array_equal(x, y)


ADDR_CASES = [
    # case, mode, extents, tau0, u0, dtype, storage, sphere
    ("tgv", "faithful", (67, 13, 11), None, 0.05, "f64", "pull", None),
    ("tgv", "faithful", (67, 13, 11), None, 0.05, "f64", "aa", None),
    ("tgv", "faithful", (67, 13, 11), None, 0.05, "f32", "aa", None),
    ("tgv", "faithful", (67, 13, 11), None, 0.05, "f16", "pull", None),
    ("tgv", "faithful", (67, 13, 11), None, 0.05, "f16", "aa", None),
    ("cavity", "config", (24, 20, 18), 0.096, 0.0, "f64", "aa", None),
    ("cavity", "config", (24, 20, 18), 0.096, 0.0, "f16", "pull", None),
    ("sphere", "config", (30, 14, 16), 0.05, 0.05, "f64", "pull", (10.0, 7.0, 8.0, 3.5)),
    ("sphere", "config", (30, 14, 16), 0.05, 0.05, "f16", "pull", (10.0, 7.0, 8.0, 3.5)),
    ("cavity", "faithful", (24, 20, 18), 0.096, 0.0, "f64", "pull", None),
    ("sphere", "faithful", (72, 12, 9), 0.08, 0.1, "f64", "pull", None),
]


@pytest.mark.parametrize("case,mode,n,tau,u0,dtype,storage,sphere", ADDR_CASES)
def test_address_modes_give_identical_bits(case, mode, n, tau, u0, dtype, storage, sphere):
    """The fused kernels' three address modes (grouped uniform bases, a base
    per population, 64-bit offsets; the launch picks the most compact one the
    lattice allows) compute the same addresses: forcing each on a small
    lattice gives the same populations and moments, bit for bit."""
    from imexlb_amd import LBM
    print(f"[case] address modes {case} {mode} {n} {dtype} {storage}", flush=True)
    tau = tgv_tau(n[0]) if tau is None else tau
    out = []
    for am in (2, 1, 0):
        for shape in ((256, 1, 1, 1), (128, 2, 0, 0)):
            g = LBM(*n, tau, u0=u0, case=case, mode=mode, dtype=dtype, storage=storage, sphere=sphere,
                    addr_mode=am)
            assert g.info()["addr_mode"] == am
            g.set_launch(*shape)
            g.Initialize()
            g.step(7, 3)
            out.append((am, shape, g.get_f(), g.moments()))
            g.close()
    for am, shape, f, m in out[1:]:
        assert np.array_equal(out[0][2], f), f"populations, mode {am}, launch {shape}"
        for name, a, b in zip("puvw", out[0][3], m):
            assert np.array_equal(a, b), f"moment {name}, mode {am}, launch {shape}"
