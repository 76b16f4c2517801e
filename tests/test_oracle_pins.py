"""CPU tests that pin the oracle (oracle/imexlb_oracle.c).

The reference ships no tests or golden field data, so the oracle is pinned by
the lattice table extracted from the reference source, by closed forms and
invariants that a dropped term, a wrong sign or a transposed index would
break, and by an independent NumPy restatement (oracle/oracle_np.py).
"""
import json
import os

import numpy as np
import pytest

from oracle.oracle import CONFIG, FAITHFUL, LDC, SPHERE, TGV, Oracle, tables
from oracle import oracle_np

HERE = os.path.dirname(os.path.abspath(__file__))
PI_R = 3.1415926
G = 3


def tgv_tau(n, re, u0=0.05):
    return u0 * n / re / (1.0 / 3.0) / 2 / PI_R


# ---------------------------------------------------------------- lattice

def test_tables_match_reference_fixture():
    fx = json.load(open(os.path.join(HERE, "golden", "lattice_d3q27.json")))
    e, bb, w = tables()
    assert e.tolist() == fx["e"]
    assert bb.tolist() == fx["bb"]
    assert np.array_equal(w, np.array([a / b for a, b in fx["weights_num_den"]]))
    # the NumPy restatement rebuilt its tables from the velocity list alone
    assert oracle_np.E.tolist() == fx["e"]
    assert oracle_np.BB.tolist() == fx["bb"]
    assert np.array_equal(oracle_np.W, w)


def test_lattice_isotropy():
    e, bb, w = tables()
    e = e.astype(float)
    assert abs(w.sum() - 1.0) < 1e-15
    assert np.abs(w @ e).max() < 1e-16
    second = np.einsum("q,qa,qb->ab", w, e, e)
    assert np.allclose(second, np.eye(3) / 3.0, atol=1e-15)
    third = np.einsum("q,qa,qb,qc->abc", w, e, e, e)
    assert np.abs(third).max() < 1e-16
    d = np.eye(3)
    fourth = np.einsum("q,qa,qb,qc,qd->abcd", w, e, e, e, e)
    want = (np.einsum("ab,cd->abcd", d, d) + np.einsum("ac,bd->abcd", d, d)
            + np.einsum("ad,bc->abcd", d, d)) / 9.0
    assert np.allclose(fourth, want, atol=1e-15)
    assert np.array_equal(e[bb], -e)
    assert np.array_equal(w[bb], w)
    assert len({tuple(v) for v in e.astype(int)}) == 27


# ---------------------------------------------------------------- init / eq

def test_tgv_initial_moments_are_the_analytic_fields():
    n = 12
    o = Oracle(TGV, n, n, n, tau0=0.05, u0=0.05)
    o.Initialize()
    p0, u0, v0, w0, _ = (a.copy() for a in o.fields_full())
    # closed form of the TGV field at the cell centres (tgv/src/lbm.cpp:197-204)
    x = (np.arange(n + 2 * G) - G) / n * 2.0 * PI_R
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    assert np.allclose(u0, 0.05 * np.sin(X) * np.cos(Y) * np.cos(Z), rtol=0, atol=1e-17)
    assert np.allclose(v0, -0.05 * np.cos(X) * np.sin(Y) * np.cos(Z), rtol=0, atol=1e-17)
    pex = 1.0 / 3.0 + 0.05 ** 2 / 16.0 * (np.cos(2 * X) + np.cos(2 * Y)) * (np.cos(2 * Z) + 2.0)
    assert np.allclose(p0, pex, rtol=0, atol=1e-16)
    # f = feq(p, u) carries exactly those moments: sum f = 3p, sum f e = u
    o.Update()
    p1, u1, v1, w1, _ = o.fields_full()
    assert np.abs(p1 - p0).max() < 1e-15
    # sums of 27 terms of size ~0.3: rounding ~1e-16
    assert np.abs(u1 - u0).max() < 2e-16 and np.abs(v1 - v0).max() < 2e-16
    assert np.abs(w1).max() < 2e-16


def test_collision_is_identity_at_equilibrium():
    """The first reference Collision is the identity: f0 is the feq of the
    fields it relaxes towards (the fused engine's initial state relies on it)."""
    o = Oracle(TGV, 10, 9, 8, tau0=0.02, u0=0.05)
    o.Initialize()
    f0 = o.get_f_full()
    o.Collision()
    assert np.array_equal(o.get_f_full(), f0)


def test_collision_conserves_mass_and_momentum_and_relaxes():
    rng = np.random.default_rng(1)
    o = Oracle(TGV, 6, 5, 7, tau0=0.3, u0=0.05)
    o.Initialize()
    f = o.get_f_full() * (1 + 0.05 * rng.standard_normal(o.get_f_full().shape))
    o.set_f_full(f)
    o.Update()
    p, u, v, w, _ = (a.copy() for a in o.fields_full())
    o.Collision()
    post = o.get_f_full()
    sl = (slice(G, -G),) * 3
    e, _, wt = tables()
    fi, pi = f[(slice(None),) + sl], post[(slice(None),) + sl]
    # sums of 27 terms of size ~0.3 each: a few ulp of 1
    assert np.allclose(pi.sum(0), fi.sum(0), rtol=0, atol=4e-15)
    for a in range(3):
        assert np.allclose(np.tensordot(e[:, a], pi, 1), np.tensordot(e[:, a], fi, 1), rtol=0, atol=4e-15)
    # BGK: post = f - (f - feq)/(tau + 1/2), feq written out independently
    uu = u[sl] ** 2 + v[sl] ** 2 + w[sl] ** 2
    for q in range(27):
        eu = e[q, 0] * u[sl] + e[q, 1] * v[sl] + e[q, 2] * w[sl]
        feq = wt[q] * (3 * p[sl] + 3 * eu + 4.5 * eu ** 2 - 1.5 * uu)
        assert np.allclose(pi[q], fi[q] - (fi[q] - feq) / 0.8, rtol=0, atol=1e-15)
    # ghosts untouched
    assert np.array_equal(post[:, :G], f[:, :G])


# ---------------------------------------------------------------- halo / streaming

def test_tgv_halo_is_modular_indexing():
    """The 26 self-messages of the one-rank halo (tgv/src/mpi_view_transfer.cpp)
    leave every ghost cell equal to its periodic image."""
    rng = np.random.default_rng(2)
    nx, ny, nz = 7, 8, 9
    o = Oracle(TGV, nx, ny, nz, tau0=0.1)
    o.Initialize()
    f = rng.standard_normal(o.get_f_full().shape)
    o.set_f_full(f)
    o.pack()
    o.exchange()
    o.unpack()
    g = o.get_f_full()
    ii = (np.arange(nx + 6) - G) % nx + G
    jj = (np.arange(ny + 6) - G) % ny + G
    kk = (np.arange(nz + 6) - G) % nz + G
    want = f[:, ii][:, :, jj][:, :, :, kk]
    assert np.array_equal(g, want)


def test_streaming_is_a_permutation():
    rng = np.random.default_rng(3)
    n = (6, 5, 4)
    o = Oracle(TGV, *n, tau0=0.1)
    o.Initialize()
    f = rng.standard_normal((27,) + n[::-1])
    o.set_f(f)
    o.pack(), o.exchange(), o.unpack()
    o.Streaming()
    s = o.get_f()
    e, _, _ = tables()
    for q in range(27):
        # f_q(x) <- f_q(x - e_q), arrays indexed [z][y][x]
        assert np.array_equal(s[q], np.roll(f[q], shift=(e[q, 2], e[q, 1], e[q, 0]), axis=(0, 1, 2)))


def test_cavity_faithful_walls_follow_the_reference_rule():
    """f_i(g) = f_bb(i)(g + 2 e_i) on the x-low wall; the top wall reads
    j + e_y (lid_driven_cavity/src/lbm.cpp:298)."""
    rng = np.random.default_rng(4)
    o = Oracle(LDC, 6, 6, 6, tau0=0.1, u0=0.1, mode=FAITHFUL)
    o.Initialize()
    f = rng.standard_normal(o.get_f_full().shape)
    o.set_f_full(f)
    o.Streaming()
    e, bb, _ = tables()
    g = o.get_f_full()  # interior streamed, ghosts hold the wall values
    q = 7  # e = (1, 1, 0)
    # walls run in order: bcl first, later kernels may overwrite shared ghost
    # rows, so check rows only bcl writes: j, k in [4, 7]
    for j in range(4, 8):
        for k in range(4, 8):
            assert g[q, G - 1, j, k] == f[bb[q], G + 1, j + 2 * e[q, 1], k + 2 * e[q, 2]]
    q = 18  # e = (0, -1, 1)  -> bcd (z low) writes it; 13: e = (1, 0, -1) -> bcu
    q = 16  # e = (0, -1, -1): top wall (e_z < 0) uses j + e_y
    le2 = G + 6
    for i in range(4, 8):
        for j in range(4, 8):
            assert g[q, i, j, le2] == f[bb[q], i + 2 * e[q, 0], j + e[q, 1], le2 - 2]


# ---------------------------------------------------------------- cross-check

CROSS = [
    (TGV, FAITHFUL, (8, 6, 10), dict(tau0=0.05, u0=0.05)),
    (LDC, FAITHFUL, (8, 7, 9), dict(tau0=0.1, u0=0.1)),
    (LDC, CONFIG, (8, 7, 9), dict(tau0=0.1, u0=0.0, lid_u=0.1)),
    (SPHERE, FAITHFUL, (40, 12, 12), dict(tau0=0.1, u0=0.1)),
    # the ball reaches the first y ghost layer (solid ghost neighbours of solid cells)
    (SPHERE, FAITHFUL, (72, 9, 12), dict(tau0=0.1, u0=0.1)),
    # nx <= 8: the one-voxel ball lies on the inlet plane
    (SPHERE, FAITHFUL, (8, 3, 4), dict(tau0=0.1, u0=0.1)),
    (SPHERE, CONFIG, (24, 12, 14), dict(tau0=0.1, u0=0.05, sphere=(8.0, 6.0, 7.0, 3.0))),
]


def _rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("case,mode,n,kw", CROSS)
def test_oracle_matches_numpy_restatement(case, mode, n, kw):
    o = Oracle(case, *n, mode=mode, **kw)
    m = oracle_np.NpModel(case, *n, mode=mode, **kw)
    o.Initialize()
    m.initialize()
    o.run(20)
    m.run(20)
    mo, mn = o.moments(), m.moments()
    # p, u relative to their field maximum; w of the TGV grows from zero
    for k, (a, b) in enumerate(zip(mn, mo)):
        scale = max(np.abs(b).max(), 1e-3 if k else 0.0)
        assert np.abs(a - b).max() <= 1e-12 * scale, (k, _rel(a, b))


def test_cross_check_catches_planted_mistakes(monkeypatch):
    """A dropped -1.5 u.u term, or a transposed bounce-back table, in one of the
    two restatements makes them disagree far beyond rounding."""
    case, mode, n, kw = CROSS[2]
    o = Oracle(case, *n, mode=mode, **kw)
    o.Initialize()
    o.run(20)
    ref = o.moments()[1]

    def run_np():
        m = oracle_np.NpModel(case, *n, mode=mode, **kw)
        m.initialize()
        m.run(20)
        return m.moments()[1]

    good = run_np()
    assert np.abs(good - ref).max() <= 1e-12 * np.abs(ref).max()

    orig_feq = oracle_np.feq

    def feq_no_uu(p, u, v, w):
        return orig_feq(p, u, v, w) + oracle_np.W.reshape((27,) + (1,) * np.ndim(p)) * 1.5 * (u * u + v * v + w * w)

    monkeypatch.setattr(oracle_np, "feq", feq_no_uu)
    bad = run_np()
    assert not np.abs(bad - ref).max() <= 1e-6 * np.abs(ref).max()
    monkeypatch.setattr(oracle_np, "feq", orig_feq)
    monkeypatch.setattr(oracle_np, "BB", np.roll(oracle_np.BB, 1))
    with np.errstate(all="ignore"):
        bad = run_np()
    # (a wrong partner table may even blow up: NaN also fails the check)
    assert not np.abs(bad - ref).max() <= 1e-6 * np.abs(ref).max()


# ---------------------------------------------------------------- physics

def test_tgv_mass_is_conserved():
    o = Oracle(TGV, 16, 16, 16, tau0=tgv_tau(16, 100))
    o.Initialize()
    o.run(1)
    _, m0 = o.diagnostics()
    o.run(60)
    _, m1 = o.diagnostics()
    assert abs(m1 - m0) <= 1e-13 * m0


@pytest.mark.slow
def test_tgv_early_time_viscous_decay():
    """E(t)/E(0) = exp(-6 nu k^2 t), k = 2 pi / N, at low Reynolds number."""
    n, re, u0 = 32, 10.0, 0.02
    tau = tgv_tau(n, re, u0)
    nu = tau / 3.0
    o = Oracle(TGV, n, n, n, tau0=tau, u0=u0)
    o.Initialize()
    o.run(1)
    ke0, _ = o.diagnostics()
    steps = 300
    o.run(steps)
    ke1, _ = o.diagnostics()
    k = 2 * np.pi / n
    want = np.exp(-6 * nu * k * k * steps)
    assert abs(ke1 / ke0 - want) < 0.01 * want, (ke1 / ke0, want)


def test_sphere_faithful_perturbation_is_seeded():
    from oracle.oracle import lib
    vals = [lib().orc_sphere_ran(5374857, i, j, k, 20, 20) for i in range(4) for j in range(4) for k in range(4)]
    assert all(0 <= v < 100 for v in vals)
    assert len(set(vals)) > 20
    assert vals == [lib().orc_sphere_ran(5374857, i, j, k, 20, 20)
                    for i in range(4) for j in range(4) for k in range(4)]


def test_sphere_faithful_mask_matches_reference_formula():
    # sphere/src/lbm.cpp:238: centre (glx-1)/8 (integer) on all axes, r = (glx-1)/64
    n = (72, 20, 20)
    o = Oracle(SPHERE, *n, tau0=0.1, u0=0.1)
    o.Initialize()
    usr = o.fields_full()[4]
    c, r = (n[0] - 1) // 8, (n[0] - 1) / 64.0
    idx = np.arange(-G, max(n) + G)
    X, Y, Z = np.meshgrid(idx[:n[0] + 6], idx[:n[1] + 6], idx[:n[2] + 6], indexing="ij")
    want = np.where((X - c) ** 2 + (Y - c) ** 2 + (Z - c) ** 2 <= r * r, 0, 1)
    assert np.array_equal(usr, want)
    assert (usr == 0).sum() > 0


# ---------------------------------------------------------------- closed forms of the config-mode terms
#
# Neither term exists in the reference: its lid is an overwrite of u in Update
# (lid_driven_cavity/src/lbm.cpp:343-344) and its sphere has no drag
# (sphere/src/lbm.cpp:350-356).  Both are pinned here by what the mathematics
# fixes, on the C oracle and on the NumPy restatement, and each check is shown
# to fail on a planted mistake.

def _lid_one_step(moms, rho0, lid_u):
    """One config-cavity step from rest.  A top-layer cell away from the side
    walls pulls its nine e_z = -1 populations off the lid:
    f_i = w_i rho0 + 6 w_i rho0 e_ix u_lid, so u_x = 6 rho0 u_lid sum w_i e_ix^2
    = rho0 u_lid / 3 (the sum over e_z = -1 is 2/54 + 4/216 = 1/18), and
    p = sum f / 3 = rho0 / 3 (sum w_i e_ix = 0 there).  Every cell below the
    top layer is still at rest."""
    p, u, v, w = moms  # [z][y][x]
    top = (-1, slice(1, -1), slice(1, -1))
    ok = (np.abs(u[top] - rho0 * lid_u / 3.0).max() <= 1e-15
          and np.abs(p[top] - rho0 / 3.0).max() <= 1e-15
          and np.abs(v[top]).max() <= 1e-16 and np.abs(w[top]).max() <= 1e-16)
    below = (slice(0, -1),)
    ok = ok and all(np.abs(a[below]).max() <= 1e-16 for a in (u, v, w))
    return ok and np.abs(p[below] - rho0 / 3.0).max() <= 1e-15


LID = dict(tau0=0.1, rho0=1.25, u0=0.0, lid_u=0.08)


def test_config_lid_term_one_step_closed_form():
    n = (7, 6, 5)
    o = Oracle(LDC, *n, mode=CONFIG, **LID)
    o.Initialize()
    o.run(1)
    assert _lid_one_step(o.moments(), LID["rho0"], LID["lid_u"])
    m = oracle_np.NpModel(LDC, *n, mode=CONFIG, **LID)
    m.initialize()
    m.run(1)
    assert _lid_one_step(m.moments(), LID["rho0"], LID["lid_u"])
    # the lid's momentum reaches only the top layer in one step
    assert abs(o.moments()[1][-1, 2, 3] - LID["rho0"] * LID["lid_u"] / 3.0) <= 1e-15


def test_lid_closed_form_catches_planted_mistakes(monkeypatch):
    """A flipped lid sign, a missing 1/cs^2, or the lid velocity taken along y
    in the restatement's lid term each break the one-step closed form."""
    n = (7, 6, 5)
    orig = oracle_np.lid_term
    bad_terms = [
        lambda q, r, u: -orig(q, r, u),
        lambda q, r, u: orig(q, r, u) / 3.0,
        lambda q, r, u: 6.0 * oracle_np.W[q] * r * oracle_np.E[q, 1] * u,
    ]
    for bad in bad_terms:
        monkeypatch.setattr(oracle_np, "lid_term", bad)
        m = oracle_np.NpModel(LDC, *n, mode=CONFIG, **LID)
        m.initialize()
        m.run(1)
        assert not _lid_one_step(m.moments(), LID["rho0"], LID["lid_u"])


SPH = dict(tau0=0.1, rho0=1.0, u0=0.05)
SPH_N = (9, 8, 7)
SPH_BALL = (4.0, 3.0, 3.0, 0.4)  # one solid voxel, at (4, 3, 3)


def _np_equilibrium_force(vel):
    """Drag proxy of the NumPy restatement on a uniform equilibrium field of
    velocity vel around the one-voxel ball."""
    m = oracle_np.NpModel(SPHERE, *SPH_N, mode=CONFIG, sphere=SPH_BALL, **SPH)
    m.initialize()
    assert (m.usr == 0).sum() == 1
    f = oracle_np.feq(np.full(SPH_N, SPH["rho0"] / 3.0), *(np.full(SPH_N, c) for c in vel))
    return oracle_np.sphere_force(f, m.usr)


def test_sphere_force_closed_form():
    """A one-voxel solid in a uniform equilibrium stream of velocity U: each of
    its 26 fluid neighbours s - e_i hands it 2 feq_i(U) e_i, and
    sum_i feq_i e_i = U (isotropy of the weights), so F = 2 U."""
    u0 = SPH["u0"]
    o = Oracle(SPHERE, *SPH_N, mode=CONFIG, sphere=SPH_BALL, **SPH)
    o.Initialize()
    assert (o.fields_full()[4][G:-G, G:-G, G:-G] == 0).sum() == 1
    o.Collision()  # the identity at equilibrium
    assert np.abs(o.sphere_force() - np.array([2 * u0, 0.0, 0.0])).max() <= 1e-15
    # any uniform equilibrium velocity, set directly as post-collision populations
    vel = (0.031, -0.047, 0.019)
    f = oracle_np.feq(np.full((o.lx, o.ly, o.lz), SPH["rho0"] / 3.0),
                      *(np.full((o.lx, o.ly, o.lz), c) for c in vel))
    o.set_f_full(f)
    assert np.abs(o.sphere_force() - 2 * np.array(vel)).max() <= 1e-15
    assert np.abs(_np_equilibrium_force((u0, 0.0, 0.0)) - np.array([2 * u0, 0.0, 0.0])).max() <= 1e-15
    assert np.abs(_np_equilibrium_force(vel) - 2 * np.array(vel)).max() <= 1e-15


def test_sphere_force_random_field_matches_restatement():
    """On random post-collision populations around a ball of several voxels the
    C oracle (enumerated from the fluid side) equals the restatement
    (enumerated from the solid side)."""
    rng = np.random.default_rng(11)
    ball = (6.0, 5.0, 5.0, 1.8)
    n = (13, 11, 11)
    o = Oracle(SPHERE, *n, mode=CONFIG, sphere=ball, **SPH)
    o.Initialize()
    f = rng.standard_normal((27, o.lx, o.ly, o.lz))
    o.set_f_full(f)
    m = oracle_np.NpModel(SPHERE, *n, mode=CONFIG, sphere=ball, **SPH)
    m.initialize()
    assert (m.usr == 0).sum() > 10
    want = oracle_np.sphere_force(f[:, G:-G, G:-G, G:-G], m.usr)
    assert np.abs(o.sphere_force() - want).max() <= 1e-12 * np.abs(want).max()


def test_sphere_force_closed_form_catches_planted_mistakes(monkeypatch):
    """A missing factor 2 or a reversed link direction in the restatement's
    drag proxy breaks F = 2 U."""
    vel = (0.031, -0.047, 0.019)
    orig = oracle_np.sphere_force
    for bad in (lambda f, usr: orig(f, usr) / 2.0, lambda f, usr: -orig(f, usr)):
        monkeypatch.setattr(oracle_np, "sphere_force", bad)
        got = _np_equilibrium_force(vel)
        assert not np.abs(got - 2 * np.array(vel)).max() <= 1e-15


def test_sphere_force_cross_check_catches_upstream_downstream_swap():
    """Taking the populations of the solid's downstream neighbours s + e_i
    leaves a uniform field's F = 2 U unchanged, so the random-field
    cross-check is what catches it."""
    rng = np.random.default_rng(11)
    ball = (6.0, 5.0, 5.0, 1.8)
    n = (13, 11, 11)
    o = Oracle(SPHERE, *n, mode=CONFIG, sphere=ball, **SPH)
    o.Initialize()
    f = rng.standard_normal((27, o.lx, o.ly, o.lz))
    o.set_f_full(f)
    m = oracle_np.NpModel(SPHERE, *n, mode=CONFIG, sphere=ball, **SPH)
    m.initialize()
    fi, usr = f[:, G:-G, G:-G, G:-G], m.usr
    bad = np.zeros(3)
    for s_ in np.argwhere(usr == 0):
        for q in range(1, 27):
            x, y, z = s_ + oracle_np.E[q]
            if usr[x, y, z] == 1:
                bad += 2.0 * fi[q, x, y, z] * oracle_np.E[q]
    assert not np.abs(o.sphere_force() - bad).max() <= 1e-12 * np.abs(bad).max()
