"""Second, independent restatement of the D3Q27 BGK step in NumPy.

Test infrastructure only.  It exists to catch transcription mistakes in the C
oracle (oracle/imexlb_oracle.c): the tests run both on small grids and require
agreement to rounding.  It is written differently on purpose:

* the lattice is rebuilt from the velocity list alone: weights from |e|^2,
  bounce-back partners by searching for -e;
* periodic axes use np.roll on the interior instead of the reference's 26
  ghost messages (tgv/src/mpi_view_transfer.cpp:117-383);
* the config-mode boundaries are applied as direct pull rules on the interior
  instead of ghost fills;
* the faithful-mode wall kernels are vectorised slices of the ghost-3 array
  (lid_driven_cavity/src/lbm.cpp:233-301, sphere/src/lbm.cpp:277-356).

Cases and modes match the C oracle: TGV (0), cavity (1), sphere (2); faithful
(0) or config (1).  Arrays are indexed [i, j, k] = [x, y, z] internally and
returned as [z, y, x] like the oracle's getters.
"""
import numpy as np

# velocity list in the reference's order, tgv/src/lbm.cpp:84-190
_VEL = """0 0 0 | 1 0 0 | -1 0 0 | 0 1 0 | 0 -1 0 | 0 0 1 | 0 0 -1 | 1 1 0 | -1 -1 0 |
1 -1 0 | -1 1 0 | 1 0 1 | -1 0 -1 | 1 0 -1 | -1 0 1 | 0 1 1 | 0 -1 -1 | 0 1 -1 | 0 -1 1 |
1 1 1 | -1 -1 -1 | 1 -1 1 | -1 1 -1 | 1 1 -1 | -1 -1 1 | 1 -1 -1 | -1 1 1"""
E = np.array([[int(t) for t in blk.split()] for blk in _VEL.replace("\n", " ").split("|")],
             dtype=np.int64)
Q = len(E)
_WBY = {0: 8.0 / 27.0, 1: 2.0 / 27.0, 2: 1.0 / 54.0, 3: 1.0 / 216.0}
W = np.array([_WBY[int((e * e).sum())] for e in E])
BB = np.array([int(np.where((E == -e).all(axis=1))[0][0]) for e in E])
G = 3
PI_R = 3.1415926


def _mix64(x):
    x = x.astype(np.uint64)
    x ^= x >> np.uint64(30)
    x *= np.uint64(0xBF58476D1CE4E5B9)
    x ^= x >> np.uint64(27)
    x *= np.uint64(0x94D049BB133111EB)
    x ^= x >> np.uint64(31)
    return x


def feq(p, u, v, w):
    """Equilibrium for every direction, stacked on axis 0."""
    out = np.empty((Q,) + np.shape(p))
    uu = u * u + v * v + w * w
    for q in range(Q):
        eu = E[q, 0] * u + E[q, 1] * v + E[q, 2] * w
        out[q] = W[q] * p * 3.0 + W[q] * (3.0 * eu + 4.5 * eu * eu - 1.5 * uu)
    return out


def lid_term(q, rho0, lid_u):
    """Momentum carried into the fluid by halfway bounce-back off the moving
    lid (config-mode cavity; the reference's own lid is an overwrite in
    Update, lid_driven_cavity/src/lbm.cpp:343-344): 2 w_q rho0 (e_q . u_w) / cs^2
    for the direction q leaving the lid, cs^2 = 1/3, lid velocity (lid_u, 0, 0)."""
    return 2.0 * W[q] * rho0 * float(E[q] @ np.array([lid_u, 0.0, 0.0])) * 3.0


def sphere_force(f, usr):
    """Momentum-exchange drag proxy, enumerated from the solid side: every
    solid cell s and direction q whose upstream cell s - e_q is fluid and
    inside the box hands the body 2 f*_q(s - e_q) e_q.  f is post-collision,
    [q, x, y, z] over the interior; usr is 1 for fluid, 0 for solid."""
    nx, ny, nz = usr.shape
    out = np.zeros(3)
    for s in np.argwhere(usr == 0):
        for q in range(1, Q):
            x, y, z = s - E[q]
            if 0 <= x < nx and 0 <= y < ny and 0 <= z < nz and usr[x, y, z] == 1:
                out += 2.0 * f[q, x, y, z] * E[q]
    return out


def moments(f):
    p = np.zeros(f.shape[1:])
    u = np.zeros_like(p)
    v = np.zeros_like(p)
    w = np.zeros_like(p)
    for q in range(Q):
        p = p + f[q] / 3.0
        u = u + f[q] * E[q, 0]
        v = v + f[q] * E[q, 1]
        w = w + f[q] * E[q, 2]
    return p, u, v, w


class NpModel:
    def __init__(self, case_id, nx, ny, nz, tau0, rho0=1.0, u0=0.05, mode=0, lid_u=0.1,
                 sphere=(0.0, 0.0, 0.0, 0.0), seed=5374857):
        self.case, self.mode = case_id, mode
        self.n = (nx, ny, nz)
        self.tau0, self.rho0, self.u0, self.lid_u = tau0, rho0, u0, lid_u
        self.sphere, self.seed = sphere, seed
        # faithful cavity/sphere keep the reference's ghost-3 array; the rest
        # live on the interior only
        self.ghosted = case_id in (1, 2) and mode == 0

    # ------------------------------------------------------------ init
    def initialize(self):
        nx, ny, nz = self.n
        if self.ghosted:
            lx, ly, lz = nx + 2 * G, ny + 2 * G, nz + 2 * G
            i, j, k = np.meshgrid(np.arange(lx), np.arange(ly), np.arange(lz), indexing="ij")
            gi, gj, gk = i - G, j - G, k - G
        else:
            gi, gj, gk = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
        shape = gi.shape
        self.usr = np.ones(shape, dtype=np.int64)
        if self.case == 0:
            ax = gi / nx * 2.0 * PI_R
            ay = gj / ny * 2.0 * PI_R
            az = gk / nz * 2.0 * PI_R
            u = self.u0 * np.sin(ax) * np.cos(ay) * np.cos(az)
            v = -self.u0 * np.cos(ax) * np.sin(ay) * np.cos(az)
            w = np.zeros(shape)
            p = self.rho0 / 3.0 + self.rho0 * self.u0 * self.u0 / 16.0 * (
                np.cos(2.0 * ax) + np.cos(2.0 * ay)) * (np.cos(2.0 * az) + 2.0)
        elif self.case == 1:
            u = np.zeros(shape)
            v = np.zeros(shape)
            w = np.zeros(shape)
            p = np.full(shape, 0.0 if self.mode == 0 else self.rho0 / 3.0)
        else:
            v = np.zeros(shape)
            w = np.zeros(shape)
            if self.mode == 0:
                lx, ly, lz = shape
                idx = (i.astype(np.uint64) * np.uint64(ly) + j.astype(np.uint64)) * np.uint64(lz) \
                    + k.astype(np.uint64)
                ran = (_mix64(np.uint64(self.seed) + idx * np.uint64(0x9E3779B97F4A7C15))
                       % np.uint64(100)).astype(np.int64)
                c = (nx - 1) // 8
                r = (nx - 1) / 64.0
                d2 = (gi - c) ** 2 + (gj - c) ** 2 + (gk - c) ** 2
                self.usr = np.where(d2 <= r * r, 0, 1)
                u = (self.u0 * 4.0 * gk * (nz - 1 - gk) / float((nz - 1) ** 2)) * self.usr \
                    * (1.0 + ran * 0.001)
                p = np.zeros(shape)
            else:
                cx, cy, cz, rad = self.sphere
                d2 = (gi - cx) ** 2 + (gj - cy) ** 2 + (gk - cz) ** 2
                self.usr = np.where(d2 <= rad * rad, 0, 1)
                u = np.where(self.usr == 1, self.u0, 0.0)
                p = np.full(shape, self.rho0 / 3.0)
        self.p, self.u, self.v, self.w = p, u, v, w
        self.f = feq(p, u, v, w)

    # ------------------------------------------------------------ phases
    def _interior(self):
        if self.ghosted:
            nx, ny, nz = self.n
            return (slice(G, G + nx), slice(G, G + ny), slice(G, G + nz))
        return (slice(None),) * 3

    def collide(self):
        s = self._interior()
        fe = feq(self.p[s], self.u[s], self.v[s], self.w[s])
        fi = self.f[(slice(None),) + s]
        post = fi - (fi - fe) / (self.tau0 + 0.5)
        if self.case == 2 and self.mode == 1:
            post = np.where(self.usr[s][None] == 1, post, fi)
        self.f[(slice(None),) + s] = post

    def step(self):
        self.collide()
        if self.case == 0:
            self._stream_periodic()
        elif self.case == 1 and self.mode == 1:
            self._stream_cavity_config()
        elif self.case == 2 and self.mode == 1:
            self._stream_sphere_config()
        else:
            self._faithful_walls()
            self._stream_ghosted()
        self._update()

    def _stream_periodic(self):
        f = self.f
        for q in range(Q):
            f[q] = np.roll(f[q], shift=tuple(int(t) for t in E[q]), axis=(0, 1, 2))

    def _pull_padded(self, fpost, fill):
        """Pull from a one-cell padded copy; returns (pulled, source-was-pad masks)."""
        nx, ny, nz = self.n
        out = np.empty_like(fpost)
        pad = np.full((nx + 2, ny + 2, nz + 2), fill)
        for q in range(Q):
            pad[1:-1, 1:-1, 1:-1] = fpost[q]
            ex, ey, ez = E[q]
            out[q] = pad[1 - ex:1 - ex + nx, 1 - ey:1 - ey + ny, 1 - ez:1 - ez + nz]
        return out

    def _stream_cavity_config(self):
        nx, ny, nz = self.n
        fpost = self.f
        pulled = self._pull_padded(fpost, np.nan)
        X, Y, Z = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
        for q in range(Q):
            sx, sy, sz = X - E[q, 0], Y - E[q, 1], Z - E[q, 2]
            out = (sx < 0) | (sx >= nx) | (sy < 0) | (sy >= ny) | (sz < 0) | (sz >= nz)
            corr = lid_term(q, self.rho0, self.lid_u)
            bounce = np.where(sz >= nz, fpost[BB[q]] + corr, fpost[BB[q]])
            pulled[q] = np.where(out, bounce, pulled[q])
        self.f = pulled

    def _stream_sphere_config(self):
        nx, ny, nz = self.n
        fpost = self.f
        fin = feq(self.rho0 / 3.0, self.u0, 0.0, 0.0)
        X = np.arange(nx)[:, None, None] * np.ones((1, ny, nz), dtype=np.int64)
        out = np.empty_like(fpost)
        for q in range(Q):
            ex, ey, ez = (int(t) for t in E[q])
            g = np.roll(fpost[q], shift=(0, ey, ez), axis=(0, 1, 2))  # y/z periodic
            pulled = np.empty_like(g)
            if ex == 1:
                pulled[1:] = g[:-1]
                pulled[0] = fin[q]
            elif ex == -1:
                pulled[:-1] = g[1:]
                pulled[-1] = g[-1]  # zero-gradient: ghost = copy of last layer
            else:
                pulled[:] = g
            # halfway bounce-back where the source is solid (source x inside)
            solid_src = np.roll(self.usr, shift=(ex, ey, ez), axis=(0, 1, 2)) == 0
            sx = X - ex
            solid_src &= (sx >= 0) & (sx < nx)
            pulled = np.where(solid_src, fpost[BB[q]], pulled)
            out[q] = np.where(self.usr == 1, pulled, fpost[q])
        self.f = out

    def _faithful_walls(self):
        nx, ny, nz = self.n
        f = self.f
        lx, ly, lz = nx + 2 * G, ny + 2 * G, nz + 2 * G
        lo, hi = G, (G + nx, G + ny, G + nz)
        if self.case == 1:
            # Q10: one-rank unpack writes never-filled buffers (zeros)
            f[:, :, hi[1], :] = 0.0
            f[:, :, :, hi[2]] = 0.0
            f[:, lo - 1, :, :] = 0.0
            f[:, hi[0], :, :] = 0.0
        a, b = G - 1, None  # index ranges [G-1, l-G+1)
        rj = np.arange(G - 1, ly - G + 1)
        rk = np.arange(G - 1, lz - G + 1)
        ri = np.arange(G - 1, lx - G + 1)

        def bc(axis, plane, src_plane, sel, step_other=(2, 2, 2)):
            for q in range(Q):
                if not sel(E[q]):
                    continue
                others = [d for d in range(3) if d != axis]
                rng = {0: ri, 1: rj, 2: rk}
                A = rng[others[0]]
                B = rng[others[1]]
                AA, BBi = np.meshgrid(A, B, indexing="ij")
                dst = [None] * 3
                src = [None] * 3
                dst[axis] = plane
                src[axis] = src_plane
                dst[others[0]], dst[others[1]] = AA, BBi
                src[others[0]] = AA + step_other[others[0]] * E[q, others[0]]
                src[others[1]] = BBi + step_other[others[1]] * E[q, others[1]]
                f[q][tuple(dst)] = f[BB[q]][tuple(src)]

        if self.case == 1:
            bc(0, lo - 1, lo + 1, lambda e: e[0] > 0)
            bc(0, hi[0], hi[0] - 2, lambda e: e[0] < 0)
            bc(1, lo - 1, lo + 1, lambda e: e[1] > 0)
            bc(1, hi[1], hi[1] - 2, lambda e: e[1] < 0)
            bc(2, lo - 1, lo + 1, lambda e: e[2] > 0)
            bc(2, hi[2], hi[2] - 2, lambda e: e[2] < 0, step_other=(2, 1, 2))
        else:
            bc(1, lo - 1, lo + 1, lambda e: e[1] > 0)
            bc(1, hi[1], hi[1] - 2, lambda e: e[1] < 0)
            bc(2, lo - 1, lo + 1, lambda e: e[2] > 0)
            bc(2, hi[2], hi[2] - 2, lambda e: e[2] < 0)
            J, K = np.meshgrid(rj, rk, indexing="ij")
            for q in range(Q):
                if E[q, 0] > 0:
                    f[q, G - 1][J, K] = f[q, G][J, K]
            for q in range(Q):
                if E[q, 0] < 0:
                    f[q, hi[0]][J, K] = f[q, hi[0] + E[q, 0]][J + E[q, 1], K + E[q, 2]]
            usr = self.usr
            I3, J3, K3 = np.meshgrid(np.arange(G, G + nx), np.arange(G, G + ny),
                                     np.arange(G, G + nz), indexing="ij")
            for q in range(Q):
                ex, ey, ez = (int(t) for t in E[q])
                cond = (usr[I3, J3, K3] == 0) & (usr[I3 + ex, J3 + ey, K3 + ez] == 1)
                if not cond.any():
                    continue
                ii, jj, kk = I3[cond], J3[cond], K3[cond]
                f[q, ii, jj, kk] = f[BB[q], ii + 2 * ex, jj + 2 * ey, kk + 2 * ez]

    def _stream_ghosted(self):
        nx, ny, nz = self.n
        f = self.f
        new = f.copy()
        for q in range(Q):
            ex, ey, ez = E[q]
            new[q, G:G + nx, G:G + ny, G:G + nz] = \
                f[q, G - ex:G - ex + nx, G - ey:G - ey + ny, G - ez:G - ez + nz]
        self.f = new

    def _update(self):
        s = self._interior()
        p, u, v, w = moments(self.f[(slice(None),) + s])
        if self.case == 2:
            m = self.usr[s]
            p, u, v, w = p * m, u * m, v * m, w * m
        if self.mode == 0 and self.case == 1:
            u[:, :, -1] = 0.1
        if self.mode == 0 and self.case == 2:
            nz = self.n[2]
            kz = np.arange(nz)
            u[0, :, :] = (self.u0 * 4.0 * kz * (nz - 1 - kz) / float((nz - 1) ** 2))[None, :]
            v[0] = 0.0
            w[0] = 0.0
            p[-1] = 0.0
        if self.ghosted:
            self.p[s], self.u[s], self.v[s], self.w[s] = p, u, v, w
        else:
            self.p, self.u, self.v, self.w = p, u, v, w

    # ------------------------------------------------------------ access
    def run(self, nsteps):
        for _ in range(nsteps):
            self.step()

    def moments(self):
        s = self._interior()
        return [np.ascontiguousarray(a[s].transpose(2, 1, 0)) for a in (self.p, self.u, self.v, self.w)]

    def get_f(self):
        s = self._interior()
        return np.ascontiguousarray(self.f[(slice(None),) + s].transpose(0, 3, 2, 1))
