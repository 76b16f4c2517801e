"""Python mirror of the reference's host interface over the C-ABI.

``LBM`` keeps the reference's method names and call order
(tgv/src/lbm.hpp:238-251, driven as in tgv/main.cpp:70-119); ``System``
parses ``input.in`` as tgv/src/System.cpp:5-24 does (per case);
``comm_layout`` is CommHelper (tgv/src/lbm.hpp:34-117).  Every call goes to
libimexlb.so; a failing status raises ``ImexlbError``.
"""
import ctypes

import numpy as np

from . import _lib as L
from ._lib import ImexlbError, as_f64, dptr, lib

CASES = {"tgv": L.CASE_TGV, "cavity": L.CASE_CAVITY, "ldc": L.CASE_CAVITY, "sphere": L.CASE_SPHERE}
MODES = {"faithful": L.MODE_FAITHFUL, "config": L.MODE_CONFIG}
ENGINES = {"fused": L.ENGINE_FUSED, "phased": L.ENGINE_PHASED}
DTYPES = {"f64": L.F64, "f32": L.F32, "f16": L.F16}
STORAGE = {"pull": L.PULL, "aa": L.AA}


def _code(table, v):
    return table[v] if isinstance(v, str) else int(v)


def check(status, handle=None):
    if status != L.OK:
        msg = lib().imexlb_last_error(handle)
        msg = msg.decode() if msg else lib().imexlb_strerror(status).decode()
        raise ImexlbError(status, msg)


class LBM:
    """One subdomain (one rank) of the reference's struct LBM, on the GPU."""

    def __init__(self, sx, sy, sz, tau, rho0=1.0, u0=0.05, case="tgv", mode="faithful",
                 engine="fused", dtype="f64", storage="pull", lid_u=0.1, sphere=None,
                 seed=5374857, block=0, cells_per_thread=0, xcd_remap=-1, device=0, stride_pad=-1,
                 nt_loads=-1, addr_mode=-1):
        c = L.Config()
        check(lib().imexlb_config_init(ctypes.byref(c)))
        c.case_id = _code(CASES, case)
        c.mode = _code(MODES, mode)
        c.engine = _code(ENGINES, engine)
        c.dtype = _code(DTYPES, dtype)
        c.storage = _code(STORAGE, storage)
        c.nx, c.ny, c.nz = int(sx), int(sy), int(sz)
        c.tau0, c.rho0, c.u0, c.lid_u = float(tau), float(rho0), float(u0), float(lid_u)
        if sphere is not None:
            c.sphere_cx, c.sphere_cy, c.sphere_cz, c.sphere_r = (float(v) for v in sphere)
        c.seed = int(seed)
        c.block, c.cells_per_thread, c.xcd_remap, c.device = int(block), int(cells_per_thread), \
            int(xcd_remap), int(device)
        c.stride_pad = int(stride_pad)
        c.nt_loads = int(nt_loads)
        c.addr_mode = int(addr_mode)
        self.cfg = c
        self.shape = (c.nz, c.ny, c.nx)
        h = ctypes.c_void_p()
        check(lib().imexlb_create(ctypes.byref(c), ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().imexlb_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _chk(self, s):
        check(s, self.h)

    # ---- the reference's surface (tgv/src/lbm.hpp:238-246)
    def Initialize(self):
        self._chk(lib().imexlb_initialize(self.h))

    def setup_subdomain(self):
        self._chk(lib().imexlb_setup_subdomain(self.h))

    def Collision(self):
        self._chk(lib().imexlb_collide(self.h))

    def pack(self):
        self._chk(lib().imexlb_pack(self.h))

    def exchange(self):
        self._chk(lib().imexlb_exchange(self.h))

    def unpack(self):
        self._chk(lib().imexlb_unpack(self.h))

    def Streaming(self):
        self._chk(lib().imexlb_stream(self.h))

    def Update(self):
        self._chk(lib().imexlb_update(self.h))

    # ---- fast path and data access
    def step(self, n, moments_every=None):
        if moments_every is None:
            moments_every = max(1, int(n))
        self._chk(lib().imexlb_step(self.h, int(n), int(moments_every)))

    def step_timed(self, n, moments_every=0):
        ms = ctypes.c_float()
        self._chk(lib().imexlb_step_timed(self.h, int(n), int(moments_every), ctypes.byref(ms)))
        return ms.value

    def sync(self):
        self._chk(lib().imexlb_sync(self.h))

    def moments(self):
        out = [np.empty(self.shape) for _ in range(4)]
        self._chk(lib().imexlb_get_moments(self.h, *[dptr(a) for a in out]))
        return out

    def macro(self):
        """rho = 3 p (= sum f), u, v, w of the last step that wrote moments."""
        out = [np.empty(self.shape) for _ in range(4)]
        self._chk(lib().imexlb_get_macro(self.h, *[dptr(a) for a in out]))
        return out

    def get_f(self):
        out = np.empty((L.Q,) + self.shape)
        self._chk(lib().imexlb_get_f(self.h, dptr(out)))
        return out

    def set_f(self, f):
        f = as_f64(f)
        if f.shape != (L.Q,) + self.shape:
            raise ValueError(f"set_f expects shape {(L.Q,) + self.shape}")
        self._chk(lib().imexlb_set_f(self.h, dptr(f)))

    def diagnostics(self):
        ke, mass = ctypes.c_double(), ctypes.c_double()
        self._chk(lib().imexlb_diagnostics(self.h, ctypes.byref(ke), ctypes.byref(mass)))
        return ke.value, mass.value

    def sphere_force(self):
        out = np.zeros(3)
        self._chk(lib().imexlb_sphere_force(self.h, dptr(out)))
        return out

    def checkpoint(self, path):
        self._chk(lib().imexlb_checkpoint_write(self.h, str(path).encode()))

    def restart(self, path):
        self._chk(lib().imexlb_checkpoint_read(self.h, str(path).encode()))

    def MPIoutput(self, n, directory="."):
        """Write output<n>.plt (Tecplot, x y z u v w p) as the reference's
        LBM::MPIoutput does (tgv/src/lbm.cpp:335-669)."""
        import os
        p, u, v, w = self.moments()
        path = os.path.join(directory, f"output{int(n)}.plt")
        tecplot_write(path, float(n), p, u, v, w)
        return path

    def autotune(self, cache_path=None, steps=10):
        """Pick the fastest launch shape (results are shape-independent); the
        handle is re-initialised afterwards.  Returns (block, cpt, xcd, ntl)."""
        ch = (ctypes.c_int32 * 4)()
        self._chk(lib().imexlb_autotune(self.h, cache_path.encode() if cache_path else None, int(steps), ch))
        return tuple(ch)

    def set_launch(self, block, cells_per_thread=1, xcd_order=1, nt_loads=0):
        """Launch shape of the fused kernels (imexlb_set_launch): block 64..512
        in steps of 64; cells per thread 1, 2 (blockDim apart) or 3 (two
        adjacent x cells, periodic lattices with nx even; one cell elsewhere);
        XCD tile order 0/1; non-temporal loads 0/1.  Results do not depend on it."""
        self._chk(lib().imexlb_set_launch(self.h, int(block), int(cells_per_thread), int(xcd_order), int(nt_loads)))

    def tune_placement(self, trials=6, steps=10):
        """Time the step on up to `trials` lattice allocations (each with its
        own HBM pages) and keep the fastest; the state is re-initialised.
        Returns (index kept, [ms per step of each candidate, 0 if none])."""
        ms = (ctypes.c_float * int(trials))()
        ch = ctypes.c_int32()
        self._chk(lib().imexlb_tune_placement(self.h, int(trials), int(steps), ms, ctypes.byref(ch)))
        return ch.value, list(ms)

    def tune_placement_window(self, slack_gib=32, step_gib=4, steps=10):
        """Time the step with the lattice at offsets 0, step, 2 step .. slack
        inside one address range of the lattice plus `slack_gib` GiB and keep
        the fastest; the chunks of the range the lattice does not touch are
        given back.  Returns (index kept, [ms per step at each offset])."""
        step = int(step_gib * (1 << 30))
        slack = int(slack_gib * (1 << 30)) // step * step
        n = slack // step + 1
        ms = (ctypes.c_float * n)()
        ch = ctypes.c_int32()
        self._chk(lib().imexlb_tune_placement_window(self.h, slack, step, int(steps), ms, n, ctypes.byref(ch)))
        return ch.value, list(ms)

    def info(self):
        i = L.Info()
        self._chk(lib().imexlb_get_info(self.h, ctypes.byref(i)))
        return {k: getattr(i, k) for k, _ in L.Info._fields_}


class System:
    """input.in parser (tgv/src/System.cpp:5-24 and the per-case copies)."""

    def __init__(self, path="input.in", case="tgv", keep_hardcoded=True):
        s = L.SystemC()
        check(lib().imexlb_system_read(str(path).encode(), _code(CASES, case), int(bool(keep_hardcoded)),
                                       ctypes.byref(s)))
        for k, _ in L.SystemC._fields_:
            setattr(self, k, getattr(s, k))


def comm_layout(nranks, me, periodic, glx, gly, glz):
    c = L.Comm()
    check(lib().imexlb_comm_layout(int(nranks), int(me), int(bool(periodic)), int(glx), int(gly), int(glz),
                                   ctypes.byref(c)))
    return {"nranks": c.nranks, "me": c.me, "rx": c.rx, "ry": c.ry, "rz": c.rz, "px": c.px, "py": c.py,
            "pz": c.pz, "neighbours": list(c.neighbours), "l_l": list(c.l_l), "lo": list(c.lo),
            "hi": list(c.hi)}


def tecplot_write(path, solution_time, p, u, v, w):
    """Tecplot #!TDV112 file of x y z u v w p; fields shaped [z][y][x]."""
    arrs = [as_f64(a) for a in (p, u, v, w)]
    nz, ny, nx = arrs[0].shape
    check(lib().imexlb_tecplot_write(str(path).encode(), float(solution_time), nx, ny, nz,
                                     *[dptr(a) for a in arrs]))


def copy_bandwidth(nbytes, reps=20, device=0):
    g, ms = ctypes.c_double(), ctypes.c_float()
    check(lib().imexlb_copy_bandwidth(int(device), int(nbytes), int(reps), ctypes.byref(g), ctypes.byref(ms)))
    return g.value, ms.value


BANDWIDTH_PROBES = ("copy", "copy_nt", "copy_x4", "copy_x4_nt", "copy_gridstride", "copy_gridstride_nt",
                    "read_x4", "read_x4_nt", "write_x4", "write_x4_nt")


def bandwidth_probe(kind, nbytes, reps=20, device=0):
    """GB/s (every byte read and written) and ms per launch of one of the
    BANDWIDTH_PROBES stream kernels over nbytes bytes per stream."""
    k = BANDWIDTH_PROBES.index(kind) if isinstance(kind, str) else int(kind)
    g, ms = ctypes.c_double(), ctypes.c_float()
    check(lib().imexlb_bandwidth_probe(int(device), k, int(nbytes), int(reps), ctypes.byref(g), ctypes.byref(ms)))
    return g.value, ms.value


class Ensemble:
    """Many independent TGV / config-mode cavity domains stepped together by
    one persistent launch per step (optionally replayed from a HIP graph)."""

    MODES = {"persistent": L.ENS_PERSISTENT, "graph": L.ENS_GRAPH, "per_domain": L.ENS_PER_DOMAIN}

    def __init__(self, domains, dtype="f32", storage="pull", device=0):
        n = len(domains)
        arr = (L.EnsembleDomain * n)()
        for k, d in enumerate(domains):
            arr[k] = d if isinstance(d, L.EnsembleDomain) else L.EnsembleDomain(**d)
        self.domains = [arr[k] for k in range(n)]
        h = ctypes.c_void_p()
        check(lib().imexlb_ensemble_create(arr, n, _code(DTYPES, dtype), _code(STORAGE, storage), int(device),
                                           ctypes.byref(h)))
        self.h = h

    @staticmethod
    def generate(seed, n):
        arr = (L.EnsembleDomain * n)()
        check(lib().imexlb_ensemble_generate(int(seed), int(n), arr))
        return [arr[k] for k in range(n)]

    def _chk(self, s):
        if s != L.OK:
            msg = lib().imexlb_ensemble_last_error(self.h)
            raise ImexlbError(s, (msg.decode() if msg else "") or lib().imexlb_strerror(s).decode())

    def close(self):
        if getattr(self, "h", None):
            lib().imexlb_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def Initialize(self):
        self._chk(lib().imexlb_ensemble_initialize(self.h))

    def step(self, n, mode="graph", moments_every=None):
        if moments_every is None:
            moments_every = max(1, int(n))
        self._chk(lib().imexlb_ensemble_step(self.h, int(n), self.MODES[mode], int(moments_every)))

    def step_timed(self, n, mode="graph", moments_every=0):
        ms = ctypes.c_float()
        self._chk(lib().imexlb_ensemble_step_timed(self.h, int(n), self.MODES[mode], int(moments_every),
                                                   ctypes.byref(ms)))
        return ms.value

    def moments(self, d):
        dom = self.domains[d]
        shape = (dom.nz, dom.ny, dom.nx)
        out = [np.empty(shape) for _ in range(4)]
        self._chk(lib().imexlb_ensemble_get_moments(self.h, int(d), *[dptr(a) for a in out]))
        return out

    def status(self):
        """Per-domain health after a step that wrote moments: a list of
        booleans, True while the domain's moments are all finite."""
        n = len(self.domains)
        fin = (ctypes.c_int32 * n)()
        bad = ctypes.c_int32()
        self._chk(lib().imexlb_ensemble_status(self.h, fin, ctypes.byref(bad)))
        out = [bool(v) for v in fin]
        assert out.count(False) == bad.value
        return out

    def set_launch(self, block, cpt, grid):
        self._chk(lib().imexlb_ensemble_set_launch(self.h, int(block), int(cpt), int(grid)))

    def autotune(self, cache_path=None, steps=6):
        ch = (ctypes.c_int32 * 3)()
        self._chk(lib().imexlb_ensemble_autotune(self.h, cache_path.encode() if cache_path else None,
                                                 int(steps), ch))
        return tuple(ch)

    def tune_placement(self, trials=6, steps=6):
        """Time graph-replayed steps on up to `trials` allocations of the
        ensemble's lattices and keep the fastest; the ensemble is reset to
        its initial state.  Returns (index kept, [ms per step of each
        candidate, 0 if none])."""
        ms = (ctypes.c_float * int(trials))()
        ch = ctypes.c_int32()
        self._chk(lib().imexlb_ensemble_tune_placement(self.h, int(trials), int(steps), ms, ctypes.byref(ch)))
        return ch.value, list(ms)

    def tune_placement_window(self, slack_gib=32, step_gib=1, steps=6):
        """As LBM.tune_placement_window, for the pool that holds every
        member's lattices; the ensemble is reset to its initial state."""
        step = int(step_gib * (1 << 30))
        slack = int(slack_gib * (1 << 30)) // step * step
        n = slack // step + 1
        ms = (ctypes.c_float * n)()
        ch = ctypes.c_int32()
        self._chk(lib().imexlb_ensemble_tune_placement_window(self.h, slack, step, int(steps), ms, n,
                                                              ctypes.byref(ch)))
        return ch.value, list(ms)

    def info(self):
        out = (ctypes.c_int64 * 7)()
        self._chk(lib().imexlb_ensemble_info(self.h, out))
        keys = ("cells", "tiles", "grid", "block", "cells_per_thread", "device_bytes", "steps")
        return dict(zip(keys, list(out)))
