"""ctypes loader for the CPU oracle (oracle/imexlb_oracle.c).

Test infrastructure only: the tests, ``__graft_entry__.smoke()`` and the
CPU-baseline leg of ``bench.py`` use it as the checker.  The product package
``imexlb_amd`` never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "libimexlb_oracle.so")

TGV, LDC, SPHERE = 0, 1, 2
FAITHFUL, CONFIG = 0, 1
Q = 27
GHOST = 3


class OrcConfig(ctypes.Structure):
    _fields_ = [
        ("case_id", ctypes.c_int),
        ("mode", ctypes.c_int),
        ("nx", ctypes.c_int),
        ("ny", ctypes.c_int),
        ("nz", ctypes.c_int),
        ("tau0", ctypes.c_double),
        ("rho0", ctypes.c_double),
        ("u0", ctypes.c_double),
        ("lid_u", ctypes.c_double),
        ("sph_cx", ctypes.c_double),
        ("sph_cy", ctypes.c_double),
        ("sph_cz", ctypes.c_double),
        ("sph_r", ctypes.c_double),
        ("seed", ctypes.c_ulonglong),
        ("threads", ctypes.c_int),
    ]


_lib = None


def build():
    """Compile the oracle with its Makefile (gcc only)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def build_native(out_dir):
    """Compile the same source for this host (-march=native) into out_dir:
    the timing build of the CPU baseline.  Returns the library's path."""
    subprocess.run(["make", "-s", "-C", HERE, "native", f"NATIVE_OUT={out_dir}"], check=True,
                   stdout=subprocess.DEVNULL, stderr=subprocess.PIPE)
    return os.path.join(out_dir, "libimexlb_oracle.so")


def use_library(path):
    """Load the oracle from `path` instead of oracle/build/ (before the first
    use in this process)."""
    global LIB_PATH
    if _lib is not None and path != LIB_PATH:
        raise RuntimeError("oracle: a library is already loaded")
    LIB_PATH = path


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        vp, dp, ip = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
        L.orc_create.restype = vp
        L.orc_create.argtypes = [ctypes.POINTER(OrcConfig)]
        for name in ("orc_free", "orc_initialize", "orc_collision", "orc_pack", "orc_exchange",
                     "orc_unpack", "orc_streaming", "orc_update"):
            getattr(L, name).argtypes = [vp]
            getattr(L, name).restype = None
        L.orc_run.argtypes = [vp, ctypes.c_int]
        L.orc_dims.argtypes = [vp, ip]
        L.orc_get_moments.argtypes = [vp, dp, dp, dp, dp]
        L.orc_get_f.argtypes = [vp, dp]
        L.orc_set_f.argtypes = [vp, dp]
        L.orc_get_f_full.argtypes = [vp, dp]
        L.orc_set_f_full.argtypes = [vp, dp]
        L.orc_get_fields_full.argtypes = [vp, dp, dp, dp, dp, ip]
        L.orc_set_fields_full.argtypes = [vp, dp, dp, dp, dp]
        L.orc_diagnostics.argtypes = [vp, dp, dp]
        L.orc_sphere_force.argtypes = [vp, dp]
        L.orc_tables.argtypes = [ip, ip, dp]
        L.orc_sphere_ran.argtypes = [ctypes.c_ulonglong, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_int]
        L.orc_sphere_ran.restype = ctypes.c_int
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def tables():
    e = np.zeros(Q * 3, dtype=np.int32)
    bb = np.zeros(Q, dtype=np.int32)
    w = np.zeros(Q, dtype=np.float64)
    lib().orc_tables(e.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                     bb.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _dp(w))
    return e.reshape(Q, 3), bb, w


class Oracle:
    """One subdomain of the reference's LBM struct, one rank, on the CPU."""

    def __init__(self, case_id, nx, ny, nz, tau0, rho0=1.0, u0=0.05, mode=FAITHFUL, lid_u=0.1,
                 sphere=(0.0, 0.0, 0.0, 0.0), seed=5374857, threads=0):
        self.cfg = OrcConfig(case_id, mode, nx, ny, nz, tau0, rho0, u0, lid_u,
                             sphere[0], sphere[1], sphere[2], sphere[3], seed, threads)
        self.h = lib().orc_create(ctypes.byref(self.cfg))
        if not self.h:
            raise ValueError("oracle: invalid configuration")
        self.nx, self.ny, self.nz = nx, ny, nz
        d = np.zeros(4, dtype=np.int32)
        lib().orc_dims(self.h, d.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        self.lx, self.ly, self.lz = int(d[0]), int(d[1]), int(d[2])

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_free(self.h)
            self.h = None

    # the reference's method names (tgv/src/lbm.hpp:238-251)
    def Initialize(self):
        lib().orc_initialize(self.h)

    def Collision(self):
        lib().orc_collision(self.h)

    def pack(self):
        lib().orc_pack(self.h)

    def exchange(self):
        lib().orc_exchange(self.h)

    def unpack(self):
        lib().orc_unpack(self.h)

    def Streaming(self):
        lib().orc_streaming(self.h)

    def Update(self):
        lib().orc_update(self.h)

    def run(self, nsteps):
        lib().orc_run(self.h, int(nsteps))

    def moments(self):
        n = self.nx * self.ny * self.nz
        out = [np.empty(n) for _ in range(4)]
        lib().orc_get_moments(self.h, *[_dp(a) for a in out])
        return [a.reshape(self.nz, self.ny, self.nx) for a in out]

    def get_f(self):
        out = np.empty(Q * self.nx * self.ny * self.nz)
        lib().orc_get_f(self.h, _dp(out))
        return out.reshape(Q, self.nz, self.ny, self.nx)

    def set_f(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        lib().orc_set_f(self.h, _dp(f))

    def get_f_full(self):
        out = np.empty(Q * self.lx * self.ly * self.lz)
        lib().orc_get_f_full(self.h, _dp(out))
        return out.reshape(Q, self.lx, self.ly, self.lz)

    def set_f_full(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        lib().orc_set_f_full(self.h, _dp(f))

    def fields_full(self):
        n = self.lx * self.ly * self.lz
        p, u, v, w = (np.empty(n) for _ in range(4))
        usr = np.empty(n, dtype=np.int32)
        lib().orc_get_fields_full(self.h, _dp(p), _dp(u), _dp(v), _dp(w),
                                  usr.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        sh = (self.lx, self.ly, self.lz)
        return p.reshape(sh), u.reshape(sh), v.reshape(sh), w.reshape(sh), usr.reshape(sh)

    def set_fields_full(self, p, u, v, w):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (p, u, v, w)]
        lib().orc_set_fields_full(self.h, *[_dp(a) for a in arrs])

    def diagnostics(self):
        ke, mass = ctypes.c_double(), ctypes.c_double()
        lib().orc_diagnostics(self.h, ctypes.byref(ke), ctypes.byref(mass))
        return ke.value, mass.value

    def sphere_force(self):
        out = np.zeros(3)
        lib().orc_sphere_force(self.h, _dp(out))
        return out
