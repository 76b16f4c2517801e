"""ctypes binding of libimexlb.so (include/imexlb.h).

The library is loaded from imexlb_amd/lib/ only; there is no fallback.  A
missing library raises at import of the binding.
"""
import ctypes
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("IMEXLB_LIB") or os.path.join(PKG, "lib", "libimexlb.so")

Q = 27
OK, EINVAL, EHIP, ENOMEM, EUNSUPPORTED, ESTATE, EIO, ENODEV = 0, -1, -2, -3, -4, -5, -6, -7
CASE_TGV, CASE_CAVITY, CASE_SPHERE = 0, 1, 2
ENS_PERSISTENT, ENS_GRAPH, ENS_PER_DOMAIN = 0, 1, 2
MODE_FAITHFUL, MODE_CONFIG = 0, 1
ENGINE_FUSED, ENGINE_PHASED = 0, 1
F64, F32, F16 = 0, 1, 2
PULL, AA = 0, 1


class Config(ctypes.Structure):
    _fields_ = [
        ("case_id", ctypes.c_int32), ("mode", ctypes.c_int32), ("engine", ctypes.c_int32),
        ("dtype", ctypes.c_int32), ("storage", ctypes.c_int32),
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("tau0", ctypes.c_double), ("rho0", ctypes.c_double), ("u0", ctypes.c_double),
        ("lid_u", ctypes.c_double),
        ("sphere_cx", ctypes.c_double), ("sphere_cy", ctypes.c_double),
        ("sphere_cz", ctypes.c_double), ("sphere_r", ctypes.c_double),
        ("seed", ctypes.c_uint64),
        ("block", ctypes.c_int32), ("cells_per_thread", ctypes.c_int32),
        ("xcd_remap", ctypes.c_int32), ("device", ctypes.c_int32),
        ("stride_pad", ctypes.c_int32), ("nt_loads", ctypes.c_int32),
        ("addr_mode", ctypes.c_int32),
    ]


class Info(ctypes.Structure):
    _fields_ = [("cells", ctypes.c_int64), ("steps", ctypes.c_int64),
                ("bytes_per_lup", ctypes.c_double), ("device_bytes", ctypes.c_int64),
                ("block", ctypes.c_int32), ("cells_per_thread", ctypes.c_int32),
                ("xcd_remap", ctypes.c_int32), ("nt_loads", ctypes.c_int32),
                ("addr_mode", ctypes.c_int32)]


class SystemC(ctypes.Structure):
    _fields_ = [("rho0", ctypes.c_double), ("R", ctypes.c_double), ("Re", ctypes.c_double),
                ("u0", ctypes.c_double), ("Time", ctypes.c_int32), ("inter", ctypes.c_int32),
                ("sx", ctypes.c_int32), ("sy", ctypes.c_int32), ("sz", ctypes.c_int32),
                ("cs2", ctypes.c_double), ("cs", ctypes.c_double), ("Ma", ctypes.c_double),
                ("miu", ctypes.c_double), ("tau", ctypes.c_double)]


class Comm(ctypes.Structure):
    _fields_ = [("nranks", ctypes.c_int32), ("me", ctypes.c_int32), ("rx", ctypes.c_int32),
                ("ry", ctypes.c_int32), ("rz", ctypes.c_int32), ("px", ctypes.c_int32),
                ("py", ctypes.c_int32), ("pz", ctypes.c_int32),
                ("neighbours", ctypes.c_int32 * 26), ("l_l", ctypes.c_int32 * 3),
                ("lo", ctypes.c_int32 * 3), ("hi", ctypes.c_int32 * 3)]


class EnsembleDomain(ctypes.Structure):
    _fields_ = [("case_id", ctypes.c_int32), ("nx", ctypes.c_int32), ("ny", ctypes.c_int32),
                ("nz", ctypes.c_int32), ("tau0", ctypes.c_double), ("rho0", ctypes.c_double),
                ("u0", ctypes.c_double), ("lid_u", ctypes.c_double)]


# (name, restype, argtypes) of every entry point in include/imexlb.h
_vp, _dp, _fp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_float)
_i = ctypes.c_int
SIGNATURES = [
    ("imexlb_config_init", _i, [ctypes.POINTER(Config)]),
    ("imexlb_create", _i, [ctypes.POINTER(Config), ctypes.POINTER(_vp)]),
    ("imexlb_destroy", _i, [_vp]),
    ("imexlb_initialize", _i, [_vp]),
    ("imexlb_setup_subdomain", _i, [_vp]),
    ("imexlb_collide", _i, [_vp]),
    ("imexlb_pack", _i, [_vp]),
    ("imexlb_exchange", _i, [_vp]),
    ("imexlb_unpack", _i, [_vp]),
    ("imexlb_stream", _i, [_vp]),
    ("imexlb_update", _i, [_vp]),
    ("imexlb_step", _i, [_vp, _i, _i]),
    ("imexlb_step_timed", _i, [_vp, _i, _i, _fp]),
    ("imexlb_get_moments", _i, [_vp, _dp, _dp, _dp, _dp]),
    ("imexlb_get_macro", _i, [_vp, _dp, _dp, _dp, _dp]),
    ("imexlb_get_f", _i, [_vp, _dp]),
    ("imexlb_set_f", _i, [_vp, _dp]),
    ("imexlb_diagnostics", _i, [_vp, _dp, _dp]),
    # device pointers as integers: p, u, v, w, ncell, ke_mass, stream
    ("imexlb_diagnostics_noalloc", _i, [_vp, _vp, _vp, _vp, ctypes.c_int64, _vp, _vp]),
    ("imexlb_sphere_force", _i, [_vp, _dp]),
    ("imexlb_sync", _i, [_vp]),
    ("imexlb_get_info", _i, [_vp, ctypes.POINTER(Info)]),
    ("imexlb_last_error", ctypes.c_char_p, [_vp]),
    ("imexlb_strerror", ctypes.c_char_p, [_i]),
    ("imexlb_copy_bandwidth", _i, [_i, ctypes.c_size_t, _i, _dp, _fp]),
    ("imexlb_bandwidth_probe", _i, [_i, _i, ctypes.c_size_t, _i, _dp, _fp]),
    ("imexlb_system_read", _i, [ctypes.c_char_p, _i, _i, ctypes.POINTER(SystemC)]),
    ("imexlb_comm_layout", _i, [_i, _i, _i, _i, _i, _i, ctypes.POINTER(Comm)]),
    ("imexlb_checkpoint_write", _i, [_vp, ctypes.c_char_p]),
    ("imexlb_checkpoint_read", _i, [_vp, ctypes.c_char_p]),
    ("imexlb_tecplot_write", _i, [ctypes.c_char_p, ctypes.c_double, _i, _i, _i, _dp, _dp, _dp, _dp]),
    ("imexlb_autotune", _i, [_vp, ctypes.c_char_p, _i, ctypes.POINTER(ctypes.c_int32)]),
    ("imexlb_set_launch", _i, [_vp, _i, _i, _i, _i]),
    ("imexlb_tune_placement", _i, [_vp, _i, _i, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_int32)]),
    ("imexlb_tune_placement_window", _i, [_vp, ctypes.c_int64, ctypes.c_int64, _i, _fp, ctypes.c_int32,
                                          ctypes.POINTER(ctypes.c_int32)]),
    ("imexlb_ensemble_generate", _i, [ctypes.c_uint64, _i, ctypes.POINTER(EnsembleDomain)]),
    ("imexlb_ensemble_create", _i, [ctypes.POINTER(EnsembleDomain), _i, _i, _i, _i, ctypes.POINTER(_vp)]),
    ("imexlb_ensemble_destroy", _i, [_vp]),
    ("imexlb_ensemble_initialize", _i, [_vp]),
    ("imexlb_ensemble_step", _i, [_vp, _i, _i, _i]),
    ("imexlb_ensemble_step_timed", _i, [_vp, _i, _i, _i, _fp]),
    ("imexlb_ensemble_get_moments", _i, [_vp, _i, _dp, _dp, _dp, _dp]),
    ("imexlb_ensemble_info", _i, [_vp, ctypes.POINTER(ctypes.c_int64)]),
    ("imexlb_ensemble_status", _i, [_vp, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]),
    ("imexlb_ensemble_set_launch", _i, [_vp, _i, _i, _i]),
    ("imexlb_ensemble_autotune", _i, [_vp, ctypes.c_char_p, _i, ctypes.POINTER(ctypes.c_int32)]),
    ("imexlb_ensemble_tune_placement", _i, [_vp, _i, _i, ctypes.POINTER(ctypes.c_float),
                                            ctypes.POINTER(ctypes.c_int32)]),
    ("imexlb_ensemble_tune_placement_window", _i, [_vp, ctypes.c_int64, ctypes.c_int64, _i, _fp, ctypes.c_int32,
                                                   ctypes.POINTER(ctypes.c_int32)]),
    ("imexlb_ensemble_last_error", ctypes.c_char_p, [_vp]),
]

_lib = None


def lib():
    """The loaded library (raises if it was never built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m imexlb_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class ImexlbError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{msg} (status {status})")
        self.status = status


def dptr(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def as_f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)
