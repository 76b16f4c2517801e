"""imexlb_diagnostics_noalloc (the allocation-free form of imexlb_diagnostics)
through the C-ABI: on caller-owned device arrays it returns the bits the
handle's imexlb_diagnostics returns for the same moments, on two calls in a
row, the second over a smaller domain."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def hip_runtime():
    """The HIP runtime libimexlb.so itself runs on (the test allocates the
    caller-owned device arrays with it)."""
    from imexlb_amd import lib
    lib()
    with open("/proc/self/maps") as f:
        paths = {line.split()[-1] for line in f if "libamdhip64" in line}
    assert len(paths) == 1, paths
    H = ctypes.CDLL(paths.pop())
    vp, sz = ctypes.c_void_p, ctypes.c_size_t
    H.hipMalloc.argtypes = [ctypes.POINTER(vp), sz]
    H.hipFree.argtypes = [vp]
    H.hipMemcpy.argtypes = [vp, vp, sz, ctypes.c_int]
    H.hipDeviceSynchronize.argtypes = []
    for name in ("hipMalloc", "hipFree", "hipMemcpy", "hipDeviceSynchronize"):
        getattr(H, name).restype = ctypes.c_int
    return H


H2D, D2H = 1, 2


class DeviceArray:
    def __init__(self, H, nbytes):
        self.H, self.ptr = H, ctypes.c_void_p()
        assert H.hipMalloc(ctypes.byref(self.ptr), max(nbytes, 8)) == 0

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert self.H.hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, H2D) == 0

    def download(self, n):
        out = np.empty(n)
        assert self.H.hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H) == 0
        return out

    def free(self):
        self.H.hipFree(self.ptr)


def noalloc_diagnostics(H, fields):
    from imexlb_amd import lib
    n = fields[0].size
    arrs = [DeviceArray(H, 8 * n) for _ in range(4)]
    out = DeviceArray(H, 16)
    try:
        for d, a in zip(arrs, fields):
            d.upload(a.ravel())
        rc = lib().imexlb_diagnostics_noalloc(*[d.ptr for d in arrs], n, out.ptr, None)
        assert rc == 0, lib().imexlb_last_error(None)
        assert H.hipDeviceSynchronize() == 0
        return tuple(out.download(2))
    finally:
        for d in arrs + [out]:
            d.free()


def handle_run(shape, case, steps, **kw):
    from imexlb_amd import LBM
    g = LBM(*shape, 0.02, case=case, engine="fused", **kw)
    g.Initialize()
    g.step(steps, moments_every=1)
    want = g.diagnostics()
    fields = g.moments()
    g.close()
    return want, fields


def test_noalloc_diagnostics_equals_handle_diagnostics():
    H = hip_runtime()
    # first call: TGV 40x36x28, AA, 9 steps
    print("[noalloc] call 1: TGV 40x36x28", flush=True)
    want, fields = handle_run((40, 36, 28), "tgv", 9, storage="aa", u0=0.05)
    got = noalloc_diagnostics(H, fields)
    assert got == want, (got, want)
    # second call, smaller: config cavity 12x10x8, pull, 5 steps; the scratch
    # still holds the first call's partial sums
    print("[noalloc] call 2: cavity 12x10x8", flush=True)
    want2, fields2 = handle_run((12, 10, 8), "cavity", 5, mode="config", lid_u=0.08)
    got2 = noalloc_diagnostics(H, fields2)
    assert got2 == want2, (got2, want2)
    assert want2 != want


def test_noalloc_diagnostics_rejects_bad_arguments():
    from imexlb_amd import lib
    from imexlb_amd._lib import EINVAL
    H = hip_runtime()
    out = DeviceArray(H, 16)
    try:
        L = lib()
        big = 2 ** 31 + 1
        assert L.imexlb_diagnostics_noalloc(out.ptr, out.ptr, out.ptr, out.ptr, big, out.ptr, None) == EINVAL
        assert b"at most" in L.imexlb_last_error(None)
        assert L.imexlb_diagnostics_noalloc(None, None, None, None, 5, out.ptr, None) == EINVAL
        assert L.imexlb_diagnostics_noalloc(out.ptr, out.ptr, out.ptr, out.ptr, 1, None, None) == EINVAL
        # zero cells: zero energy and mass, no reads
        assert L.imexlb_diagnostics_noalloc(None, None, None, None, 0, out.ptr, None) == 0
        assert H.hipDeviceSynchronize() == 0
        assert tuple(out.download(2)) == (0.0, 0.0)
    finally:
        out.free()
