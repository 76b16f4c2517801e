// imexlb_diagnostics_noalloc: the FP64 kinetic energy and mass of
// imexlb_diagnostics, over caller-owned device arrays, for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc).  Its only scratch memory -- the per-block
// partial sums of the two-pass reduction -- is the file-scope device array
// below, which every GPU the library runs on holds until the process ends.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>

#include "../../include/imexlb.h"
#include "kernels.h"

namespace {

// The handle's partial count, so both entry points add the same numbers in
// the same order and return the same bits.
constexpr int NOALLOC_PARTS = 1024;
constexpr int MAX_DEVICES = 64;

// 2 doubles (kinetic energy, mass) per block of the first pass: 16 KiB.
// Every block writes its own two slots on every call before the second pass
// reads them, so nothing has to start at zero.
static __device__ __align__(256) double scratch[2 * NOALLOC_PARTS];

std::mutex g_mu;
double *g_scratch[MAX_DEVICES] = {};  // scratch's device address, per device

// The device address of `scratch` on `dev`, looked up once per device with
// that device current (each GPU holds its own copy at its own address).  The
// name `scratch` in host code is a host-side stand-in, never passed to a
// kernel or a copy.
int scratch_of(int dev, double **out)
{
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            imx::set_call_error("imexlb_diagnostics_noalloc: hipGetSymbolAddress(scratch) on device %d: %s", dev,
                                hipGetErrorString(e));
            return IMEXLB_EHIP;
        }
        g_scratch[dev] = (double *)ptr;
    }
    *out = g_scratch[dev];
    return IMEXLB_OK;
}

}  // namespace

extern "C" int imexlb_diagnostics_noalloc(const double *p, const double *u, const double *v, const double *w,
                                          int64_t ncell, double *ke_mass, void *stream)
{
    if (ncell < 0 || ncell > IMEXLB_NOALLOC_MAX_CELLS) {
        imx::set_call_error("imexlb_diagnostics_noalloc: %lld cells; at most %lld", (long long)ncell,
                            (long long)IMEXLB_NOALLOC_MAX_CELLS);
        return IMEXLB_EINVAL;
    }
    if (!ke_mass || (ncell > 0 && (!p || !u || !v || !w))) {
        imx::set_call_error("imexlb_diagnostics_noalloc: null device pointer");
        return IMEXLB_EINVAL;
    }
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess || dev < 0 || dev >= MAX_DEVICES) {
        imx::set_call_error("imexlb_diagnostics_noalloc: current device %d: %s", dev, hipGetErrorString(e));
        return IMEXLB_ENODEV;
    }
    double *partials = nullptr;
    const int rc = scratch_of(dev, &partials);
    if (rc) return rc;
    e = imx::launch_diag(p, u, v, w, nullptr, (uint32_t)ncell, partials, NOALLOC_PARTS, ke_mass,
                         (hipStream_t)stream);
    if (e != hipSuccess) {
        imx::set_call_error("imexlb_diagnostics_noalloc: launch: %s", hipGetErrorString(e));
        return IMEXLB_EHIP;
    }
    return IMEXLB_OK;
}
