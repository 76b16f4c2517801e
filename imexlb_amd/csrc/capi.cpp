// C-ABI of the MI355X D3Q27 BGK step: handle lifetime, the reference's
// per-phase surface, the fused fast path and data access.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/imexlb.h"
#include "window.h"
#include "kernels.h"
#include "phased.h"

using namespace imx;

struct imexlb_handle {
    imexlb_config cfg{};
    hipStream_t stream = nullptr;
    int device = 0;
    bool own_stream = false;
    // fused engine
    void *lat[2] = {nullptr, nullptr};
    // set by imexlb_tune_placement_window: lat[] lie in this address range,
    // of which only the chunks under the lattice(s) are still held
    imx::ChunkedWindow win;
    int cur = 0;             // pull: lattice holding the current state
    double *mom[4] = {nullptr, nullptr, nullptr, nullptr};
    uint8_t *mask = nullptr;
    double *partials = nullptr;
    double *result = nullptr;  // device scratch for reductions (4 doubles)
    int nparts = 1024;
    uint64_t pstride = 0;
    uint32_t ncell = 0;
    size_t elem = 8;
    double shift = 0.0;      // S: FP32/FP16 lattices store f - s_i with sum_i s_i = S
    double ush = 0.0;        // the shift's free-stream x velocity (reduced-precision channel)
    double sh[Q] = {};       // s_i (storage_shift)
    double corr[Q] = {};
    double feq_in[Q] = {};
    double *tgv_tab = nullptr;
    double *ring = nullptr;  // faithful cavity: the first ghost layer (FusedParams::ring)
    double *stage = nullptr; // faithful mode: what the x-face boundary kernels read (FusedParams::stage)
    uint32_t ring_a = 0, ring_b = 0, ring_c = 0;
    int block = 256, cpt = 1, xcd = 1, ntl = 0;
    int64_t steps = 0;
    int64_t mom_step = -1;   // step whose moments are in mom[]
    bool initialized = false;
    bool phase_collided = false;
    int64_t device_bytes = 0;
    // HIP graph of GRAPH_STEPS moment-free steps, replayed by imexlb_step;
    // valid for the lattices, start state and launch shape in gkey
    hipGraphExec_t gexec = nullptr;
    uint64_t gkey[8] = {};
    // phased engine
    PhasedState *ph = nullptr;
    std::string err;
};

static thread_local std::string g_create_err;

static int fail(imexlb_handle *h, int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (h) h->err = buf;
    else g_create_err = buf;
    return code;
}

// The message imexlb_last_error(NULL) returns, for entry points that take no
// handle (imexlb_diagnostics_noalloc).
void imx::set_call_error(const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_create_err = buf;
}

#define HIPCHK(h, expr)                                                                           \
    do {                                                                                          \
        hipError_t e_ = (expr);                                                                   \
        if (e_ != hipSuccess) return fail((h), IMEXLB_EHIP, "%s: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

extern "C" {

int imexlb_config_init(imexlb_config *c)
{
    if (!c) return IMEXLB_EINVAL;
    std::memset(c, 0, sizeof *c);
    c->case_id = IMEXLB_CASE_TGV;
    c->mode = IMEXLB_MODE_FAITHFUL;
    c->engine = IMEXLB_ENGINE_FUSED;
    c->dtype = IMEXLB_F64;
    c->storage = IMEXLB_PULL;
    c->nx = c->ny = c->nz = 64;
    c->tau0 = 0.05 * 64 / 100.0 / (1.0 / 3.0) / 2 / 3.1415926;
    c->rho0 = 1.0;
    c->u0 = 0.05;
    c->lid_u = 0.1;
    c->seed = 5374857ULL;
    c->xcd_remap = -1;
    c->stride_pad = -1;
    c->nt_loads = -1;
    c->addr_mode = -1;
    return IMEXLB_OK;
}

const char *imexlb_strerror(int s)
{
    switch (s) {
    case IMEXLB_OK: return "ok";
    case IMEXLB_EINVAL: return "invalid argument";
    case IMEXLB_EHIP: return "HIP runtime error";
    case IMEXLB_ENOMEM: return "out of memory";
    case IMEXLB_EUNSUPPORTED: return "not supported by this engine";
    case IMEXLB_ESTATE: return "call out of order";
    case IMEXLB_EIO: return "input file error";
    case IMEXLB_ENODEV: return "no HIP device";
    default: return "unknown status";
    }
}

const char *imexlb_last_error(const imexlb_handle *h)
{
    return h ? h->err.c_str() : g_create_err.c_str();
}

}  // extern "C"

// ---------------------------------------------------------------- helpers

static int validate(const imexlb_config *c)
{
    if (c->case_id < 0 || c->case_id > 2) return fail(nullptr, IMEXLB_EINVAL, "case_id %d", c->case_id);
    if (c->mode < 0 || c->mode > 1) return fail(nullptr, IMEXLB_EINVAL, "mode %d", c->mode);
    if (c->engine < 0 || c->engine > 1) return fail(nullptr, IMEXLB_EINVAL, "engine %d", c->engine);
    if (c->dtype < 0 || c->dtype > 2) return fail(nullptr, IMEXLB_EINVAL, "dtype %d", c->dtype);
    if (c->storage < 0 || c->storage > 1) return fail(nullptr, IMEXLB_EINVAL, "storage %d", c->storage);
    if (c->nx < 2 || c->ny < 2 || c->nz < 2)
        return fail(nullptr, IMEXLB_EINVAL, "extents must be >= 2 (got %d %d %d)", c->nx, c->ny, c->nz);
    const uint64_t n = (uint64_t)c->nx * c->ny * c->nz;
    if (n >= (1ull << 31)) return fail(nullptr, IMEXLB_EINVAL, "subdomain of %llu cells exceeds 2^31", (unsigned long long)n);
    if (!(c->tau0 > -0.5) || !std::isfinite(c->tau0))
        return fail(nullptr, IMEXLB_EINVAL, "tau0 = %g gives no stable BGK rate", c->tau0);
    if (!std::isfinite(c->rho0) || !std::isfinite(c->u0)) return fail(nullptr, IMEXLB_EINVAL, "non-finite rho0/u0");
    if (c->block != 0 && (c->block < 64 || c->block > 512 || c->block % 64))
        return fail(nullptr, IMEXLB_EINVAL, "block %d must be a multiple of 64 in [64, 512]", c->block);
    if (c->stride_pad < -1 || c->stride_pad > (1 << 24) || (c->stride_pad > 0 && c->stride_pad % 256))
        return fail(nullptr, IMEXLB_EINVAL, "stride_pad %d must be -1 or a multiple of 256 below 16 MiB", c->stride_pad);
    if (c->cells_per_thread < 0 || c->cells_per_thread > 3)
        return fail(nullptr, IMEXLB_EINVAL, "cells_per_thread %d", c->cells_per_thread);
    if (c->addr_mode < -1 || c->addr_mode > 2) return fail(nullptr, IMEXLB_EINVAL, "addr_mode %d", c->addr_mode);
    if (c->xcd_remap < -1 || c->xcd_remap > 1) return fail(nullptr, IMEXLB_EINVAL, "xcd_remap %d", c->xcd_remap);
    if (c->nt_loads < -1 || c->nt_loads > 1) return fail(nullptr, IMEXLB_EINVAL, "nt_loads %d", c->nt_loads);
    if (c->engine == IMEXLB_ENGINE_FUSED) {
        if (c->mode == IMEXLB_MODE_FAITHFUL && c->case_id != IMEXLB_CASE_TGV &&
            (c->dtype != IMEXLB_F64 || c->storage != IMEXLB_PULL))
            return fail(nullptr, IMEXLB_EUNSUPPORTED,
                        "the faithful cavity and sphere on the fused engine are FP64 with pull storage");
        if (c->case_id == IMEXLB_CASE_SPHERE && c->storage == IMEXLB_AA)
            return fail(nullptr, IMEXLB_EUNSUPPORTED, "the open channel runs with pull storage");
        if (c->case_id == IMEXLB_CASE_SPHERE && c->mode == IMEXLB_MODE_CONFIG) {
            const double r = c->sphere_r;
            if (!(r > 0) || c->sphere_cx - r < 2 || c->sphere_cx + r > c->nx - 3 || c->sphere_cy - r < 1 ||
                c->sphere_cy + r > c->ny - 2 || c->sphere_cz - r < 1 || c->sphere_cz + r > c->nz - 2)
                return fail(nullptr, IMEXLB_EINVAL, "sphere must lie inside the channel with a margin");
        }
    } else {
        if (c->dtype != IMEXLB_F64 || c->storage != IMEXLB_PULL)
            return fail(nullptr, IMEXLB_EUNSUPPORTED, "the phased engine is FP64 with its own layout");
    }
    return IMEXLB_OK;
}

static void fill_fused_params(imexlb_handle *h, FusedParams &P, int src, int dst)
{
    const imexlb_config &c = h->cfg;
    std::memset(&P, 0, sizeof P);
    P.nx = c.nx;
    P.ny = c.ny;
    P.nz = c.nz;
    P.ncell = h->ncell;
    P.pstride = h->pstride;
    P.src = h->lat[src];
    P.dst = h->lat[dst];
    P.mask = h->mask;
    P.tau0 = c.tau0;
    P.omega_f = (float)(1.0 / (c.tau0 + 0.5));
    P.shift = h->shift;
    P.ush = h->ush;
    std::memcpy(P.sh, h->sh, sizeof P.sh);
    std::memcpy(P.corr, h->corr, sizeof P.corr);
    std::memcpy(P.feq_in, h->feq_in, sizeof P.feq_in);
    const uint32_t tile = (uint32_t)h->block * (uint32_t)h->cpt;
    P.nblocks = (h->ncell + tile - 1) / tile;
    P.xcd_remap = h->xcd;
    P.cpt = h->cpt;
    P.ntl = h->ntl;
    P.am_max = c.addr_mode < 0 ? 2 : c.addr_mode;
    P.divx.init(c.nx);
    P.divy.init(c.ny);
    P.ring = h->ring;
    P.stage = h->stage;
    if (h->ring) P.ring_rel = ((intptr_t)h->ring - (intptr_t)h->lat[src]) / (intptr_t)sizeof(double);
    P.ring_a = h->ring_a;
    P.ring_b = h->ring_b;
    P.ring_c = h->ring_c;
    P.ring_n = 2 * (h->ring_a + h->ring_b + h->ring_c);
    P.faithful_sphere = c.case_id == IMEXLB_CASE_SPHERE && c.mode == IMEXLB_MODE_FAITHFUL;
    P.u0 = c.u0;
    P.seed = c.seed;
}

namespace imx {
hipError_t launch_fused_f64(int bc, int pattern, const FusedParams &P, int block, hipStream_t st);
hipError_t launch_fused_f32(int bc, int pattern, const FusedParams &P, int block, hipStream_t st);
hipError_t launch_fused_f16(int bc, int pattern, const FusedParams &P, int block, hipStream_t st);
hipError_t launch_fused(int dtype, int bc, int pattern, const FusedParams &P, int block, hipStream_t st)
{
    if (dtype == DT_F64) return launch_fused_f64(bc, pattern, P, block, st);
    if (dtype == DT_F32) return launch_fused_f32(bc, pattern, P, block, st);
    return launch_fused_f16(bc, pattern, P, block, st);
}
}  // namespace imx

static int bc_of(const imexlb_config &c)
{
    if (c.case_id == IMEXLB_CASE_TGV) return BC_PERIODIC;
    if (c.mode == IMEXLB_MODE_FAITHFUL) return BC_FAITHFUL;
    return c.case_id == IMEXLB_CASE_CAVITY ? BC_CAVITY : BC_CHANNEL;
}

// One fused step: step index k = h->steps + 1.
static int fused_step(imexlb_handle *h, bool write_moments)
{
    const imexlb_config &c = h->cfg;
    FusedParams P;
    int pattern;
    if (c.storage == IMEXLB_PULL) {
        fill_fused_params(h, P, h->cur, 1 - h->cur);
        pattern = PAT_PULL;
    } else {
        fill_fused_params(h, P, 0, 0);
        // state after an even number of steps is the swapped layout, so the
        // next (odd) step streams through neighbours
        pattern = (h->steps % 2 == 0) ? PAT_AA_ODD : PAT_AA_EVEN;
    }
    if (write_moments) {
        P.mom_p = h->mom[0];
        P.mom_u = h->mom[1];
        P.mom_v = h->mom[2];
        P.mom_w = h->mom[3];
    }
    // the faithful cavity's unpack and walls fill the ghost ring from the
    // post-collision lattice first, as the reference's Streaming does
    if (bc_of(c) == BC_FAITHFUL) HIPCHK(h, launch_faithful_boundaries(P, h->stream));
    HIPCHK(h, launch_fused(c.dtype, bc_of(c), pattern, P, h->block, h->stream));
    if (c.storage == IMEXLB_PULL) h->cur = 1 - h->cur;
    h->steps++;
    if (write_moments) h->mom_step = h->steps;
    return IMEXLB_OK;
}

static int set_device(imexlb_handle *h)
{
    HIPCHK(h, hipSetDevice(h->device));
    return IMEXLB_OK;
}

// Steps per captured graph: even, so that the pull lattice index and the AA
// layout parity are where they started when a replay ends.
constexpr int GRAPH_STEPS = 16;

static void graph_key(const imexlb_handle *h, uint64_t (&key)[8])
{
    key[0] = (uint64_t)(uintptr_t)h->lat[0];
    key[1] = (uint64_t)(uintptr_t)h->lat[1];
    key[2] = (uint64_t)h->cur;
    key[3] = (uint64_t)(h->steps & 1);
    key[4] = (uint64_t)h->block;
    key[5] = (uint64_t)h->cpt;
    key[6] = (uint64_t)h->xcd;
    key[7] = (uint64_t)h->ntl;
}

// Capture GRAPH_STEPS moment-free fused steps from the current state (the
// state itself is left as it was) unless a graph for this state is cached.
// Only kernel launches are captured.
static int ensure_graph(imexlb_handle *h)
{
    uint64_t key[8];
    graph_key(h, key);
    if (h->gexec && std::memcmp(key, h->gkey, sizeof key) == 0) return IMEXLB_OK;
    if (h->gexec) {
        hipGraphExecDestroy(h->gexec);
        h->gexec = nullptr;
    }
    const int cur0 = h->cur;
    const int64_t steps0 = h->steps, mom0 = h->mom_step;
    hipGraph_t g = nullptr;
    HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    int rc = IMEXLB_OK;
    for (int s = 0; s < GRAPH_STEPS && !rc; ++s) rc = fused_step(h, false);
    const hipError_t e = hipStreamEndCapture(h->stream, &g);
    h->cur = cur0;
    h->steps = steps0;
    h->mom_step = mom0;
    if (rc || e != hipSuccess || !g) {
        if (g) hipGraphDestroy(g);
        return rc ? rc : fail(h, IMEXLB_EHIP, "step graph capture: %s", hipGetErrorString(e));
    }
    const hipError_t ei = hipGraphInstantiate(&h->gexec, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (ei != hipSuccess) {
        h->gexec = nullptr;
        return fail(h, IMEXLB_EHIP, "step graph instantiate: %s", hipGetErrorString(ei));
    }
    std::memcpy(h->gkey, key, sizeof key);
    return IMEXLB_OK;
}

extern "C" {

int imexlb_create(const imexlb_config *cfg, imexlb_handle **out)
{
    if (!cfg || !out) return fail(nullptr, IMEXLB_EINVAL, "null argument");
    *out = nullptr;
    int rc = validate(cfg);
    if (rc) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(nullptr, IMEXLB_ENODEV, "no HIP device");
    if (cfg->device < 0 || cfg->device >= ndev) return fail(nullptr, IMEXLB_EINVAL, "device %d of %d", cfg->device, ndev);

    imexlb_handle *h = new (std::nothrow) imexlb_handle;
    if (!h) return fail(nullptr, IMEXLB_ENOMEM, "handle");
    h->cfg = *cfg;
    h->device = cfg->device;
    if ((rc = set_device(h))) {
        g_create_err = h->err;
        delete h;
        return rc;
    }
    auto bail = [&](int code) {
        g_create_err = h->err;
        imexlb_destroy(h);
        return code;
    };
    if (hipStreamCreate(&h->stream) != hipSuccess) return bail(fail(h, IMEXLB_EHIP, "hipStreamCreate"));
    h->own_stream = true;
    const imexlb_config &c = h->cfg;

    if (c.engine == IMEXLB_ENGINE_PHASED) {
        rc = phased_create(c, h->stream, &h->ph, h->err);
        if (rc) return bail(rc);
        h->device_bytes = phased_device_bytes(h->ph);
        *out = h;
        return IMEXLB_OK;
    }

    h->ncell = (uint32_t)((uint64_t)c.nx * c.ny * c.nz);
    h->elem = c.dtype == IMEXLB_F64 ? 8 : (c.dtype == IMEXLB_F32 ? 4 : 2);
    {
        // population stride: the cell count rounded up to 256 bytes, plus a
        // pad that staggers the 27 streams across HBM channels
        const uint64_t per = 256 / h->elem;
        const int pad = c.stride_pad < 0 ? IMEXLB_DEFAULT_PAD : c.stride_pad;
        h->pstride = ((uint64_t)h->ncell + per - 1) / per * per + (uint64_t)pad / h->elem;
    }
    h->block = c.block ? c.block : 256;
    h->cpt = c.cells_per_thread ? c.cells_per_thread : 1;
    h->xcd = c.xcd_remap < 0 ? 1 : c.xcd_remap;
    h->ntl = c.nt_loads < 0 ? 1 : (c.nt_loads ? 1 : 0);
    // reduced precision stores g_i = f_i - s_i (storage_shift); FP64 stores f itself
    if (c.dtype != IMEXLB_F64) {
        h->shift = c.rho0;
        // the open channel stores its departure from the inflow equilibrium
        if (c.case_id == IMEXLB_CASE_SPHERE && c.mode == IMEXLB_MODE_CONFIG) h->ush = c.u0;
        storage_shift(h->shift, h->ush, h->sh);
    }
    const size_t lat_bytes = (size_t)Q * h->pstride * h->elem;
    const int nlat = c.storage == IMEXLB_PULL ? 2 : 1;
    for (int k = 0; k < nlat; ++k) {
        if (hipMalloc(&h->lat[k], lat_bytes) != hipSuccess)
            return bail(fail(h, IMEXLB_ENOMEM, "lattice of %zu bytes", lat_bytes));
        h->device_bytes += lat_bytes;
    }
    for (int k = 0; k < 4; ++k) {
        if (hipMalloc(&h->mom[k], (size_t)h->ncell * 8) != hipSuccess)
            return bail(fail(h, IMEXLB_ENOMEM, "moments"));
        h->device_bytes += (int64_t)h->ncell * 8;
        if (hipMemsetAsync(h->mom[k], 0, (size_t)h->ncell * 8, h->stream) != hipSuccess)
            return bail(fail(h, IMEXLB_EHIP, "memset"));
    }
    if (hipMalloc(&h->partials, sizeof(double) * 4 * h->nparts) != hipSuccess ||
        hipMalloc(&h->result, sizeof(double) * 4) != hipSuccess)
        return bail(fail(h, IMEXLB_ENOMEM, "reduction scratch"));
    if (bc_of(c) == BC_FAITHFUL) {
        h->ring_a = (uint32_t)(c.ny + 2) * (uint32_t)(c.nz + 2);
        h->ring_b = (uint32_t)c.nx * (uint32_t)(c.nz + 2);
        h->ring_c = (uint32_t)c.nx * (uint32_t)c.ny;
        const size_t rb = (size_t)Q * 2 * (h->ring_a + h->ring_b + h->ring_c) * 8;
        if (hipMalloc(&h->ring, rb) != hipSuccess) return bail(fail(h, IMEXLB_ENOMEM, "ghost ring"));
        h->device_bytes += (int64_t)rb;
        const size_t sb = (size_t)2 * STAGE_ROW * c.ny * c.nz * 8;
        if (hipMalloc(&h->stage, sb) != hipSuccess) return bail(fail(h, IMEXLB_ENOMEM, "x-face stage"));
        h->device_bytes += (int64_t)sb;
    }

    // boundary constants in the oracle's operation order
    const double cs2 = 1.0 / 3.0;
    if (c.case_id == IMEXLB_CASE_CAVITY) {
        for (int i = 0; i < Q; ++i) h->corr[i] = 6.0 * w(i) * c.rho0 * ((double)EX[i] * c.lid_u);
    }
    if (c.case_id == IMEXLB_CASE_SPHERE && c.mode == IMEXLB_MODE_FAITHFUL) {
        // the reference's ball (sphere/src/lbm.cpp:238): 1 solid, 0 fluid
        std::vector<uint8_t> m((size_t)h->ncell, 0);
        for (int z = 0; z < c.nz; ++z)
            for (int y = 0; y < c.ny; ++y)
                for (int x = 0; x < c.nx; ++x)
                    if (faithful_sphere_solid(c.nx, x, y, z)) m[((size_t)z * c.ny + y) * c.nx + x] = 1;
        if (hipMalloc(&h->mask, h->ncell) != hipSuccess) return bail(fail(h, IMEXLB_ENOMEM, "mask"));
        h->device_bytes += h->ncell;
        if (hipMemcpy(h->mask, m.data(), h->ncell, hipMemcpyHostToDevice) != hipSuccess)
            return bail(fail(h, IMEXLB_EHIP, "mask upload"));
    } else if (c.case_id == IMEXLB_CASE_SPHERE) {
        const double pin = c.rho0 * cs2;
        for (int i = 0; i < Q; ++i) {
            const double u = c.u0, v = 0.0, ww = 0.0;
            const double edu = (double)EX[i] * u + (double)EY[i] * v + (double)EZ[i] * ww;
            const double udu = u * u + v * v + ww * ww;
            const double eu2 = edu * edu;
            const double feq = w(i) * pin * 3.0 + w(i) * (3.0 * edu + 4.5 * eu2 - 1.5 * udu);
            h->feq_in[i] = c.dtype == IMEXLB_F64 ? feq : feq - h->sh[i];
        }
        // mask: 1 solid, 2 fluid with a solid neighbour (y, z periodic)
        std::vector<uint8_t> m((size_t)h->ncell, 0);
        const double r2 = c.sphere_r * c.sphere_r;
        auto solid = [&](int x, int y, int z) {
            const double dx = (double)x - c.sphere_cx, dy = (double)y - c.sphere_cy, dz = (double)z - c.sphere_cz;
            return dx * dx + dy * dy + dz * dz <= r2;
        };
        const int x0 = std::max(0, (int)std::floor(c.sphere_cx - c.sphere_r) - 2);
        const int x1 = std::min(c.nx - 1, (int)std::ceil(c.sphere_cx + c.sphere_r) + 2);
        const int y0 = std::max(0, (int)std::floor(c.sphere_cy - c.sphere_r) - 2);
        const int y1 = std::min(c.ny - 1, (int)std::ceil(c.sphere_cy + c.sphere_r) + 2);
        const int z0 = std::max(0, (int)std::floor(c.sphere_cz - c.sphere_r) - 2);
        const int z1 = std::min(c.nz - 1, (int)std::ceil(c.sphere_cz + c.sphere_r) + 2);
        auto at = [&](int x, int y, int z) -> uint8_t & { return m[((size_t)z * c.ny + y) * c.nx + x]; };
        for (int z = z0; z <= z1; ++z)
            for (int y = y0; y <= y1; ++y)
                for (int x = x0; x <= x1; ++x)
                    if (solid(x, y, z)) at(x, y, z) = 1;
        for (int z = z0; z <= z1; ++z)
            for (int y = y0; y <= y1; ++y)
                for (int x = x0; x <= x1; ++x) {
                    if (at(x, y, z) == 1) continue;
                    for (int i = 1; i < Q; ++i) {
                        const int xx = x + EX[i], yy = y + EY[i], zz = z + EZ[i];
                        if (xx < 0 || xx >= c.nx || yy < 0 || yy >= c.ny || zz < 0 || zz >= c.nz) continue;
                        if (at(xx, yy, zz) == 1) {
                            at(x, y, z) = 2;
                            break;
                        }
                    }
                }
        if (hipMalloc(&h->mask, h->ncell) != hipSuccess) return bail(fail(h, IMEXLB_ENOMEM, "mask"));
        h->device_bytes += h->ncell;
        if (hipMemcpy(h->mask, m.data(), h->ncell, hipMemcpyHostToDevice) != hipSuccess)
            return bail(fail(h, IMEXLB_EHIP, "mask upload"));
    }
    if (c.case_id == IMEXLB_CASE_TGV) {
        // per-axis tables of the TGV initial field (tgv/src/lbm.cpp:197-204)
        std::vector<double> tab(3 * ((size_t)c.nx + c.ny + c.nz));
        size_t o = 0;
        const int ext[3] = {c.nx, c.ny, c.nz};
        for (int a = 0; a < 3; ++a) {
            const int n = ext[a];
            for (int k = 0; k < n; ++k) tab[o + k] = std::sin((double)k / (double)n * 2.0 * PI_REF);
            for (int k = 0; k < n; ++k) tab[o + n + k] = std::cos((double)k / (double)n * 2.0 * PI_REF);
            for (int k = 0; k < n; ++k) tab[o + 2 * n + k] = std::cos((double)k / (double)n * 2.0 * PI_REF * 2.0);
            o += 3 * (size_t)n;
        }
        if (hipMalloc(&h->tgv_tab, tab.size() * 8) != hipSuccess) return bail(fail(h, IMEXLB_ENOMEM, "tables"));
        if (hipMemcpy(h->tgv_tab, tab.data(), tab.size() * 8, hipMemcpyHostToDevice) != hipSuccess)
            return bail(fail(h, IMEXLB_EHIP, "table upload"));
    }
    if (hipStreamSynchronize(h->stream) != hipSuccess) return bail(fail(h, IMEXLB_EHIP, "sync"));
    *out = h;
    return IMEXLB_OK;
}

int imexlb_destroy(imexlb_handle *h)
{
    if (!h) return IMEXLB_OK;
    hipSetDevice(h->device);
    if (h->stream) hipStreamSynchronize(h->stream);
    if (!h->win.empty()) imx::window_destroy(h->win);
    else
        for (auto &p : h->lat)
            if (p) hipFree(p);
    for (auto &p : h->mom)
        if (p) hipFree(p);
    if (h->mask) hipFree(h->mask);
    if (h->partials) hipFree(h->partials);
    if (h->result) hipFree(h->result);
    if (h->tgv_tab) hipFree(h->tgv_tab);
    if (h->ring) hipFree(h->ring);
    if (h->stage) hipFree(h->stage);
    if (h->ph) phased_destroy(h->ph);
    if (h->gexec) hipGraphExecDestroy(h->gexec);
    if (h->stream && h->own_stream) hipStreamDestroy(h->stream);
    delete h;
    return IMEXLB_OK;
}

int imexlb_initialize(imexlb_handle *h)
{
    if (!h) return IMEXLB_EINVAL;
    int rc = set_device(h);
    if (rc) return rc;
    const imexlb_config &c = h->cfg;
    if (h->ph) {
        rc = phased_initialize(h->ph, h->err);
        if (rc) return rc;
        h->initialized = true;
        h->steps = 0;
        return IMEXLB_OK;
    }
    InitParams P;
    std::memset(&P, 0, sizeof P);
    P.nx = c.nx;
    P.ny = c.ny;
    P.nz = c.nz;
    P.ncell = h->ncell;
    P.pstride = h->pstride;
    P.dst = h->lat[0];
    P.layout = c.storage == IMEXLB_AA ? 1 : 0;
    P.kind = c.case_id == IMEXLB_CASE_TGV ? 0 : (bc_of(c) == BC_FAITHFUL && c.case_id == IMEXLB_CASE_SPHERE ? 2 : 1);
    P.seed = c.seed;
    P.tab = h->tgv_tab;
    P.u0 = c.case_id == IMEXLB_CASE_SPHERE ? c.u0 : (c.case_id == IMEXLB_CASE_CAVITY ? 0.0 : c.u0);
    P.rho0 = c.rho0;
    P.cs2 = 1.0 / 3.0;
    // the reference's cavity starts from p = 0, so f = 0 (lid_driven_cavity/src/lbm.cpp:196)
    P.p0 = bc_of(c) == BC_FAITHFUL ? 0.0 : c.rho0 * P.cs2;
    std::memcpy(P.sh, h->sh, sizeof P.sh);
    P.mask = h->mask;
    for (int k = 0; k < 4; ++k) P.mom[k] = h->mom[k];
    P.divx.init(c.nx);
    P.divy.init(c.ny);
    HIPCHK(h, launch_init(c.dtype, P, h->stream));
    if (h->ring) {  // the first ghost layer starts from the initial fields too
        FusedParams F;
        fill_fused_params(h, F, 0, 0);
        HIPCHK(h, launch_faithful_ring_init(F, h->stream));
        if (h->stage) HIPCHK(h, launch_faithful_stage_fill(F, h->stream));
    }
    h->cur = 0;
    if (c.storage == IMEXLB_PULL && c.case_id == IMEXLB_CASE_SPHERE) {
        // solid cells are never written by the step; give both lattices the
        // same initial contents
        HIPCHK(h, hipMemcpyAsync(h->lat[1], h->lat[0], (size_t)Q * h->pstride * h->elem, hipMemcpyDeviceToDevice,
                                 h->stream));
    }
    h->steps = 0;
    h->mom_step = 0;  // k_init wrote the initial fields
    h->initialized = true;
    h->phase_collided = false;
    return IMEXLB_OK;
}

int imexlb_setup_subdomain(imexlb_handle *h) { return h ? IMEXLB_OK : IMEXLB_EINVAL; }

int imexlb_collide(imexlb_handle *h)
{
    if (!h) return IMEXLB_EINVAL;
    if (!h->initialized) return fail(h, IMEXLB_ESTATE, "Collision before Initialize");
    if (h->ph) return phased_collide(h->ph, h->err);
    h->phase_collided = true;  // fused: the collision happens inside Streaming
    return IMEXLB_OK;
}

int imexlb_pack(imexlb_handle *h)
{
    if (!h) return IMEXLB_EINVAL;
    if (h->ph) return phased_pack(h->ph, h->err);
    return IMEXLB_OK;  // fused: periodic images are read through the index wrap
}

int imexlb_exchange(imexlb_handle *h)
{
    if (!h) return IMEXLB_EINVAL;
    if (h->ph) return phased_exchange(h->ph, h->err);
    return IMEXLB_OK;
}

int imexlb_unpack(imexlb_handle *h)
{
    if (!h) return IMEXLB_EINVAL;
    if (h->ph) return phased_unpack(h->ph, h->err);
    return IMEXLB_OK;
}

int imexlb_stream(imexlb_handle *h)
{
    if (!h) return IMEXLB_EINVAL;
    if (!h->initialized) return fail(h, IMEXLB_ESTATE, "Streaming before Initialize");
    int rc = set_device(h);
    if (rc) return rc;
    if (h->ph) return phased_stream(h->ph, h->err);
    if (!h->phase_collided) return fail(h, IMEXLB_ESTATE, "Streaming without a preceding Collision");
    h->phase_collided = false;
    return fused_step(h, true);
}

int imexlb_update(imexlb_handle *h)
{
    if (!h) return IMEXLB_EINVAL;
    if (h->ph) {
        int rc = phased_update(h->ph, h->err);
        if (!rc) h->steps++;
        return rc;
    }
    return IMEXLB_OK;  // fused: moments were written by Streaming
}

int imexlb_step(imexlb_handle *h, int nsteps, int moments_every)
{
    if (!h || nsteps < 0 || moments_every < 0) return IMEXLB_EINVAL;
    if (!h->initialized) return fail(h, IMEXLB_ESTATE, "step before Initialize");
    int rc = set_device(h);
    if (rc) return rc;
    if (h->ph) {
        if ((rc = phased_steps(h->ph, nsteps, h->err))) return rc;
        h->steps += nsteps;
        return IMEXLB_OK;
    }
    // steps that write no moments go GRAPH_STEPS at a time through a replayed
    // graph (one launch instead of GRAPH_STEPS); the others one by one
    auto writes_moments = [&](int64_t idx, int k) {
        return moments_every > 0 && (idx % moments_every == 0 || k == nsteps - 1);
    };
    int k = 0;
    while (k < nsteps) {
        bool plain_run = nsteps - k >= GRAPH_STEPS;
        for (int j = 0; plain_run && j < GRAPH_STEPS; ++j) plain_run = !writes_moments(h->steps + 1 + j, k + j);
        if (plain_run) {
            if ((rc = ensure_graph(h))) return rc;
            HIPCHK(h, hipGraphLaunch(h->gexec, h->stream));
            h->steps += GRAPH_STEPS;  // an even count: cur and the AA parity are unchanged
            k += GRAPH_STEPS;
            continue;
        }
        if ((rc = fused_step(h, writes_moments(h->steps + 1, k)))) return rc;
        ++k;
    }
    return IMEXLB_OK;
}

int imexlb_step_timed(imexlb_handle *h, int nsteps, int moments_every, float *ms)
{
    if (!h || !ms) return IMEXLB_EINVAL;
    int rc = set_device(h);
    if (rc) return rc;
    hipEvent_t e0, e1;
    HIPCHK(h, hipEventCreate(&e0));
    HIPCHK(h, hipEventCreate(&e1));
    HIPCHK(h, hipEventRecord(e0, h->stream));
    rc = imexlb_step(h, nsteps, moments_every);
    if (rc) {
        hipEventDestroy(e0);
        hipEventDestroy(e1);
        return rc;
    }
    HIPCHK(h, hipEventRecord(e1, h->stream));
    HIPCHK(h, hipEventSynchronize(e1));
    HIPCHK(h, hipEventElapsedTime(ms, e0, e1));
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return IMEXLB_OK;
}

int imexlb_sync(imexlb_handle *h)
{
    if (!h) return IMEXLB_EINVAL;
    int rc = set_device(h);
    if (rc) return rc;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return IMEXLB_OK;
}

int imexlb_get_moments(imexlb_handle *h, double *p, double *u, double *v, double *w)
{
    if (!h) return IMEXLB_EINVAL;
    int rc = set_device(h);
    if (rc) return rc;
    if (h->ph) return phased_get_moments(h->ph, p, u, v, w, h->err);
    if (h->mom_step != h->steps)
        return fail(h, IMEXLB_ESTATE, "no moments for step %lld (last written at %lld)", (long long)h->steps,
                    (long long)h->mom_step);
    double *out[4] = {p, u, v, w};
    for (int k = 0; k < 4; ++k)
        if (out[k]) HIPCHK(h, hipMemcpyAsync(out[k], h->mom[k], (size_t)h->ncell * 8, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return IMEXLB_OK;
}

int imexlb_get_macro(imexlb_handle *h, double *rho, double *u, double *v, double *w)
{
    const int rc = imexlb_get_moments(h, rho, u, v, w);
    if (rc || !rho) return rc;
    const size_t n = (size_t)h->cfg.nx * h->cfg.ny * h->cfg.nz;
    for (size_t k = 0; k < n; ++k) rho[k] = 3.0 * rho[k];
    return IMEXLB_OK;
}

static int xfer(imexlb_handle *h, double *host, bool get)
{
    const imexlb_config &c = h->cfg;
    XferParams P;
    std::memset(&P, 0, sizeof P);
    P.nx = c.nx;
    P.ny = c.ny;
    P.nz = c.nz;
    P.ncell = h->ncell;
    P.pstride = h->pstride;
    std::memcpy(P.sh, h->sh, sizeof P.sh);
    P.bc = bc_of(c);
    std::memcpy(P.corr, h->corr, sizeof P.corr);
    if (c.storage == IMEXLB_PULL) {
        P.lat = h->lat[h->cur];
        P.layout = 0;
    } else {
        P.lat = h->lat[0];
        P.layout = (h->steps % 2 == 0) ? 1 : 2;
    }
    P.divx.init(c.nx);
    P.divy.init(c.ny);
    const size_t bytes = (size_t)Q * h->ncell * 8;
    double *d = nullptr;
    HIPCHK(h, hipMalloc(&d, bytes));
    P.buf = d;
    hipError_t e = hipSuccess;
    if (get) {
        e = launch_get_f(c.dtype, P, h->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(host, d, bytes, hipMemcpyDeviceToHost, h->stream);
    } else {
        e = hipMemcpyAsync(d, host, bytes, hipMemcpyHostToDevice, h->stream);
        if (e == hipSuccess) e = launch_set_f(c.dtype, P, h->stream);
        if (e == hipSuccess && h->stage) {  // the x-face kernels' copy of the new populations
            FusedParams F;
            fill_fused_params(h, F, h->cur, 1 - h->cur);
            e = launch_faithful_stage_fill(F, h->stream);
        }
    }
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    hipFree(d);
    if (e != hipSuccess) return fail(h, IMEXLB_EHIP, "population transfer: %s", hipGetErrorString(e));
    return IMEXLB_OK;
}

int imexlb_get_f(imexlb_handle *h, double *f)
{
    if (!h || !f) return IMEXLB_EINVAL;
    if (!h->initialized) return fail(h, IMEXLB_ESTATE, "get_f before Initialize");
    int rc = set_device(h);
    if (rc) return rc;
    if (h->ph) return phased_get_f(h->ph, f, h->err);
    return xfer(h, f, true);
}

int imexlb_set_f(imexlb_handle *h, const double *f)
{
    if (!h || !f) return IMEXLB_EINVAL;
    if (!h->initialized) return fail(h, IMEXLB_ESTATE, "set_f before Initialize");
    int rc = set_device(h);
    if (rc) return rc;
    if (h->ph) return phased_set_f(h->ph, f, h->err);
    h->mom_step = -1;
    return xfer(h, const_cast<double *>(f), false);
}

static const char CK_MAGIC[8] = {'I', 'M', 'X', 'L', 'B', 'C', 'K', '1'};
static const char CK_MAGIC_PHASED[8] = {'I', 'M', 'X', 'L', 'B', 'P', 'H', '1'};
// the fused engine's faithful runs: the populations, then the ghost ring
static const char CK_MAGIC_RING[8] = {'I', 'M', 'X', 'L', 'B', 'C', 'R', '1'};

static size_t ring_doubles(const imexlb_handle *h)
{
    return h->ring ? (size_t)Q * 2 * (h->ring_a + h->ring_b + h->ring_c) : 0;
}

int imexlb_checkpoint_write(imexlb_handle *h, const char *path)
{
    if (!h || !path) return IMEXLB_EINVAL;
    if (!h->initialized) return fail(h, IMEXLB_ESTATE, "checkpoint before Initialize");
    int rc = set_device(h);
    if (rc) return rc;
    // fused: the post-collision populations of the interior (and, in faithful
    // mode, the ghost ring after them); phased: the engine's whole state,
    // ghost cells and moments included
    const size_t nf = h->ph ? (size_t)phased_state_doubles(h->ph) : (size_t)Q * h->ncell;
    const size_t n = nf + ring_doubles(h);
    std::vector<double> f(n);
    rc = h->ph ? phased_get_state(h->ph, f.data(), h->err) : imexlb_get_f(h, f.data());
    if (rc) return rc;
    if (h->ring) {
        HIPCHK(h, hipMemcpyAsync(f.data() + nf, h->ring, ring_doubles(h) * 8, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
    }
    FILE *fp = std::fopen(path, "wb");
    if (!fp) return fail(h, IMEXLB_EIO, "cannot open %s", path);
    const int64_t steps = h->steps;
    const char *magic = h->ph ? CK_MAGIC_PHASED : (h->ring ? CK_MAGIC_RING : CK_MAGIC);
    bool ok = std::fwrite(magic, 1, 8, fp) == 8 &&
              std::fwrite(&h->cfg, sizeof h->cfg, 1, fp) == 1 && std::fwrite(&steps, 8, 1, fp) == 1 &&
              std::fwrite(f.data(), 8, n, fp) == n;
    if (std::fclose(fp) != 0) ok = false;
    return ok ? IMEXLB_OK : fail(h, IMEXLB_EIO, "short write to %s", path);
}

int imexlb_checkpoint_read(imexlb_handle *h, const char *path)
{
    if (!h || !path) return IMEXLB_EINVAL;
    if (!h->initialized) {
        int rc = imexlb_initialize(h);
        if (rc) return rc;
    }
    FILE *fp = std::fopen(path, "rb");
    if (!fp) return fail(h, IMEXLB_EIO, "cannot open %s", path);
    char magic[8];
    imexlb_config c;
    int64_t steps = 0;
    const size_t nf = h->ph ? (size_t)phased_state_doubles(h->ph) : (size_t)Q * h->ncell;
    const size_t n = nf + ring_doubles(h);
    std::vector<double> f(n);
    bool ok = std::fread(magic, 1, 8, fp) == 8 && std::fread(&c, sizeof c, 1, fp) == 1 &&
              std::fread(&steps, 8, 1, fp) == 1;
    const char *want = h->ph ? CK_MAGIC_PHASED : (h->ring ? CK_MAGIC_RING : CK_MAGIC);
    if (ok && std::memcmp(magic, want, 8) != 0) {
        std::fclose(fp);
        const bool known = std::memcmp(magic, CK_MAGIC, 8) == 0 || std::memcmp(magic, CK_MAGIC_PHASED, 8) == 0 ||
                           std::memcmp(magic, CK_MAGIC_RING, 8) == 0;
        if (!known) return fail(h, IMEXLB_EIO, "%s is not a restart file", path);
        return fail(h, IMEXLB_EINVAL, "checkpoint %s holds another engine's or mode's state", path);
    }
    if (ok && (c.case_id != h->cfg.case_id || c.mode != h->cfg.mode || c.nx != h->cfg.nx || c.ny != h->cfg.ny ||
               c.nz != h->cfg.nz)) {
        std::fclose(fp);
        return fail(h, IMEXLB_EINVAL, "checkpoint %s is for another case or extent", path);
    }
    // the physics the populations were advanced with: the BGK rate and the
    // density everywhere; the lid of the config cavity; the inlet speed, the
    // config ball (its mask must match the populations) and the faithful
    // sphere's seed (its deeper ghost cells are recomputed from the initial
    // fields on every step).  The TGV amplitude only sets the initial field.
    if (ok) {
        const char *what = nullptr;
        if (c.tau0 != h->cfg.tau0) what = "tau0";
        else if (c.rho0 != h->cfg.rho0) what = "rho0";
        else if (c.case_id == IMEXLB_CASE_CAVITY && c.mode == IMEXLB_MODE_CONFIG && c.lid_u != h->cfg.lid_u)
            what = "lid speed";
        else if (c.case_id == IMEXLB_CASE_SPHERE && c.u0 != h->cfg.u0) what = "inlet speed u0";
        else if (c.case_id == IMEXLB_CASE_SPHERE && c.mode == IMEXLB_MODE_CONFIG &&
                 (c.sphere_cx != h->cfg.sphere_cx || c.sphere_cy != h->cfg.sphere_cy ||
                  c.sphere_cz != h->cfg.sphere_cz || c.sphere_r != h->cfg.sphere_r))
            what = "sphere centre or radius";
        else if (c.case_id == IMEXLB_CASE_SPHERE && c.mode == IMEXLB_MODE_FAITHFUL && c.seed != h->cfg.seed)
            what = "seed";
        if (what) {
            std::fclose(fp);
            return fail(h, IMEXLB_EINVAL, "checkpoint %s was written with another %s", path, what);
        }
    }
    ok = ok && steps >= 0 && std::fread(f.data(), 8, n, fp) == n;
    std::fclose(fp);
    if (!ok) return fail(h, IMEXLB_EIO, "%s is not a complete checkpoint", path);
    int rc = set_device(h);
    if (rc) return rc;
    h->steps = steps;  // fixes the AA layout parity before the populations go in
    if (h->ph) {
        h->phase_collided = false;
        return phased_set_state(h->ph, f.data(), h->err);
    }
    if (h->ring) {
        HIPCHK(h, hipMemcpyAsync(h->ring, f.data() + nf, ring_doubles(h) * 8, hipMemcpyHostToDevice, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
    }
    return imexlb_set_f(h, f.data());
}

int imexlb_diagnostics(imexlb_handle *h, double *ke, double *mass)
{
    if (!h || !ke || !mass) return IMEXLB_EINVAL;
    int rc = set_device(h);
    if (rc) return rc;
    if (h->ph) return phased_diagnostics(h->ph, ke, mass, h->err);
    if (h->mom_step != h->steps) return fail(h, IMEXLB_ESTATE, "diagnostics need the moments of the current step");
    HIPCHK(h, launch_diag(h->mom[0], h->mom[1], h->mom[2], h->mom[3], h->mask, h->ncell, h->partials, h->nparts,
                          h->result, h->stream));
    double out[2];
    HIPCHK(h, hipMemcpyAsync(out, h->result, sizeof out, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    *ke = out[0];
    *mass = out[1];
    return IMEXLB_OK;
}

int imexlb_sphere_force(imexlb_handle *h, double *force3)
{
    if (!h || !force3) return IMEXLB_EINVAL;
    if (h->ph || h->cfg.case_id != IMEXLB_CASE_SPHERE || h->cfg.mode != IMEXLB_MODE_CONFIG)
        return fail(h, IMEXLB_EUNSUPPORTED, "sphere force is defined for the config-mode channel");
    int rc = set_device(h);
    if (rc) return rc;
    FusedParams P;
    fill_fused_params(h, P, h->cur, h->cur);
    HIPCHK(h, launch_force(h->cfg.dtype, P, h->partials, h->nparts, h->result, h->stream));
    HIPCHK(h, hipMemcpyAsync(force3, h->result, 3 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return IMEXLB_OK;
}

int imexlb_get_info(imexlb_handle *h, imexlb_info *info)
{
    if (!h || !info) return IMEXLB_EINVAL;
    std::memset(info, 0, sizeof *info);
    const imexlb_config &c = h->cfg;
    info->cells = (int64_t)c.nx * c.ny * c.nz;
    info->steps = h->steps;
    const double e = c.dtype == IMEXLB_F64 ? 8.0 : (c.dtype == IMEXLB_F32 ? 4.0 : 2.0);
    info->bytes_per_lup = 2.0 * Q * (h->ph ? 8.0 : e) + (c.case_id == IMEXLB_CASE_SPHERE && !h->ph ? 1.0 : 0.0);
    info->device_bytes = h->device_bytes;
    info->block = h->block;
    info->xcd_remap = h->xcd;
    info->nt_loads = h->ntl;
    info->addr_mode = h->ph ? -1 : fused_addr_mode(h->elem, h->pstride, c.addr_mode < 0 ? 2 : c.addr_mode);
    info->cells_per_thread = h->ph ? h->cpt : fused_cells_per_thread(h->cpt, bc_of(c), (uint32_t)c.nx, info->addr_mode);
    return IMEXLB_OK;
}

int imexlb_autotune(imexlb_handle *h, const char *cache_path, int steps, int32_t *chosen)
{
    if (!h) return IMEXLB_EINVAL;
    if (h->ph) return fail(h, IMEXLB_EUNSUPPORTED, "the phased engine has no launch shapes to tune");
    int rc = set_device(h);
    if (rc) return rc;
    const imexlb_config &c = h->cfg;
    char shape[96];
    snprintf(shape, sizeof shape, "bc%d_%dx%dx%d", bc_of(c), c.nx, c.ny, c.nz);
    const std::string key = tune_key(h->device, "single", c.dtype, c.storage, shape);
    TuneChoice best{};
    if (!tune_cache_read(cache_path, key, best)) {
        if (steps < 2) steps = 2;
        steps += steps & 1;  // whole AA periods
        if (!h->initialized && (rc = imexlb_initialize(h))) return rc;
        auto time_shape = [&](const TuneChoice &t, int n, float &ms) -> int {
            h->block = t.block;
            h->cpt = t.cpt;
            h->xcd = t.grid_or_xcd;
            h->ntl = t.ntl;
            int r = imexlb_step(h, 2, 0);
            if (!r) r = imexlb_step_timed(h, n, 0, &ms);
            return r;
        };
        std::vector<TuneChoice> cand;
        const int blocks[4] = {64, 128, 256, 512};
        // cell pairs (3) only where their kernel runs (periodic, nx even,
        // grouped bases: address mode 2); elsewhere they would time the
        // one-cell kernel again
        const int am = fused_addr_mode(h->elem, h->pstride, c.addr_mode < 0 ? 2 : c.addr_mode);
        const int max_cpt = fused_cells_per_thread(3, bc_of(c), (uint32_t)c.nx, am) == 3 ? 3 : 2;
        for (int b : blocks)
            for (int cpt = 1; cpt <= max_cpt; ++cpt)
                for (int xcd = 0; xcd <= 1; ++xcd)
                    for (int ntl = 0; ntl <= 1; ++ntl) {
                        TuneChoice t{b, cpt, xcd, ntl, 0.0f};
                        if ((rc = time_shape(t, steps, t.ms))) return rc;
                        cand.push_back(t);
                    }
        // the shapes differ by a few per cent at most: time the four fastest
        // again, interleaved and twice as long, and keep the fastest sum
        std::sort(cand.begin(), cand.end(), [](const TuneChoice &x, const TuneChoice &y) { return x.ms < y.ms; });
        cand.resize(std::min<size_t>(cand.size(), 4));
        std::vector<float> sum(cand.size(), 0.0f);
        for (int rep = 0; rep < 2; ++rep)
            for (size_t k = 0; k < cand.size(); ++k) {
                float ms = 0;
                if ((rc = time_shape(cand[k], 2 * steps, ms))) return rc;
                sum[k] += ms;
            }
        size_t kb = 0;
        for (size_t k = 1; k < cand.size(); ++k)
            if (sum[k] < sum[kb]) kb = k;
        best = cand[kb];
        best.ms = sum[kb] / 4.0f;  // time of `steps` steps, as the first pass records it
        tune_cache_write(cache_path, key, best);
    }
    h->block = best.block;
    h->cpt = best.cpt;
    h->xcd = best.grid_or_xcd;
    h->ntl = best.ntl;
    if (chosen) {
        chosen[0] = best.block;
        chosen[1] = best.cpt;
        chosen[2] = best.grid_or_xcd;
        chosen[3] = best.ntl;
    }
    return imexlb_initialize(h);
}

int imexlb_set_launch(imexlb_handle *h, int block, int cpt, int xcd, int ntl)
{
    if (!h) return IMEXLB_EINVAL;
    if (h->ph) return fail(h, IMEXLB_EUNSUPPORTED, "the phased engine has no launch shapes");
    if (block < 64 || block > 512 || block % 64 || (cpt < 1 || cpt > 3) || (xcd != 0 && xcd != 1) ||
        (ntl != 0 && ntl != 1))
        return fail(h, IMEXLB_EINVAL, "launch shape %d/%d/%d/%d", block, cpt, xcd, ntl);
    h->block = block;
    h->cpt = cpt;
    h->xcd = xcd;
    h->ntl = ntl;
    return IMEXLB_OK;
}

int imexlb_tune_placement(imexlb_handle *h, int trials, int steps, float *ms_per_step, int32_t *chosen)
{
    if (!h) return IMEXLB_EINVAL;
    if (trials < 1 || trials > 16) return fail(h, IMEXLB_EINVAL, "placement trials %d must be in [1, 16]", trials);
    if (h->ph) return fail(h, IMEXLB_EUNSUPPORTED, "the phased engine keeps its own layout");
    int rc = set_device(h);
    if (rc) return rc;
    if (steps < 2) steps = 2;
    steps += steps & 1;  // whole AA periods
    const size_t lat_bytes = (size_t)Q * h->pstride * h->elem;
    const int nlat = h->cfg.storage == IMEXLB_PULL ? 2 : 1;
    // candidate 0 is the current lattice, which may lie in a window
    imx::ChunkedWindow win0 = std::move(h->win);
    h->win = imx::ChunkedWindow{};
    std::vector<std::array<void *, 2>> cand{{h->lat[0], h->lat[1]}};
    std::vector<float> ms(trials, 0.0f);
    auto time_current = [&](float &out) -> int {
        int r = imexlb_initialize(h);
        if (!r) r = imexlb_step(h, 2, 0);
        if (!r) r = imexlb_step_timed(h, steps, 0, &out);
        if (!r) out /= (float)steps;
        return r;
    };
    rc = time_current(ms[0]);
    for (int t = 1; !rc && t < trials; ++t) {
        size_t free_b = 0, total_b = 0;
        if (hipMemGetInfo(&free_b, &total_b) != hipSuccess || free_b < nlat * lat_bytes + (2ull << 30)) break;
        std::array<void *, 2> p{nullptr, nullptr};
        bool ok = true;
        for (int k = 0; k < nlat && ok; ++k) ok = hipMalloc(&p[k], lat_bytes) == hipSuccess;
        if (!ok) {
            (void)hipGetLastError();
            for (void *q : p)
                if (q) hipFree(q);
            break;
        }
        cand.push_back(p);
        h->lat[0] = p[0];
        h->lat[1] = p[1];
        rc = time_current(ms[t]);
    }
    int best = 0;
    if (!rc)
        for (int t = 1; t < (int)cand.size(); ++t)
            if (ms[t] < ms[best]) best = t;
    (void)hipStreamSynchronize(h->stream);
    for (int t = 0; t < (int)cand.size(); ++t) {
        if (t == best) continue;
        if (t == 0 && !win0.empty()) {
            h->device_bytes -= (int64_t)win0.held() - (int64_t)(nlat * lat_bytes);
            imx::window_destroy(win0);
            continue;
        }
        for (void *q : cand[t])
            if (q) hipFree(q);
    }
    h->lat[0] = cand[best][0];
    h->lat[1] = cand[best][1];
    if (best == 0) h->win = std::move(win0);
    const int rc2 = imexlb_initialize(h);
    if (rc) return rc;
    if (ms_per_step)
        for (int t = 0; t < trials; ++t) ms_per_step[t] = ms[t];
    if (chosen) *chosen = best;
    return rc2;
}

int imexlb_tune_placement_window(imexlb_handle *h, int64_t slack_bytes, int64_t step_bytes, int steps,
                                 float *ms_per_pos, int32_t npos_max, int32_t *chosen)
{
    if (!h) return IMEXLB_EINVAL;
    if (h->ph) return fail(h, IMEXLB_EUNSUPPORTED, "the phased engine keeps its own layout");
    const int64_t two_mib = 2ll << 20;
    if (step_bytes < two_mib || step_bytes % two_mib || slack_bytes < 0 || slack_bytes % step_bytes)
        return fail(h, IMEXLB_EINVAL, "window step must be a multiple of 2 MiB dividing the slack");
    const int64_t npos = slack_bytes / step_bytes + 1;
    if (npos > 256 || npos > npos_max)
        return fail(h, IMEXLB_EINVAL, "%lld window positions (at most %d)", (long long)npos,
                    std::min(256, (int)npos_max));
    int rc = set_device(h);
    if (rc) return rc;
    if (steps < 2) steps = 2;
    steps += steps & 1;  // whole AA periods
    const size_t lat_bytes = (size_t)Q * h->pstride * h->elem;
    const int nlat = h->cfg.storage == IMEXLB_PULL ? 2 : 1;
    const size_t total = (size_t)nlat * lat_bytes + (size_t)slack_bytes;
    size_t free_b = 0, total_b = 0;
    HIPCHK(h, hipMemGetInfo(&free_b, &total_b));
    if (free_b < total + (2ull << 30))
        return fail(h, IMEXLB_ENOMEM, "window of %zu bytes does not fit (%zu free)", total, free_b);
    // one address range in chunks of physical memory: the chunks the kept
    // offset does not use are given back after the search
    imx::ChunkedWindow w;
    if (imx::window_create(w, h->device, total, (size_t)step_bytes) != hipSuccess) {
        (void)hipGetLastError();
        return fail(h, IMEXLB_ENOMEM, "window of %zu bytes", total);
    }
    char *const big = w.base;
    void *old_lat[2] = {h->lat[0], h->lat[1]};
    std::vector<float> ms((size_t)npos, 0.0f);
    for (int64_t k = 0; k < npos && !rc; ++k) {
        h->lat[0] = (char *)big + k * step_bytes;
        h->lat[1] = nlat == 2 ? (char *)h->lat[0] + lat_bytes : nullptr;
        rc = imexlb_initialize(h);
        if (!rc) rc = imexlb_step(h, 2, 0);
        if (!rc) rc = imexlb_step_timed(h, steps, 0, &ms[k]);
        if (!rc) ms[k] /= (float)steps;
    }
    (void)hipStreamSynchronize(h->stream);
    if (rc) {  // keep the lattice the handle had
        h->lat[0] = old_lat[0];
        h->lat[1] = old_lat[1];
        imx::window_destroy(w);
        const int rc2 = imexlb_initialize(h);
        return rc ? rc : rc2;
    }
    int64_t best = 0;
    for (int64_t k = 1; k < npos; ++k)
        if (ms[k] < ms[best]) best = k;
    h->lat[0] = (char *)big + best * step_bytes;
    h->lat[1] = nlat == 2 ? (char *)h->lat[0] + lat_bytes : nullptr;
    if (!h->win.empty()) {
        h->device_bytes -= (int64_t)h->win.held();
        imx::window_destroy(h->win);
    } else {
        h->device_bytes -= (int64_t)(nlat * lat_bytes);
        for (void *q : old_lat)
            if (q) hipFree(q);
    }
    // the lattice stays on the pages it was timed on; the rest goes back
    imx::window_keep(w, (size_t)(best * step_bytes), (size_t)nlat * lat_bytes);
    h->device_bytes += (int64_t)w.held();
    h->win = std::move(w);
    if (ms_per_pos)
        for (int64_t k = 0; k < npos; ++k) ms_per_pos[k] = ms[k];
    if (chosen) *chosen = (int32_t)best;
    return imexlb_initialize(h);
}

// Times `reps` launches of bandwidth probe `kind` (-1: the plain copy of
// launch_copy) over `bytes` bytes per stream.
static int time_stream_kernel(int device, int kind, size_t bytes, int reps, double *gbps, float *ms_per_launch)
{
    if (bytes < 4096 || reps < 1 || !gbps || !ms_per_launch) return IMEXLB_EINVAL;
    const int streams = kind < 0 ? 2 : bandwidth_probe_streams(kind);
    if (streams == 0) return IMEXLB_EINVAL;
    if (hipSetDevice(device) != hipSuccess) return IMEXLB_ENODEV;
    int cus = 256;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus < 1) cus = 256;
    const size_t n = bytes / 8 / 2 * 2;
    double *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, n * 8) != hipSuccess) return IMEXLB_ENOMEM;
    if (hipMalloc(&b, n * 8) != hipSuccess) {
        hipFree(a);
        return IMEXLB_ENOMEM;
    }
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = IMEXLB_OK;
    if (hipStreamCreate(&st) != hipSuccess || hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        rc = IMEXLB_EHIP;
    if (!rc && (hipMemsetAsync(a, 0, n * 8, st) != hipSuccess || hipMemsetAsync(b, 0, n * 8, st) != hipSuccess))
        rc = IMEXLB_EHIP;
    auto launch = [&]() { return kind < 0 ? launch_copy(a, b, n, st) : launch_bandwidth_probe(kind, a, b, n, cus, st); };
    if (!rc && launch() != hipSuccess) rc = IMEXLB_EHIP;  // warm-up
    if (!rc) {
        hipEventRecord(e0, st);
        for (int r = 0; r < reps && !rc; ++r)
            if (launch() != hipSuccess) rc = IMEXLB_EHIP;
        hipEventRecord(e1, st);
        if (hipEventSynchronize(e1) != hipSuccess) rc = IMEXLB_EHIP;
        float ms = 0;
        hipEventElapsedTime(&ms, e0, e1);
        *ms_per_launch = ms / reps;
        *gbps = (double)streams * (double)n * 8.0 / (*ms_per_launch * 1e-3) / 1e9;
    }
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    if (st) {
        hipStreamSynchronize(st);
        hipStreamDestroy(st);
    }
    hipFree(a);
    hipFree(b);
    return rc;
}

int imexlb_copy_bandwidth(int device, size_t bytes, int reps, double *gbps, float *ms_per_launch)
{
    return time_stream_kernel(device, -1, bytes, reps, gbps, ms_per_launch);
}

int imexlb_bandwidth_probe(int device, int kind, size_t bytes, int reps, double *gbps, float *ms_per_launch)
{
    if (kind < 0) return IMEXLB_EINVAL;
    return time_stream_kernel(device, kind, bytes, reps, gbps, ms_per_launch);
}

}  // extern "C"
