// Initialisation, population transfer, FP64 reductions and the copy kernel
// used to calibrate bandwidth.  Compiled with -ffp-contract=off: the initial
// state is the oracle's, bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "fused.h"
#include "kernels.h"

namespace imx {

// feq in the reference's evaluation order (tgv/src/lbm.cpp:209-215).
__device__ __forceinline__ double feq_f64(int i, double p, double u, double v, double w)
{
    const double edu = edu_of<double>(i, u, v, w);
    const double udu = u * u + v * v + w * w;
    const double eu2 = edu * edu;
    return WVAL[WCLASS[i]] * p * 3.0 + WVAL[WCLASS[i]] * (3.0 * edu + 4.5 * eu2 - 1.5 * udu);
}

// Population f of direction i (weight class of slot `slot`, which is i or
// bb(i)) in the stored convention, g = f - s with s the direction's shift.
template <typename S>
__device__ __forceinline__ void put(S *lat, uint64_t ps, int slot, uint32_t c, double f, double s)
{
    S *p = lat + (uint64_t)slot * ps + c;
    if constexpr (std::is_same<S, double>::value)
        *p = f;
    else if constexpr (std::is_same<S, float>::value)
        *p = (float)(f - s);
    else
        *p = (_Float16)((float)(f - s) * H_SCALE[WCLASS[slot]]);
}

template <typename S>
__device__ __forceinline__ double get(const S *lat, uint64_t ps, int slot, uint32_t c, double s)
{
    const S *p = lat + (uint64_t)slot * ps + c;
    if constexpr (std::is_same<S, double>::value)
        return *p;
    else if constexpr (std::is_same<S, float>::value)
        return (double)*p + s;
    else
        return (double)((float)*p * H_INV_SCALE[WCLASS[slot]]) + s;
}

// Initial state = f0 = feq(p, u) of the analytic fields.  The reference's
// first Collision relaxes f0 towards the feq of the same fields, which is f0
// itself, so f0 is also the first post-collision state.
template <typename S>
__global__ void k_init(const InitParams P)
{
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= P.ncell) return;
    const uint32_t t = P.divx.div(c);
    const uint32_t x = c - t * P.nx;
    const uint32_t z = P.divy.div(t);
    const uint32_t y = t - z * P.ny;
    double u, v, w, p;
    if (P.kind == 0) {
        // tgv/src/lbm.cpp:197-204 with per-axis sin/cos tables
        const double *T = P.tab;
        const double sx = T[x], cx = T[P.nx + x], c2x = T[2 * P.nx + x];
        const double *Ty = T + 3 * P.nx;
        const double sy = Ty[y], cy = Ty[P.ny + y], c2y = Ty[2 * P.ny + y];
        const double *Tz = Ty + 3 * P.ny;
        const double cz = Tz[P.nz + z], c2z = Tz[2 * P.nz + z];
        u = P.u0 * sx * cy * cz;
        v = -P.u0 * cx * sy * cz;
        w = 0;
        p = P.rho0 * P.cs2 + P.rho0 * P.u0 * P.u0 / 16.0 * (c2x + c2y) * (c2z + 2.0);
    } else if (P.kind == 2) {
        // sphere/src/lbm.cpp:224-239: p = 0, a parabolic x velocity in z
        // masked by usr and perturbed by urand(100) (here the seeded hash of
        // the cell's index in the reference's ghost-3 layout)
        const uint64_t ly = (uint64_t)P.ny + 6, lz = (uint64_t)P.nz + 6;
        const uint64_t idx = ((uint64_t)(x + 3) * ly + (y + 3)) * lz + (z + 3);
        const int ran = (int)(mix64(P.seed + idx * 0x9E3779B97F4A7C15ULL) % 100ULL);
        const int usr = faithful_sphere_solid(P.nx, (int)x, (int)y, (int)z) ? 0 : 1;
        const int glz = P.nz, kz = (int)z;
        const double gm = (double)(glz - 1);
        u = (P.u0 * 4.0 * (double)kz * (double)(glz - 1 - kz) / (gm * gm)) * (double)usr * (1.0 + (double)ran * 0.001);
        v = 0.0;
        w = 0.0;
        p = 0.0;
    } else {
        const bool solid = P.mask && P.mask[c] == 1;
        u = solid ? 0.0 : P.u0;
        v = 0.0;
        w = 0.0;
        p = P.p0;
    }
    S *lat = (S *)P.dst;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        const double f = feq_f64(i, p, u, v, w);
        put<S>(lat, P.pstride, P.layout == 1 ? BB[i] : i, c, f, P.sh[i]);
    }
    // the fields themselves are the moments of step 0, as the reference's
    // Initialize leaves them in p, ua, va, wa; a solid cell reports a fluid at
    // rest with p = 0, as every later step leaves it
    if (P.mom[0]) {
        const bool solid = P.mask && P.mask[c] == 1;
        P.mom[0][c] = solid ? 0.0 : p;
        P.mom[1][c] = u;
        P.mom[2][c] = v;
        P.mom[3][c] = w;
    }
}

hipError_t launch_init(int dtype, const InitParams &P, hipStream_t st)
{
    dim3 blk(256), grid((P.ncell + 255) / 256);
    if (dtype == DT_F64) hipLaunchKernelGGL(k_init<double>, grid, blk, 0, st, P);
    else if (dtype == DT_F32) hipLaunchKernelGGL(k_init<float>, grid, blk, 0, st, P);
    else hipLaunchKernelGGL(k_init<_Float16>, grid, blk, 0, st, P);
    return hipGetLastError();
}

// Lattice slot holding f*_i(x) for each storage layout.
__device__ __forceinline__ void slot_of(const XferParams &P, int i, uint32_t c, int &slot, uint32_t &cell)
{
    const uint32_t nx = P.nx, ny = P.ny, nz = P.nz;
    if (P.layout == 0) {
        slot = i;
        cell = c;
        return;
    }
    if (P.layout == 1) {
        slot = BB[i];
        cell = c;
        return;
    }
    // streamed (AA after an odd step): f*_i(x) sits at x + e_i in slot i, or
    // at x in slot bb(i) when x + e_i is a cavity wall
    const uint32_t t = P.divx.div(c);
    int x = (int)(c - t * nx);
    const uint32_t zz = P.divy.div(t);
    int y = (int)(t - zz * ny), z = (int)zz;
    int xd = x + EX[i], yd = y + EY[i], zd = z + EZ[i];
    if (P.bc == BC_CAVITY) {
        if (xd < 0 || xd >= (int)nx || yd < 0 || yd >= (int)ny || zd < 0 || zd >= (int)nz) {
            slot = BB[i];
            cell = c;
            return;
        }
    } else {
        xd = xd < 0 ? xd + (int)nx : (xd >= (int)nx ? xd - (int)nx : xd);
        yd = yd < 0 ? yd + (int)ny : (yd >= (int)ny ? yd - (int)ny : yd);
        zd = zd < 0 ? zd + (int)nz : (zd >= (int)nz ? zd - (int)nz : zd);
    }
    slot = i;
    cell = ((uint32_t)zd * ny + (uint32_t)yd) * nx + (uint32_t)xd;
}

template <typename S>
__global__ void k_get_f(const XferParams P)
{
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= P.ncell) return;
    const S *lat = (const S *)P.lat;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        int slot;
        uint32_t cell;
        slot_of(P, i, c, slot, cell);
        double v = 0.0;
        // slot is i or bb(i); both share the weight class of i
        if (slot == i) v = get<S>(lat, P.pstride, i, cell, P.sh[i]);
        else v = get<S>(lat, P.pstride, BB[i], cell, P.sh[i]);
        P.buf[(uint64_t)i * P.ncell + c] = v;
    }
}

template <typename S>
__global__ void k_set_f(const XferParams P)
{
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= P.ncell) return;
    S *lat = (S *)P.lat;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        int slot;
        uint32_t cell;
        slot_of(P, i, c, slot, cell);
        const double v = P.buf[(uint64_t)i * P.ncell + c];
        if (slot == i) put<S>(lat, P.pstride, i, cell, v, P.sh[i]);
        else put<S>(lat, P.pstride, BB[i], cell, v, P.sh[i]);
    }
}

hipError_t launch_get_f(int dtype, const XferParams &P, hipStream_t st)
{
    dim3 blk(256), grid((P.ncell + 255) / 256);
    if (dtype == DT_F64) hipLaunchKernelGGL(k_get_f<double>, grid, blk, 0, st, P);
    else if (dtype == DT_F32) hipLaunchKernelGGL(k_get_f<float>, grid, blk, 0, st, P);
    else hipLaunchKernelGGL(k_get_f<_Float16>, grid, blk, 0, st, P);
    return hipGetLastError();
}

hipError_t launch_set_f(int dtype, const XferParams &P, hipStream_t st)
{
    dim3 blk(256), grid((P.ncell + 255) / 256);
    if (dtype == DT_F64) hipLaunchKernelGGL(k_set_f<double>, grid, blk, 0, st, P);
    else if (dtype == DT_F32) hipLaunchKernelGGL(k_set_f<float>, grid, blk, 0, st, P);
    else hipLaunchKernelGGL(k_set_f<_Float16>, grid, blk, 0, st, P);
    return hipGetLastError();
}

// ------------------------------------------------------------ reductions
// Two passes with a fixed assignment of cells to blocks and a fixed tree in
// each block, so the sums are reproducible run to run.

template <int N>
__device__ __forceinline__ void block_sum(double (&v)[N], double (*sh)[256])
{
    const int t = threadIdx.x;
#pragma unroll
    for (int k = 0; k < N; ++k) sh[k][t] = v[k];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s) {
#pragma unroll
            for (int k = 0; k < N; ++k) sh[k][t] += sh[k][t + s];
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] = sh[k][0];
}

__global__ void __launch_bounds__(256) k_diag_pass1(const double *p, const double *u, const double *v,
                                                    const double *w, uint32_t n, double *partials)
{
    __shared__ double sh[2][256];
    const uint32_t per = (n + gridDim.x - 1) / gridDim.x;
    const uint32_t lo = blockIdx.x * per, hi = min(n, lo + per);
    double acc[2] = {0.0, 0.0};
    for (uint32_t c = lo + threadIdx.x; c < hi; c += blockDim.x) {
        acc[0] += u[c] * u[c] + v[c] * v[c] + w[c] * w[c];
        acc[1] += 3.0 * p[c];
    }
    block_sum<2>(acc, sh);
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = acc[0];
        partials[2 * blockIdx.x + 1] = acc[1];
    }
}

template <int N>
__global__ void __launch_bounds__(256) k_sum_partials(const double *partials, int nparts, double *out, double scale0)
{
    __shared__ double sh[N][256];
    double acc[N];
#pragma unroll
    for (int k = 0; k < N; ++k) acc[k] = 0.0;
    for (int b = threadIdx.x; b < nparts; b += blockDim.x)
#pragma unroll
        for (int k = 0; k < N; ++k) acc[k] += partials[N * b + k];
    block_sum<N>(acc, sh);
    if (threadIdx.x == 0) {
        out[0] = acc[0] * scale0;
#pragma unroll
        for (int k = 1; k < N; ++k) out[k] = acc[k];
    }
}

hipError_t launch_diag(const double *p, const double *u, const double *v, const double *w, const uint8_t *,
                       uint32_t n, double *partials, int nparts, double *out2, hipStream_t st)
{
    hipLaunchKernelGGL(k_diag_pass1, dim3(nparts), dim3(256), 0, st, p, u, v, w, n, partials);
    hipLaunchKernelGGL(k_sum_partials<2>, dim3(1), dim3(256), 0, st, partials, nparts, out2, 0.5);
    return hipGetLastError();
}

// Momentum handed to the solid by halfway bounce-back, from the
// post-collision populations of the channel (pull storage, natural layout):
// F = sum over fluid x, directions i with x + e_i solid, of 2 f*_i(x) e_i.
template <typename S>
__global__ void __launch_bounds__(256) k_force_pass1(const FusedParams P, double *partials)
{
    __shared__ double sh[3][256];
    const uint32_t n = P.ncell;
    const uint32_t per = (n + gridDim.x - 1) / gridDim.x;
    const uint32_t lo = blockIdx.x * per, hi = min(n, lo + per);
    double acc[3] = {0.0, 0.0, 0.0};
    const S *lat = (const S *)P.src;
    for (uint32_t c = lo + threadIdx.x; c < hi; c += blockDim.x) {
        if (P.mask[c] != 2) continue;
        const uint32_t t = P.divx.div(c);
        const int x = (int)(c - t * P.nx);
        const uint32_t zz = P.divy.div(t);
        const int y = (int)(t - zz * P.ny), z = (int)zz;
#pragma unroll
        for (int i = 1; i < Q; ++i) {
            const int xd = x + EX[i];
            if (xd < 0 || xd >= P.nx) continue;
            int yd = y + EY[i], zd = z + EZ[i];
            yd = yd < 0 ? yd + P.ny : (yd >= P.ny ? yd - P.ny : yd);
            zd = zd < 0 ? zd + P.nz : (zd >= P.nz ? zd - P.nz : zd);
            const uint32_t cd = ((uint32_t)zd * P.ny + (uint32_t)yd) * P.nx + (uint32_t)xd;
            if (P.mask[cd] != 1) continue;
            const double f = 2.0 * get<S>(lat, P.pstride, i, c, P.sh[i]);
            acc[0] += f * (double)EX[i];
            acc[1] += f * (double)EY[i];
            acc[2] += f * (double)EZ[i];
        }
    }
    block_sum<3>(acc, sh);
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; ++k) partials[3 * blockIdx.x + k] = acc[k];
}

hipError_t launch_force(int dtype, const FusedParams &P, double *partials, int nparts, double *out3, hipStream_t st)
{
    if (dtype == DT_F64) hipLaunchKernelGGL(k_force_pass1<double>, dim3(nparts), dim3(256), 0, st, P, partials);
    else if (dtype == DT_F32) hipLaunchKernelGGL(k_force_pass1<float>, dim3(nparts), dim3(256), 0, st, P, partials);
    else hipLaunchKernelGGL(k_force_pass1<_Float16>, dim3(nparts), dim3(256), 0, st, P, partials);
    hipLaunchKernelGGL(k_sum_partials<3>, dim3(1), dim3(256), 0, st, partials, nparts, out3, 1.0);
    return hipGetLastError();
}

// Domain health: *flag stays 1 while every value of the four moment fields
// is finite and is set to 0 by any lane that finds one that is not (all
// writers store the same 0).  A moment sums all 27 populations of its cell,
// so a non-finite population shows in the moments of the same step.
__global__ void __launch_bounds__(256) k_nonfinite(const double *p, const double *u, const double *v,
                                                   const double *w, uint32_t n, int32_t *flag)
{
    bool bad = false;
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < n; c += gridDim.x * blockDim.x)
        bad |= !(isfinite(p[c]) && isfinite(u[c]) && isfinite(v[c]) && isfinite(w[c]));
    if (bad) *flag = 0;
}

__global__ void k_fill_i32(int32_t *a, int n, int32_t v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

hipError_t launch_finite_flags(const double *const *mom, const uint32_t *ncell, int ndom, int32_t *flags,
                               hipStream_t st)
{
    hipLaunchKernelGGL(k_fill_i32, dim3((ndom + 255) / 256), dim3(256), 0, st, flags, ndom, 1);
    for (int d = 0; d < ndom; ++d) {
        const uint32_t n = ncell[d];
        const int grid = (int)std::min<uint32_t>(1024u, (n + 255) / 256);
        hipLaunchKernelGGL(k_nonfinite, dim3(grid), dim3(256), 0, st, mom[4 * d], mom[4 * d + 1], mom[4 * d + 2],
                           mom[4 * d + 3], n, flags + d);
    }
    return hipGetLastError();
}

// Plain copy, one 16-byte vector per lane, one pass: the reference point for
// the hardware counters and the measured bandwidth ceiling.
typedef double dvec2 __attribute__((ext_vector_type(2)));

__global__ void __launch_bounds__(256) k_copy(const dvec2 *__restrict__ src, dvec2 *__restrict__ dst, size_t n2)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n2) dst[i] = src[i];
}

hipError_t launch_copy(const double *src, double *dst, size_t n, hipStream_t st)
{
    const size_t n2 = n / 2;
    hipLaunchKernelGGL(k_copy, dim3((unsigned)((n2 + 255) / 256)), dim3(256), 0, st, (const dvec2 *)src,
                       (dvec2 *)dst, n2);
    return hipGetLastError();
}

// Bandwidth probes: copy, read-only and write-only streams with V 16-byte
// vectors per lane (a workgroup covers V * 256 consecutive vectors, each
// sub-pass coalesced), optionally non-temporal, or a grid-stride loop over a
// fixed grid.  They bound what any kernel with one read and one write stream
// per value can reach on the part.
template <int V, bool NT>
__global__ void __launch_bounds__(256) k_copy_v(const dvec2 *__restrict__ src, dvec2 *__restrict__ dst, size_t n2)
{
    const size_t base = (size_t)blockIdx.x * 256 * V + threadIdx.x;
    dvec2 x[V];
#pragma unroll
    for (int r = 0; r < V; ++r) {
        const size_t i = base + (size_t)r * 256;
        if (i < n2) x[r] = NT ? __builtin_nontemporal_load(src + i) : src[i];
    }
#pragma unroll
    for (int r = 0; r < V; ++r) {
        const size_t i = base + (size_t)r * 256;
        if (i < n2) {
            if (NT) __builtin_nontemporal_store(x[r], dst + i);
            else dst[i] = x[r];
        }
    }
}

template <bool NT>
__global__ void __launch_bounds__(256) k_copy_gs(const dvec2 *__restrict__ src, dvec2 *__restrict__ dst, size_t n2)
{
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n2; i += (size_t)gridDim.x * 256) {
        if (NT) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
        else dst[i] = src[i];
    }
}

// Read-only: the loaded values feed a sum that is stored only if it equals a
// value the data never sums to, so no load can be dropped.
template <int V, bool NT>
__global__ void __launch_bounds__(256) k_read_v(const dvec2 *__restrict__ src, size_t n2, double never,
                                                double *__restrict__ sink)
{
    const size_t base = (size_t)blockIdx.x * 256 * V + threadIdx.x;
    double acc = 0.0;
#pragma unroll
    for (int r = 0; r < V; ++r) {
        const size_t i = base + (size_t)r * 256;
        if (i < n2) {
            const dvec2 x = NT ? __builtin_nontemporal_load(src + i) : src[i];
            acc += x.x + x.y;
        }
    }
    if (acc == never) sink[threadIdx.x] = acc;
}

template <int V, bool NT>
__global__ void __launch_bounds__(256) k_write_v(dvec2 *__restrict__ dst, size_t n2, double v)
{
    const size_t base = (size_t)blockIdx.x * 256 * V + threadIdx.x;
    const dvec2 x = {v, v};
#pragma unroll
    for (int r = 0; r < V; ++r) {
        const size_t i = base + (size_t)r * 256;
        if (i < n2) {
            if (NT) __builtin_nontemporal_store(x, dst + i);
            else dst[i] = x;
        }
    }
}

int bandwidth_probe_streams(int kind)
{
    switch (kind) {
    case 0: case 1: case 2: case 3: case 4: case 5: return 2;
    case 6: case 7: case 8: case 9: return 1;
    default: return 0;
    }
}

hipError_t launch_bandwidth_probe(int kind, const double *src, double *dst, size_t n, int cus, hipStream_t st)
{
    const size_t n2 = n / 2;
    const auto grid = [&](int v) { return dim3((unsigned)((n2 + 256 * (size_t)v - 1) / (256 * (size_t)v))); };
    const dvec2 *s2 = (const dvec2 *)src;
    dvec2 *d2 = (dvec2 *)dst;
    switch (kind) {
    case 0: hipLaunchKernelGGL((k_copy_v<1, false>), grid(1), dim3(256), 0, st, s2, d2, n2); break;
    case 1: hipLaunchKernelGGL((k_copy_v<1, true>), grid(1), dim3(256), 0, st, s2, d2, n2); break;
    case 2: hipLaunchKernelGGL((k_copy_v<4, false>), grid(4), dim3(256), 0, st, s2, d2, n2); break;
    case 3: hipLaunchKernelGGL((k_copy_v<4, true>), grid(4), dim3(256), 0, st, s2, d2, n2); break;
    case 4: hipLaunchKernelGGL((k_copy_gs<false>), dim3(cus * 8), dim3(256), 0, st, s2, d2, n2); break;
    case 5: hipLaunchKernelGGL((k_copy_gs<true>), dim3(cus * 8), dim3(256), 0, st, s2, d2, n2); break;
    case 6: hipLaunchKernelGGL((k_read_v<4, false>), grid(4), dim3(256), 0, st, s2, n2, -1.0, dst); break;
    case 7: hipLaunchKernelGGL((k_read_v<4, true>), grid(4), dim3(256), 0, st, s2, n2, -1.0, dst); break;
    case 8: hipLaunchKernelGGL((k_write_v<4, false>), grid(4), dim3(256), 0, st, d2, n2, 0.5); break;
    case 9: hipLaunchKernelGGL((k_write_v<4, true>), grid(4), dim3(256), 0, st, d2, n2, 0.5); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Faithful mode on the fused engine: Streaming's boundary phase on the ring
// (the first ghost layer), in the reference's order, before the fused pass
// reads it.  Interior values come from the post-collision lattice P.src.
// Ghost cells deeper than the first layer are never written, so they keep
// Initialize's values: 0 in the cavity, whose fields start at rest with
// p = 0; in the sphere channel the initial fields reach into the ghosts
// (sphere/src/lbm.cpp:224-239 loops over the whole ghosted box), and those
// values are recomputed where a wall reads them.

// the faithful sphere's initial population q at cell (i, j, k) of the
// ghosted box, in interior coordinates (k_init's kind 2 at any cell)
__device__ __forceinline__ double faithful_sphere_f0(const FusedParams &P, uint64_t seed, int q, int i, int j, int k)
{
    const uint64_t ly = (uint64_t)P.ny + 6, lz = (uint64_t)P.nz + 6;
    const uint64_t idx = ((uint64_t)(i + 3) * ly + (uint64_t)(j + 3)) * lz + (uint64_t)(k + 3);
    const int ran = (int)(mix64(seed + idx * 0x9E3779B97F4A7C15ULL) % 100ULL);
    const int usr = faithful_sphere_solid(P.nx, i, j, k) ? 0 : 1;
    const int glz = P.nz, kz = k;
    const double gm = (double)(glz - 1);
    const double u = (P.u0 * 4.0 * (double)kz * (double)(glz - 1 - kz) / (gm * gm)) * (double)usr * (1.0 + (double)ran * 0.001);
    return feq_f64(q, 0.0, u, 0.0, 0.0);
}

// Where population q of cell (i, j, k) is read from, in two steps so that a
// kernel can first place all its sources (the deep-ghost value is computed
// under a branch, with no load in it), then issue all its loads back to back
// (a value loaded before a divergent branch and still live after it would be
// waited for there), then pick.
struct GhostSrc {
    const double *p;  // always a valid address
    double deep;      // the value when the cell lies deeper than the first layer
    bool is_deep;
};

__device__ __forceinline__ GhostSrc faithful_src(const FusedParams &P, int q, int i, int j, int k)
{
    const int64_t r = ring_index(P, i, j, k);
    GhostSrc g;
    g.is_deep = r == -1;
    g.deep = 0.0;
    if (g.is_deep && P.faithful_sphere) g.deep = faithful_sphere_f0(P, P.seed, q, i, j, k);
    const uint64_t in_cell = r == -2 ? ((uint64_t)k * P.ny + j) * P.nx + i : 0;
    g.p = r == -2 ? (const double *)P.src + (uint64_t)q * P.pstride + in_cell
                  : P.ring + (uint64_t)q * P.ring_n + (uint64_t)(r < 0 ? 0 : r);
    return g;
}

// Initialize for the ring: every first-layer ghost cell takes its initial
// populations (0 in the cavity).
__global__ void k_faithful_ring_init(const FusedParams P)
{
    uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= P.ring_n) return;
    const uint32_t rr = r;
    int i, j, k;
    if (r < 2 * P.ring_a) {
        i = r < P.ring_a ? -1 : P.nx;
        r -= r < P.ring_a ? 0 : P.ring_a;
        j = (int)(r % (P.ny + 2)) - 1;
        k = (int)(r / (P.ny + 2)) - 1;
    } else if ((r -= 2 * P.ring_a) < 2 * P.ring_b) {
        j = r < P.ring_b ? -1 : P.ny;
        r -= r < P.ring_b ? 0 : P.ring_b;
        i = (int)(r % P.nx);
        k = (int)(r / P.nx) - 1;
    } else {
        r -= 2 * P.ring_b;
        k = r < P.ring_c ? -1 : P.nz;
        r -= r < P.ring_c ? 0 : P.ring_c;
        i = (int)(r % P.nx);
        j = (int)(r / P.nx);
    }
#pragma unroll
    for (int q = 0; q < Q; ++q)
        P.ring[(uint64_t)q * P.ring_n + rr] = P.faithful_sphere ? faithful_sphere_f0(P, P.seed, q, i, j, k) : 0.0;
}

hipError_t launch_faithful_ring_init(const FusedParams &P, hipStream_t st)
{
    hipLaunchKernelGGL(k_faithful_ring_init, dim3((P.ring_n + 255) / 256), dim3(256), 0, st, P);
    return hipGetLastError();
}

// The one-rank unpack of lid_driven_cavity/src/mpi_view_transfer.cpp:617-688
// writes receive buffers nothing filled (zeros) into the ghost planes
// y = ny, z = nz, x = -1 and x = nx, all 27 populations: here their cells of
// the first layer.  Of a cell's 27 entries only those a wall writes later in
// the step (the directions pointing into the cavity across any of the cell's
// ghost axes) ever hold anything but zero: Initialize zeroes the ring and
// nothing else writes it.  Those are zeroed again here; the others already
// are.  Thread t: the two x faces, then the plane y = ny (x in [0, nx)),
// then the plane z = nz (x in [0, nx), y in [-1, ny]).
__global__ void k_faithful_unpack(const FusedParams P)
{
    const uint32_t t0 = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t nyz = (uint32_t)(P.ny + 2) * (P.nz + 2), nxz = (uint32_t)P.nx * (P.nz + 2),
                   nxy = (uint32_t)P.nx * (P.ny + 2);
    uint32_t t = t0;
    int i, j, k;
    if (t < 2 * nyz) {
        const uint32_t f = t / nyz, r = t - f * nyz;
        i = f ? P.nx : -1;
        k = (int)(r / (P.ny + 2)) - 1;
        j = (int)(r % (P.ny + 2)) - 1;
    } else if ((t -= 2 * nyz) < nxz) {
        j = P.ny;
        k = (int)(t / P.nx) - 1;
        i = (int)(t % P.nx);
    } else if ((t -= nxz) < nxy) {
        k = P.nz;
        j = (int)(t / P.nx) - 1;
        i = (int)(t % P.nx);
    } else {
        return;
    }
    const uint64_t r = (uint64_t)ring_index(P, i, j, k);
    // the inward direction across each ghost axis of the cell (0: not a
    // ghost coordinate)
    const int sx = i < 0 ? 1 : (i >= P.nx ? -1 : 0);
    const int sy = j < 0 ? 1 : (j >= P.ny ? -1 : 0);
    const int sz = k < 0 ? 1 : (k >= P.nz ? -1 : 0);
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const bool walled = (sx != 0 && EX[q] == sx) || (sy != 0 && EY[q] == sy) || (sz != 0 && EZ[q] == sz);
        if (walled) P.ring[(uint64_t)q * P.ring_n + r] = 0.0;
    }
}

// One wall of lid_driven_cavity/src/lbm.cpp:233-301: for the first-layer
// cells (a, b) of the plane and every direction q pointing into the cavity,
// f_q(plane) = f_bb(q)(src_plane, a + sa e_a, b + sb e_b).  The sources lie
// in src_plane and the targets in plane, so no source is a target.
// Thread t of row blockIdx.y: cell t of the plane, its blockIdx.y-th such q
// (a row per direction: the pairs are independent, and nine times the lanes
// hide the latency of the scattered loads).
template <int AXIS, int SIGN>
__device__ __forceinline__ void faithful_wall(const FusedParams &P, int plane, int src_plane, int sa, int sb)
{
    const int la = AXIS == 0 ? P.ny : P.nx, lb = AXIS == 2 ? P.ny : P.nz;
    const int na = la + 2, nb = lb + 2;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= na * nb) return;
    const int q = nth_dir<AXIS, SIGN>(blockIdx.y);
    int e[3], bq;
    dir_of(q, e, bq);
    const int A = t % na - 1, B = t / na - 1;
    int si, sj, sk, di, dj, dk;
    if (AXIS == 0) {
        si = src_plane; sj = A + sa * e[1]; sk = B + sb * e[2];
        di = plane; dj = A; dk = B;
    } else if (AXIS == 1) {
        si = A + sa * e[0]; sj = src_plane; sk = B + sb * e[2];
        di = A; dj = plane; dk = B;
    } else {
        si = A + sa * e[0]; sj = B + sb * e[1]; sk = src_plane;
        di = A; dj = B; dk = plane;
    }
    double v;
    if (AXIS == 0 && P.stage != nullptr && sj >= 0 && sj < P.ny && sk >= 0 && sk < P.nz) {
        // an interior cell of plane 1 or nx - 2: from the copy the last fused
        // pass wrote, y fastest (one strided cell per row of the lattice
        // otherwise)
        const int side = SIGN > 0 ? 0 : 1;
        v = P.stage[(((uint64_t)side * P.nz + sk) * P.ny + sj) * STAGE_ROW + x_rank(bq)];
    } else {
        const GhostSrc g = faithful_src(P, bq, si, sj, sk);
        v = *g.p;
        v = g.is_deep ? g.deep : v;
    }
    const uint64_t r = (uint64_t)ring_index(P, di, dj, dk);
    P.ring[(uint64_t)q * P.ring_n + r] = v;
}

// FusedParams::stage from the lattice P.src: thread t is entry t of the
// stage (side, z, y, r).
__global__ void k_faithful_stage_fill(const FusedParams P)
{
    const uint64_t plane = (uint64_t)P.ny * P.nz;
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 2 * plane * STAGE_ROW) return;
    const int r = (int)(t % STAGE_ROW);
    const uint64_t row = t / STAGE_ROW;  // side * plane + (z ny + y)
    const int side = row >= plane ? 1 : 0;
    const uint64_t yz = row - (side ? plane : 0);
    int q;
    uint32_t x;
    if (P.faithful_sphere) {  // inlet x = 0, e_x > 0; outlet x = nx - 1, e_x < 0
        q = side == 0 ? nth_dir<0, 1>(r) : nth_dir<0, -1>(r);
        x = side == 0 ? 0u : (uint32_t)P.nx - 1u;
    } else {  // walls: x = 1, e_x < 0; x = nx - 2, e_x > 0
        q = side == 0 ? nth_dir<0, -1>(r) : nth_dir<0, 1>(r);
        x = side == 0 ? 1u : (uint32_t)P.nx - 2u;
    }
    P.stage[t] = ((const double *)P.src)[(uint64_t)q * P.pstride + yz * P.nx + x];
}

hipError_t launch_faithful_stage_fill(const FusedParams &P, hipStream_t st)
{
    const uint64_t n = 2ull * P.ny * P.nz * STAGE_ROW;
    hipLaunchKernelGGL(k_faithful_stage_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, P);
    return hipGetLastError();
}

// The two walls of one axis in one launch: blockIdx.z = 0 the low wall
// (plane -1, sources on plane 1), 1 the high wall (plane n, sources on plane
// n - 2; its b-stride sb_high).  With n >= 2 both source planes lie inside the
// domain on this axis, so neither wall reads what the other writes and their
// order does not matter; the axes still follow the reference's order, since
// a later axis reads edge cells an earlier one wrote.
template <int AXIS>
__global__ void k_faithful_walls(const FusedParams P, int sb_high)
{
    const int n = AXIS == 0 ? P.nx : AXIS == 1 ? P.ny : P.nz;
    if (blockIdx.z == 0) faithful_wall<AXIS, 1>(P, -1, 1, 2, 2);
    else faithful_wall<AXIS, -1>(P, n, n - 2, 2, sb_high);
}

// sphere/src/lbm.cpp:326-347: the inlet ghost plane takes the first
// interior plane's populations with e_x > 0, the outlet ghost plane takes
// f_q(nx - 1, y + e_y, z + e_z) for e_x < 0; first-layer cells only (the
// sources lie on x = 0 or nx - 1, the targets on x = -1 or nx).  Thread t of
// row blockIdx.y: cell t of the plane, its blockIdx.y-th such q.
template <int OUTLET>
__global__ void k_faithful_inlet_outlet(const FusedParams P)
{
    const int na = P.ny + 2, nb = P.nz + 2;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= na * nb) return;
    const int q = OUTLET ? nth_dir<0, -1>(blockIdx.y) : nth_dir<0, 1>(blockIdx.y);
    int e[3], bq;
    dir_of(q, e, bq);
    const int j = t % na - 1, k = t / na - 1;
    const int sj = OUTLET ? j + e[1] : j, sk = OUTLET ? k + e[2] : k;
    double v;
    if (P.stage != nullptr && sj >= 0 && sj < P.ny && sk >= 0 && sk < P.nz) {
        // an interior cell of plane 0 or nx - 1: from the copy the last fused
        // pass wrote
        v = P.stage[(((uint64_t)OUTLET * P.nz + sk) * P.ny + sj) * STAGE_ROW + x_rank(q)];
    } else {
        const GhostSrc g = faithful_src(P, q, OUTLET ? P.nx - 1 : 0, sj, sk);
        v = *g.p;
        v = g.is_deep ? g.deep : v;
    }
    const uint64_t r = (uint64_t)ring_index(P, OUTLET ? P.nx : -1, j, k);
    P.ring[(uint64_t)q * P.ring_n + r] = v;
}

// "usrbc", sphere/src/lbm.cpp:350-356: a solid cell takes, for each
// direction q whose neighbour x + e_q is fluid, f_bb(q)(x + 2 e_q), in place
// in the post-collision lattice.  The reference's mask covers the ghost
// cells too: where the ball reaches the first ghost layer (ny or nz within r
// of its centre), a ghost neighbour can be solid, and its mask value is
// recomputed here.  A source is never a target: between two solid voxels of a
// ball the midpoint voxel is solid too, so that direction is not taken.  The
// (cell, q) pairs are therefore independent: thread t of row q = blockIdx.y
// takes cell t of the box [lo, lo + n)^3 around the ball (clamped to the
// cavity), so that the few thousand cells of the box fill 27 times as many
// lanes.  Every thread loads (threads off the box from a valid cell) and only
// solid cells store: no load sits under a divergent branch.
__global__ void k_faithful_usr(const FusedParams P, int lo, int n)
{
    const int q = blockIdx.y;
    int ex = 0, ey = 0, ez = 0, bq = 0;
#pragma unroll
    for (int k = 0; k < Q; ++k)
        if (k == q) {
            ex = EX[k];
            ey = EY[k];
            ez = EZ[k];
            bq = BB[k];
        }
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int tt = t < n * n * n ? t : 0;
    const int x0 = lo + tt % n, y0 = lo + (tt / n) % n, z0 = lo + tt / (n * n);
    const bool valid = (t < n * n * n) & (x0 < P.nx) & (y0 < P.ny) & (z0 < P.nz);
    const int x = valid ? x0 : 0, y = valid ? y0 : 0, z = valid ? z0 : 0;
    const uint32_t c = ((uint32_t)z * P.ny + y) * P.nx + x;
    const int xi = x + ex, yi = y + ey, zi = z + ez;
    const bool in = (xi >= 0) & (xi < P.nx) & (yi >= 0) & (yi < P.ny) & (zi >= 0) & (zi < P.nz);
    const uint64_t mi = in ? ((uint64_t)zi * P.ny + yi) * P.nx + xi : (uint64_t)c;
    const GhostSrc g = faithful_src(P, bq, x + 2 * ex, y + 2 * ey, z + 2 * ez);
    const uint8_t mc = P.mask[c];
    const uint8_t m = P.mask[mi];
    double v = *g.p;
    v = g.is_deep ? g.deep : v;
    const bool solid = valid & (mc == 1);
    const bool nb_solid = in ? m == 1 : faithful_sphere_solid(P.nx, xi, yi, zi);
    if (solid & !nb_solid) ((double *)P.src)[(uint64_t)q * P.pstride + c] = v;
}

hipError_t launch_faithful_boundaries(const FusedParams &P, hipStream_t st)
{
    const int nx = P.nx, ny = P.ny, nz = P.nz;
    // nine rows of workgroups: one per direction entering the domain
    auto grid = [](int la, int lb) { return dim3((unsigned)(((la + 2) * (lb + 2) + 255) / 256), 9); };
    // the same, times the two walls of an axis
    auto grid2 = [](int la, int lb) { return dim3((unsigned)(((la + 2) * (lb + 2) + 255) / 256), 9, 2); };
    if (P.faithful_sphere) {
        // bcf bcb bcd bcu, inlet, outlet, usrbc (sphere/src/lbm.cpp:277-356);
        // one rank: the unpack has nothing to write
        hipLaunchKernelGGL(k_faithful_walls<1>, grid2(nx, nz), dim3(256), 0, st, P, 2);
        hipLaunchKernelGGL(k_faithful_walls<2>, grid2(nx, ny), dim3(256), 0, st, P, 2);
        hipLaunchKernelGGL(k_faithful_inlet_outlet<0>, grid(ny, nz), dim3(256), 0, st, P);
        hipLaunchKernelGGL(k_faithful_inlet_outlet<1>, grid(ny, nz), dim3(256), 0, st, P);
        // the ball lies within [c - r, c + r] of its centre c on every axis
        const int cc = (nx - 1) / 8, rr = (int)std::ceil((nx - 1) / 64.0);
        const int lo = cc - rr < 0 ? 0 : cc - rr, n = cc + rr - lo + 1;
        hipLaunchKernelGGL(k_faithful_usr, dim3((unsigned)((n * n * n + 255) / 256), Q), dim3(256), 0, st, P, lo, n);
        return hipGetLastError();
    }
    const uint32_t nun = 2u * (P.ny + 2) * (P.nz + 2) + (uint32_t)P.nx * (P.nz + 2) + (uint32_t)P.nx * (P.ny + 2);
    hipLaunchKernelGGL(k_faithful_unpack, dim3((nun + 255) / 256), dim3(256), 0, st, P);
    // bcl bcr, bcf bcb, bcd bcu; the top wall reads j + e_y (:298)
    hipLaunchKernelGGL(k_faithful_walls<0>, grid2(ny, nz), dim3(256), 0, st, P, 2);
    hipLaunchKernelGGL(k_faithful_walls<1>, grid2(nx, nz), dim3(256), 0, st, P, 2);
    hipLaunchKernelGGL(k_faithful_walls<2>, grid2(nx, ny), dim3(256), 0, st, P, 1);
    return hipGetLastError();
}

}  // namespace imx
