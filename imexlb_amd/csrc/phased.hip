// The phased engine: one kernel per phase of the reference, on its ghost-3
// layout, FP64, compiled with -ffp-contract=off.
//
//   Initialize  tgv/src/lbm.cpp:195-218, lid_driven_cavity/src/lbm.cpp:191-210,
//               sphere/src/lbm.cpp:221-254
//   Collision   tgv/src/lbm.cpp:222-237
//   pack/exchange/unpack  one rank: tgv/src/mpi_view_transfer.cpp:117-383
//               (every ghost becomes the periodic image of the interior);
//               lid_driven_cavity/src/mpi_view_transfer.cpp:598-690 (four ghost
//               planes receive buffers nothing wrote: zeros);
//               sphere/src/mpi_view_transfer.cpp:364-439 (no effect)
//   Streaming   boundary kernels in the reference's order, then stream1 and
//               stream2 as a swap (tgv/src/lbm.cpp:239-256,
//               lid_driven_cavity/src/lbm.cpp:231-317, sphere/src/lbm.cpp:274-371)
//   Update      tgv/src/lbm.cpp:288-333, lid_driven_cavity/src/lbm.cpp:319-349,
//               sphere/src/lbm.cpp:374-417
//
// Layout: f[i][k][j][x] with every index over the ghosted extent
// (l = n + 6), x fastest -- structure of arrays instead of the reference's
// CUDA LayoutLeft f(i, x, y, z), which keeps the 27 populations of a cell
// together.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <utility>
#include <vector>

#include "fused.h"
#include "phased.h"

namespace imx {

namespace {

constexpr int G = GHOST;

struct Dims {
    int lx, ly, lz;
    int nx, ny, nz;
    uint64_t ncell;  // lx * ly * lz, below 2^32 (checked at create)
    FastDiv divx, divy;  // by lx and ly: cell index -> (i, j, k) without 64-bit division
    int case_id, mode;
    double tau0, rho0, u0, lid_u;
    double cx, cy, cz, r;
    uint64_t seed;
};

__host__ __device__ __forceinline__ uint64_t cidx(const Dims &d, int i, int j, int k)
{
    return ((uint64_t)k * d.ly + j) * d.lx + i;
}
__host__ __device__ __forceinline__ uint64_t fidx(const Dims &d, int q, int i, int j, int k)
{
    return (uint64_t)q * d.ncell + cidx(d, i, j, k);
}
__device__ __forceinline__ void decode(const Dims &d, uint64_t c, int &i, int &j, int &k)
{
    const uint32_t c32 = (uint32_t)c;
    const uint32_t t = d.divx.div(c32);
    i = (int)(c32 - t * (uint32_t)d.lx);
    const uint32_t kk = d.divy.div(t);
    j = (int)(t - kk * (uint32_t)d.ly);
    k = (int)kk;
}
__device__ __forceinline__ bool interior(const Dims &d, int i, int j, int k)
{
    return i >= G && i < G + d.nx && j >= G && j < G + d.ny && k >= G && k < G + d.nz;
}

__device__ __forceinline__ double feq_ref(int q, double p, double u, double v, double w)
{
    const double edu = edu_of<double>(q, u, v, w);
    const double udu = u * u + v * v + w * w;
    const double eu2 = edu * edu;
    return WVAL[WCLASS[q]] * p * 3.0 + WVAL[WCLASS[q]] * (3.0 * edu + 4.5 * eu2 - 1.5 * udu);
}

struct Arrays {
    double *f, *ft, *ua, *va, *wa, *p;
    int *usr;
    const double *tab;  // TGV: per-axis sin, cos, cos(2.) over the ghosted extent
};

__global__ void k_init(const Dims d, const Arrays a)
{
    const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= d.ncell) return;
    int i, j, k;
    decode(d, c, i, j, k);
    double u = 0.0, v = 0.0, w = 0.0, p = 0.0;
    int usr = 1;
    if (d.case_id == IMEXLB_CASE_TGV) {
        const double *tx = a.tab, *ty = a.tab + 3 * d.lx, *tz = ty + 3 * d.ly;
        const double sx = tx[i], cx = tx[d.lx + i], c2x = tx[2 * d.lx + i];
        const double sy = ty[j], cy = ty[d.ly + j], c2y = ty[2 * d.ly + j];
        const double cz = tz[d.lz + k], c2z = tz[2 * d.lz + k];
        u = d.u0 * sx * cy * cz;
        v = -d.u0 * cx * sy * cz;
        w = 0;
        p = d.rho0 * (1.0 / 3.0) + d.rho0 * d.u0 * d.u0 / 16.0 * (c2x + c2y) * (c2z + 2.0);
    } else if (d.case_id == IMEXLB_CASE_CAVITY) {
        p = d.mode == IMEXLB_MODE_FAITHFUL ? 0.0 : d.rho0 * (1.0 / 3.0);
    } else if (d.mode == IMEXLB_MODE_FAITHFUL) {
        // sphere/src/lbm.cpp:224-239 with a seeded hash for urand(100)
        const uint64_t idx = ((uint64_t)i * (uint64_t)d.ly + (uint64_t)j) * (uint64_t)d.lz + (uint64_t)k;
        const int ran = (int)(mix64(d.seed + idx * 0x9E3779B97F4A7C15ULL) % 100ULL);
        const int glx = d.nx, glz = d.nz;
        const int di = i - G - (glx - 1) / 8, dj = j - G - (glx - 1) / 8, dk = k - G - (glx - 1) / 8;
        const double r = (glx - 1) / 64.0;
        usr = ((double)di * (double)di + (double)dj * (double)dj + (double)dk * (double)dk <= r * r) ? 0 : 1;
        const int kz = k - G;
        const double gm = (double)(glz - 1);
        u = (d.u0 * 4.0 * (double)kz * (double)(glz - 1 - kz) / (gm * gm)) * (double)usr * (1.0 + (double)ran * 0.001);
        p = 0;
    } else {
        const double dx = (double)(i - G) - d.cx, dy = (double)(j - G) - d.cy, dz = (double)(k - G) - d.cz;
        usr = (dx * dx + dy * dy + dz * dz <= d.r * d.r) ? 0 : 1;
        u = usr ? d.u0 : 0.0;
        p = d.rho0 * (1.0 / 3.0);
    }
    a.ua[c] = u;
    a.va[c] = v;
    a.wa[c] = w;
    a.p[c] = p;
    a.usr[c] = usr;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const double fq = feq_ref(q, p, u, v, w);
        a.f[(uint64_t)q * d.ncell + c] = fq;
        // ft starts as a copy of f: after the swaps of Streaming either array
        // is f, and ghost cells deeper than one layer, which the steps
        // rewrite before reading or never write, must agree in both
        a.ft[(uint64_t)q * d.ncell + c] = fq;
    }
}

// 27 values and 27 addresses live at once: the bound of 256 threads (every
// launch here) leaves room for them; the default bound caps the registers at
// 128 and spills three values to private memory
__global__ void __launch_bounds__(256) k_collide(const Dims d, const Arrays a)
{
    const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= d.ncell) return;
    int i, j, k;
    decode(d, c, i, j, k);
    if (!interior(d, i, j, k)) return;
    if (d.case_id == IMEXLB_CASE_SPHERE && d.mode == IMEXLB_MODE_CONFIG && a.usr[c] == 0) return;
    const double u = a.ua[c], v = a.va[c], w = a.wa[c], p = a.p[c];
    // all 27 loads before any store: a store may alias a later load as far as
    // the compiler knows, and would otherwise wait for each load in turn
    double f[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) f[q] = a.f[(uint64_t)q * d.ncell + c];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const double feq = feq_ref(q, p, u, v, w);
        a.f[(uint64_t)q * d.ncell + c] = f[q] - (f[q] - feq) / (d.tau0 + 0.5);
    }
}

// one rank, periodic on the axes flagged: every ghost cell on those axes
// takes the value of its periodic image
__global__ void k_periodic(const Dims d, const Arrays a, int px, int py, int pz)
{
    const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= d.ncell) return;
    int i, j, k;
    decode(d, c, i, j, k);
    int ii = i, jj = j, kk = k;
    if (px) ii = i < G ? i + d.nx : (i >= G + d.nx ? i - d.nx : i);
    if (py) jj = j < G ? j + d.ny : (j >= G + d.ny ? j - d.ny : j);
    if (pz) kk = k < G ? k + d.nz : (k >= G + d.nz ? k - d.nz : k);
    if (ii == i && jj == j && kk == k) return;
    const uint64_t s = cidx(d, ii, jj, kk);  // an interior cell: never a target here
    double v[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) v[q] = a.f[(uint64_t)q * d.ncell + s];
#pragma unroll
    for (int q = 0; q < Q; ++q) a.f[(uint64_t)q * d.ncell + c] = v[q];
}

// the planes y = le1, z = le2, x = ls0 - 1 and x = le0, one thread per cell
// of each plane (cells on two planes are zeroed twice).
// In the two x planes only the edge cells (y or z on the first ghost layer)
// are written; the rest of those planes already hold what the zero fill
// would leave there.  Their cells outside the first ghost layer are never
// written by any kernel and stay at the zeros of Initialize (f = 0 in the
// faithful cavity).  Inside it, the populations pointing into the domain are
// rewritten by bcl / bcr, the first two boundary kernels, before anything
// reads them (neither reads the other's plane: nx >= 2, checked at create),
// and the other 18 are written by no kernel: the y and z walls write only
// their own planes, and the carry-over copies values unchanged.  This saves
// a strided write of every population of every cell of both planes.
__global__ void k_cavity_zero_planes(const Dims d, const Arrays a)
{
    const uint32_t nyz = (uint32_t)d.ly * d.lz, nxz = (uint32_t)d.lx * d.lz, nxy = (uint32_t)d.lx * d.ly;
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    int i, j, k;
    if (t < 2u * nyz) {
        const uint32_t p = t / nyz, r = t - p * nyz;
        i = p == 0 ? G - 1 : G + d.nx;
        k = (int)(r / d.ly);
        j = (int)(r - (uint32_t)k * d.ly);
        const bool ring_j = j >= G - 1 && j <= G + d.ny, ring_k = k >= G - 1 && k <= G + d.nz;
        const bool edge = j == G - 1 || j == G + d.ny || k == G - 1 || k == G + d.nz;
        if (!(ring_j && ring_k && edge)) return;
    } else if ((t -= 2u * nyz) < nxz) {
        j = G + d.ny;
        k = (int)(t / d.lx);
        i = (int)(t - (uint32_t)k * d.lx);
    } else if ((t -= nxz) < nxy) {
        k = G + d.nz;
        j = (int)(t / d.lx);
        i = (int)(t - (uint32_t)j * d.lx);
    } else {
        return;
    }
    const uint64_t c = cidx(d, i, j, k);
#pragma unroll
    for (int q = 0; q < Q; ++q) a.f[(uint64_t)q * d.ncell + c] = 0.0;
}

// Faithful wall bounce-back on one face: for the cells (a, b) of the plane in
// [G-1, l-G+1)^2 and every direction pointing into the domain,
// f_q(plane) = f_bb(q)(src_plane, a + sa e_a, b + sb e_b).  Sources lie in
// src_plane and targets in plane, so the (cell, q) pairs are independent:
// thread t of row blockIdx.y takes cell t and the blockIdx.y-th such q (nine
// times the lanes of a thread per cell hide the latency of the scattered
// loads).  The axis and the inward sign are template arguments.
template <int AXIS, int SIGN>
__global__ void k_wall(const Dims d, const Arrays a, int plane, int src_plane, int sa, int sb)
{
    const int la = AXIS == 0 ? d.ly : d.lx;
    const int lb = AXIS == 2 ? d.ly : d.lz;
    const int na = la - 2 * (G - 1), nb = lb - 2 * (G - 1);
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= na * nb) return;
    const int q = nth_dir<AXIS, SIGN>(blockIdx.y);
    int e[3], bq;
    dir_of(q, e, bq);
    const int A = G - 1 + t % na, B = G - 1 + t / na;
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
    a.f[fidx(d, q, di, dj, dk)] = a.f[fidx(d, bq, si, sj, sk)];
}

// sphere/src/lbm.cpp:326-347.  Targets lie in the ghost plane, sources in
// the first (inlet) or last (outlet) interior plane; a row of workgroups per
// direction, as for the walls.
template <int OUTLET>
__global__ void k_sphere_inlet_outlet(const Dims d, const Arrays a)
{
    const int na = d.ly - 2 * (G - 1), nb = d.lz - 2 * (G - 1);
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= na * nb) return;
    const int q = OUTLET ? nth_dir<0, -1>(blockIdx.y) : nth_dir<0, 1>(blockIdx.y);
    int e[3], bq;
    dir_of(q, e, bq);
    const int j = G - 1 + t % na, k = G - 1 + t / na;
    const int le0 = G + d.nx;
    if (!OUTLET) a.f[fidx(d, q, G - 1, j, k)] = a.f[fidx(d, q, G, j, k)];
    else a.f[fidx(d, q, le0, j, k)] = a.f[fidx(d, q, le0 + e[0], j + e[1], k + e[2])];
}

// "usrbc", sphere/src/lbm.cpp:350-356, over the box [lo, lo + n)^3 of
// interior cells around the ball: thread t of row q = blockIdx.y takes cell t
// and direction q.  Every pair loads, the masks pick the store: no load waits
// under a branch.  A source is never a target: between two solid voxels of a
// ball the midpoint voxel is solid too.
__global__ void k_sphere_usr(const Dims d, const Arrays a, int lo, int n)
{
    const int q = blockIdx.y;
    int e[3], bq;
    dir_of(q, e, bq);
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * n * n) return;
    const int x = lo + t % n, y = lo + (t / n) % n, z = lo + t / (n * n);
    if (x >= d.nx || y >= d.ny || z >= d.nz) return;
    const int i = G + x, j = G + y, k = G + z;
    const bool solid = a.usr[cidx(d, i, j, k)] == 0;
    const bool take = a.usr[cidx(d, i + e[0], j + e[1], k + e[2])] == 1;
    const double v = a.f[fidx(d, bq, i + 2 * e[0], j + 2 * e[1], k + 2 * e[2])];
    if (solid & take) a.f[fidx(d, q, i, j, k)] = v;
}

// config-mode cavity: halfway bounce-back as a fill of the depth-1 ghost shell
__global__ void k_cavity_halfway(const Dims d, const Arrays a, const double *lidcorr)
{
    const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= d.ncell) return;
    int i, j, k;
    decode(d, c, i, j, k);
    if (i < G - 1 || i > G + d.nx || j < G - 1 || j > G + d.ny || k < G - 1 || k > G + d.nz) return;
    if (interior(d, i, j, k)) return;
    double v[Q], corr[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        v[q] = a.f[fidx(d, BB[q], i + EX[q], j + EY[q], k + EZ[q])];
        corr[q] = lidcorr[q];
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        if (!interior(d, i + EX[q], j + EY[q], k + EZ[q])) continue;
        const double lid = v[q] + corr[q];
        a.f[fidx(d, q, i, j, k)] = k == G + d.nz ? lid : v[q];
    }
}

// config-mode sphere: inlet equilibrium and zero-gradient outlet ghost planes
__global__ void k_channel_planes(const Dims d, const Arrays a, const double *feq_in)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= d.ly * d.lz) return;
    const int j = t % d.ly, k = t / d.ly;
    const int ls0 = G, le0 = G + d.nx;
    double v[Q], in[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        v[q] = a.f[fidx(d, q, le0 - 1, j, k)];
        in[q] = feq_in[q];
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        a.f[fidx(d, q, ls0 - 1, j, k)] = in[q];
        a.f[fidx(d, q, le0, j, k)] = v[q];
    }
}

// config-mode sphere: halfway bounce-back as a fill of solid cells next to fluid
__global__ void k_channel_solid(const Dims d, const Arrays a)
{
    const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= d.ncell) return;
    int i, j, k;
    decode(d, c, i, j, k);
    if (!interior(d, i, j, k) || a.usr[c] != 0) return;
    bool take[Q];
    double v[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int xi = i + EX[q], xj = j + EY[q], xk = k + EZ[q];
        take[q] = a.usr[cidx(d, xi, xj, xk)] == 1;
        v[q] = a.f[fidx(d, BB[q], xi, xj, xk)];
    }
#pragma unroll
    for (int q = 0; q < Q; ++q)
        if (take[q]) a.f[fidx(d, q, i, j, k)] = v[q];
}

__global__ void k_stream1(const Dims d, const Arrays a)
{
    const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= d.ncell) return;
    int i, j, k;
    decode(d, c, i, j, k);
    if (!interior(d, i, j, k)) return;
    double v[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) v[q] = a.f[fidx(d, q, i - EX[q], j - EY[q], k - EZ[q])];
#pragma unroll
    for (int q = 0; q < Q; ++q) a.ft[fidx(d, q, i, j, k)] = v[q];
}

// The copy-back of Streaming (tgv/src/lbm.cpp:250-253: f = ft over the
// interior; the config channel leaves its solid cells alone) done as a swap:
// ft already holds the streamed interior, so it becomes f, and only the cells
// the copy-back would not have touched -- the ghost shell and the channel's
// solid cells -- are carried over from the old f.  The new f is the same
// array, bit for bit, for a sixth of the copy-back's traffic at 128^3 (a
// thirtieth at 512^3).
__global__ void k_stream2_keep(const Dims d, const double *__restrict__ f_old, double *__restrict__ f_new,
                               const int *__restrict__ usr)
{
    const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= d.ncell) return;
    int i, j, k;
    decode(d, c, i, j, k);
    const bool solid_kept = d.case_id == IMEXLB_CASE_SPHERE && d.mode == IMEXLB_MODE_CONFIG && usr[c] == 0;
    if (interior(d, i, j, k) && !solid_kept) return;
#pragma unroll
    for (int q = 0; q < Q; ++q) f_new[(uint64_t)q * d.ncell + c] = f_old[(uint64_t)q * d.ncell + c];
}

// The first ghost layer only (every case but the config channel, which also
// keeps its solid cells).  Deeper ghost cells need no carrying over: every
// step rewrites the ones it reads before reading them (the periodic halo, the
// cavity's zeroed planes, the channel's inlet and outlet planes), and the
// rest keep their initial values, which both arrays hold from Initialize.
// Thread t maps to the t-th cell of the layer: the two z faces (with their
// edges and corners) first, then the two y faces and the two x faces.
__host__ __device__ __forceinline__ uint64_t layer_cells(const Dims &d)
{
    const uint64_t a = (uint64_t)d.nx + 2, b = (uint64_t)d.ny + 2;
    return 2 * a * b + 2 * a * d.nz + 2 * (uint64_t)d.ny * d.nz;
}

__global__ void k_layer_keep(const Dims d, const double *__restrict__ f_old, double *__restrict__ f_new)
{
    const uint64_t t0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t0 >= layer_cells(d)) return;
    uint32_t t = (uint32_t)t0;
    const uint32_t a = d.nx + 2, b = d.ny + 2, ny = d.ny, nz = d.nz;
    uint32_t i, j, k;
    if (t < 2u * a * b) {
        const uint32_t sd = t / (a * b), r = t - sd * (a * b);
        k = sd ? G + nz : G - 1;
        j = G - 1 + r / a;
        i = G - 1 + (r - (j - (G - 1)) * a);
    } else if ((t -= 2u * a * b) < 2u * a * nz) {
        const uint32_t sd = t / (a * nz), r = t - sd * (a * nz);
        j = sd ? G + ny : G - 1;
        k = G + r / a;
        i = G - 1 + (r - (k - G) * a);
    } else {
        t -= 2u * a * nz;
        const uint32_t sd = t / (ny * nz), r = t - sd * (ny * nz);
        i = sd ? G + d.nx : G - 1;
        k = G + r / ny;
        j = G + (r - (k - G) * ny);
    }
    const uint64_t c = cidx(d, (int)i, (int)j, (int)k);
    double v[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) v[q] = f_old[(uint64_t)q * d.ncell + c];
#pragma unroll
    for (int q = 0; q < Q; ++q) f_new[(uint64_t)q * d.ncell + c] = v[q];
}

__global__ void k_update(const Dims d, const Arrays a)
{
    const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= d.ncell) return;
    int i, j, k;
    decode(d, c, i, j, k);
    const bool tgv = d.case_id == IMEXLB_CASE_TGV;
    if (!tgv && !interior(d, i, j, k)) {
        a.ua[c] = 0.0;
        a.va[c] = 0.0;
        a.wa[c] = 0.0;
        a.p[c] = 0.0;
        return;
    }
    double f[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) f[q] = a.f[(uint64_t)q * d.ncell + c];
    double pm, um, vm, wm;
    moments_f64(f, pm, um, vm, wm);
    if (d.case_id == IMEXLB_CASE_SPHERE) {
        const double m = (double)a.usr[c];
        um = um * m;
        vm = vm * m;
        wm = wm * m;
        pm = pm * m;
    }
    if (d.mode == IMEXLB_MODE_FAITHFUL) {
        if (d.case_id == IMEXLB_CASE_CAVITY && k == G + d.nz - 1) um = 0.1;
        if (d.case_id == IMEXLB_CASE_SPHERE) {
            if (i == G) {
                const int kz = k - G, glz = d.nz;
                const double gm = (double)(glz - 1);
                um = d.u0 * 4.0 * (double)kz * (double)(glz - 1 - kz) / (gm * gm);
                vm = 0.0;
                wm = 0.0;
            }
            if (i == d.lx - G - 1) pm = 0.0;
        }
    }
    a.ua[c] = um;
    a.va[c] = vm;
    a.wa[c] = wm;
    a.p[c] = pm;
}

// Several steps at a time (imexlb_step): Streaming's pull, Update and the next
// step's Collision in one pass per interior cell, from `src` (post-collision
// f with its ghost shell already filled by the step's boundary kernels) into
// `dst`.  The moments and their overrides are k_update's, the collision is
// k_collide's, in the same operation order, so every value is the one the
// phase-by-phase sequence produces.  last = 1 on the last step leaves dst
// holding the streamed (pre-collision) populations, the state after Update,
// and writes the moments of the interior; earlier steps write no moments, as
// nothing reads them before the last Update (the collision takes them from
// registers), and ghost moments are never read.
//
// The same pass carries over what the copy-back leaves alone: every cell of
// the rows from the first ghost layer inwards (y and z) that is not an
// interior fluid cell is copied from src to dst.  That covers the first ghost
// layer and the channel's solid cells, and it also copies the x ghost columns
// of those rows, so each row is written whole and no line of dst is left half
// written.  Deeper ghost cells need no carrying over: every step rewrites the
// ones it reads before reading them (the periodic halo, the cavity's zeroed
// planes, the channel's inlet and outlet planes), and the rest keep their
// initial values, which both arrays hold from Initialize; copying one of
// those x columns therefore changes nothing.  Copy lanes take the same path
// as the others, with their own cell as every source, so the wavefront does
// not diverge.
__global__ void k_stream_update_collide(const Dims d, const Arrays a, const double *__restrict__ src,
                                        double *__restrict__ dst, int last)
{
    const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= d.ncell) return;
    int i, j, k;
    decode(d, c, i, j, k);
    if (j < G - 1 || j > G + d.ny || k < G - 1 || k > G + d.nz) return;
    const bool in = interior(d, i, j, k);
    const bool copy = !in || (d.case_id == IMEXLB_CASE_SPHERE && d.mode == IMEXLB_MODE_CONFIG && a.usr[c] == 0);
    // as in the fused engine: lines no other wavefront reads (e_x = 0) are
    // loaded non-temporally, and dst is written non-temporally
    double f[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const uint64_t own = (uint64_t)q * d.ncell + c;
        const int64_t back = EX[q] + (int64_t)EY[q] * d.lx + (int64_t)EZ[q] * d.lx * d.ly;  // x - e_q
        const double *p = src + (copy ? own : own - back);
        f[q] = EX[q] == 0 ? __builtin_nontemporal_load(p) : *p;
    }
    double pm, um, vm, wm;
    moments_f64(f, pm, um, vm, wm);
    if (d.case_id == IMEXLB_CASE_SPHERE) {
        const double m = (double)a.usr[c];
        um = um * m;
        vm = vm * m;
        wm = wm * m;
        pm = pm * m;
    }
    if (d.mode == IMEXLB_MODE_FAITHFUL) {
        if (d.case_id == IMEXLB_CASE_CAVITY && k == G + d.nz - 1) um = 0.1;
        if (d.case_id == IMEXLB_CASE_SPHERE) {
            if (i == G) {
                const int kz = k - G, glz = d.nz;
                const double gm = (double)(glz - 1);
                um = d.u0 * 4.0 * (double)kz * (double)(glz - 1 - kz) / (gm * gm);
                vm = 0.0;
                wm = 0.0;
            }
            if (i == d.lx - G - 1) pm = 0.0;
        }
    }
    if (last && in) {
        a.ua[c] = um;
        a.va[c] = vm;
        a.wa[c] = wm;
        a.p[c] = pm;
    }
    if (!last) {
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const double feq = feq_ref(q, pm, um, vm, wm);
            const double g = f[q] - (f[q] - feq) / (d.tau0 + 0.5);
            f[q] = copy ? f[q] : g;
        }
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) __builtin_nontemporal_store(f[q], dst + (uint64_t)q * d.ncell + c);
}

__global__ void k_interior_io(const Dims d, double *f, double *buf, int nq, int to_buf)
{
    const uint64_t n = (uint64_t)d.nx * d.ny * d.nz;
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * nq) return;
    const int q = (int)(t / n);
    const uint64_t r = t - (uint64_t)q * n;
    const int x = (int)(r % d.nx);
    const uint64_t s = r / d.nx;
    const int y = (int)(s % d.ny), z = (int)(s / d.ny);
    const uint64_t g = (uint64_t)q * d.ncell + cidx(d, x + G, y + G, z + G);
    if (to_buf) buf[t] = f[g];
    else f[g] = buf[t];
}

__global__ void __launch_bounds__(256) k_diag1(const Dims d, const Arrays a, double *partials)
{
    __shared__ double sh[2][256];
    const uint64_t n = d.ncell;
    const uint64_t per = (n + gridDim.x - 1) / gridDim.x;
    const uint64_t lo = blockIdx.x * per, hi = lo + per < n ? lo + per : n;
    double acc[2] = {0.0, 0.0};
    for (uint64_t c = lo + threadIdx.x; c < hi; c += blockDim.x) {
        int i, j, k;
        decode(d, c, i, j, k);
        if (!interior(d, i, j, k) || a.usr[c] == 0) continue;
        acc[0] += a.ua[c] * a.ua[c] + a.va[c] * a.va[c] + a.wa[c] * a.wa[c];
        acc[1] += 3.0 * a.p[c];
    }
    sh[0][threadIdx.x] = acc[0];
    sh[1][threadIdx.x] = acc[1];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + s];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partials[2 * blockIdx.x] = sh[0][0];
        partials[2 * blockIdx.x + 1] = sh[1][0];
    }
}

__global__ void __launch_bounds__(256) k_diag2(const double *partials, int nparts, double *out)
{
    __shared__ double sh[2][256];
    double a0 = 0.0, a1 = 0.0;
    for (int b = threadIdx.x; b < nparts; b += blockDim.x) {
        a0 += partials[2 * b];
        a1 += partials[2 * b + 1];
    }
    sh[0][threadIdx.x] = a0;
    sh[1][threadIdx.x] = a1;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + s];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = 0.5 * sh[0][0];
        out[1] = sh[1][0];
    }
}

}  // namespace

struct PhasedState {
    Dims d{};
    Arrays a{};
    double *tab = nullptr;
    double *lidcorr = nullptr, *feq_in = nullptr;
    double *partials = nullptr, *result = nullptr;
    int nparts = 512;
    hipStream_t st = nullptr;
    int64_t bytes = 0;
};

#define PH_CHK(expr)                                                            \
    do {                                                                        \
        hipError_t e_ = (expr);                                                 \
        if (e_ != hipSuccess) {                                                 \
            err = std::string(#expr) + ": " + hipGetErrorString(e_);            \
            return IMEXLB_EHIP;                                                 \
        }                                                                       \
    } while (0)

static dim3 grid_for(uint64_t n) { return dim3((unsigned)((n + 255) / 256)); }

int phased_create(const imexlb_config &c, hipStream_t st, PhasedState **out, std::string &err)
{
    PhasedState *s = new PhasedState;
    s->st = st;
    Dims &d = s->d;
    d.nx = c.nx;
    d.ny = c.ny;
    d.nz = c.nz;
    d.lx = c.nx + 2 * G;
    d.ly = c.ny + 2 * G;
    d.lz = c.nz + 2 * G;
    d.ncell = (uint64_t)d.lx * d.ly * d.lz;
    d.divx.init((uint32_t)d.lx);
    d.divy.init((uint32_t)d.ly);
    d.case_id = c.case_id;
    d.mode = c.mode;
    d.tau0 = c.tau0;
    d.rho0 = c.rho0;
    d.u0 = c.u0;
    d.lid_u = c.lid_u;
    d.cx = c.sphere_cx;
    d.cy = c.sphere_cy;
    d.cz = c.sphere_cz;
    d.r = c.sphere_r;
    d.seed = c.seed;
    *out = s;
    if (d.ncell >= (1ull << 32)) {
        err = "the ghosted phased-engine lattice must hold fewer than 2^32 cells";
        return IMEXLB_EUNSUPPORTED;
    }
    if (c.case_id == IMEXLB_CASE_SPHERE && c.mode == IMEXLB_MODE_CONFIG) {
        const double r = c.sphere_r;
        if (!(r > 0) || c.sphere_cx - r < 2 || c.sphere_cx + r > c.nx - 3 || c.sphere_cy - r < 1 ||
            c.sphere_cy + r > c.ny - 2 || c.sphere_cz - r < 1 || c.sphere_cz + r > c.nz - 2) {
            err = "sphere must lie inside the channel with a margin";
            return IMEXLB_EINVAL;
        }
    }
    const size_t fb = (size_t)Q * d.ncell * 8, cb = (size_t)d.ncell * 8;
    if (hipMalloc(&s->a.f, fb) != hipSuccess || hipMalloc(&s->a.ft, fb) != hipSuccess ||
        hipMalloc(&s->a.ua, cb) != hipSuccess || hipMalloc(&s->a.va, cb) != hipSuccess ||
        hipMalloc(&s->a.wa, cb) != hipSuccess || hipMalloc(&s->a.p, cb) != hipSuccess ||
        hipMalloc(&s->a.usr, (size_t)d.ncell * 4) != hipSuccess || hipMalloc(&s->lidcorr, Q * 8) != hipSuccess ||
        hipMalloc(&s->feq_in, Q * 8) != hipSuccess || hipMalloc(&s->partials, 2 * 8 * s->nparts) != hipSuccess ||
        hipMalloc(&s->result, 2 * 8) != hipSuccess) {
        err = "phased engine: device allocation failed";
        return IMEXLB_ENOMEM;
    }
    s->bytes = 2 * fb + 4 * cb + d.ncell * 4;
    // host-side constants in the oracle's operation order
    double corr[Q], fin[Q];
    const double pin = c.rho0 * (1.0 / 3.0);
    for (int q = 0; q < Q; ++q) {
        corr[q] = 6.0 * w(q) * c.rho0 * ((double)EX[q] * c.lid_u);
        const double u = c.u0, v = 0.0, ww = 0.0;
        const double edu = (double)EX[q] * u + (double)EY[q] * v + (double)EZ[q] * ww;
        const double udu = u * u + v * v + ww * ww;
        const double eu2 = edu * edu;
        fin[q] = w(q) * pin * 3.0 + w(q) * (3.0 * edu + 4.5 * eu2 - 1.5 * udu);
    }
    PH_CHK(hipMemcpy(s->lidcorr, corr, sizeof corr, hipMemcpyHostToDevice));
    PH_CHK(hipMemcpy(s->feq_in, fin, sizeof fin, hipMemcpyHostToDevice));
    if (c.case_id == IMEXLB_CASE_TGV) {
        // tgv/src/lbm.cpp:197-204 over the ghosted extent (x_lo = 0)
        std::vector<double> tab(3 * ((size_t)d.lx + d.ly + d.lz));
        const int ext[3] = {d.lx, d.ly, d.lz}, gl[3] = {d.nx, d.ny, d.nz};
        size_t o = 0;
        for (int ax = 0; ax < 3; ++ax) {
            const int n = ext[ax];
            for (int k = 0; k < n; ++k) {
                const double a = (double)(k - G) / (double)gl[ax] * 2.0 * PI_REF;
                tab[o + k] = std::sin(a);
                tab[o + n + k] = std::cos(a);
                tab[o + 2 * n + k] = std::cos(a * 2.0);
            }
            o += 3 * (size_t)n;
        }
        PH_CHK(hipMalloc(&s->tab, tab.size() * 8));
        PH_CHK(hipMemcpy(s->tab, tab.data(), tab.size() * 8, hipMemcpyHostToDevice));
        s->a.tab = s->tab;
    }
    return IMEXLB_OK;
}

void phased_destroy(PhasedState *s)
{
    if (!s) return;
    for (void *p : {(void *)s->a.f, (void *)s->a.ft, (void *)s->a.ua, (void *)s->a.va, (void *)s->a.wa,
                    (void *)s->a.p, (void *)s->a.usr, (void *)s->tab, (void *)s->lidcorr, (void *)s->feq_in,
                    (void *)s->partials, (void *)s->result})
        if (p) hipFree(p);
    delete s;
}

int64_t phased_device_bytes(const PhasedState *s) { return s->bytes; }

int phased_initialize(PhasedState *s, std::string &err)
{
    hipLaunchKernelGGL(k_init, grid_for(s->d.ncell), dim3(256), 0, s->st, s->d, s->a);
    PH_CHK(hipGetLastError());
    return IMEXLB_OK;
}

int phased_collide(PhasedState *s, std::string &err)
{
    hipLaunchKernelGGL(k_collide, grid_for(s->d.ncell), dim3(256), 0, s->st, s->d, s->a);
    PH_CHK(hipGetLastError());
    return IMEXLB_OK;
}

int phased_pack(PhasedState *s, std::string &err)
{
    (void)s;
    (void)err;
    return IMEXLB_OK;  // the self-copy below reads the interior directly
}

int phased_exchange(PhasedState *s, std::string &err)
{
    (void)s;
    (void)err;
    return IMEXLB_OK;  // one rank: every message is to itself
}

int phased_unpack(PhasedState *s, std::string &err)
{
    const Dims &d = s->d;
    if (d.case_id == IMEXLB_CASE_TGV)
        hipLaunchKernelGGL(k_periodic, grid_for(d.ncell), dim3(256), 0, s->st, d, s->a, 1, 1, 1);
    else if (d.case_id == IMEXLB_CASE_CAVITY && d.mode == IMEXLB_MODE_FAITHFUL)
        hipLaunchKernelGGL(k_cavity_zero_planes, grid_for(2ull * d.ly * d.lz + (uint64_t)d.lx * d.lz + (uint64_t)d.lx * d.ly),
                           dim3(256), 0, s->st, d, s->a);
    else if (d.case_id == IMEXLB_CASE_SPHERE && d.mode == IMEXLB_MODE_CONFIG)
        hipLaunchKernelGGL(k_periodic, grid_for(d.ncell), dim3(256), 0, s->st, d, s->a, 0, 1, 1);
    PH_CHK(hipGetLastError());
    return IMEXLB_OK;
}

static void wall(PhasedState *s, int axis, int plane, int src_plane, int sign, int sa, int sb)
{
    const Dims &d = s->d;
    const int la = axis == 0 ? d.ly : d.lx, lb = axis == 2 ? d.ly : d.lz;
    const uint64_t n = (uint64_t)(la - 2 * (G - 1)) * (lb - 2 * (G - 1));
    dim3 g = grid_for(n), b(256);
    g.y = 9;  // a row of workgroups per direction entering the domain
#define IMX_WALL(AX, SG) hipLaunchKernelGGL((k_wall<AX, SG>), g, b, 0, s->st, d, s->a, plane, src_plane, sa, sb)
    if (axis == 0) {
        if (sign > 0) IMX_WALL(0, 1);
        else IMX_WALL(0, -1);
    } else if (axis == 1) {
        if (sign > 0) IMX_WALL(1, 1);
        else IMX_WALL(1, -1);
    } else {
        if (sign > 0) IMX_WALL(2, 1);
        else IMX_WALL(2, -1);
    }
#undef IMX_WALL
}

// Carry over what the copy-back leaves alone from s->a.f to s->a.ft.
static void keep(PhasedState *s)
{
    const Dims &d = s->d;
    if (d.case_id == IMEXLB_CASE_SPHERE && d.mode == IMEXLB_MODE_CONFIG)
        hipLaunchKernelGGL(k_stream2_keep, grid_for(d.ncell), dim3(256), 0, s->st, d, s->a.f, s->a.ft, s->a.usr);
    else
        hipLaunchKernelGGL(k_layer_keep, grid_for(layer_cells(d)), dim3(256), 0, s->st, d, s->a.f, s->a.ft);
}

// Streaming's boundary kernels, in the reference's order, on s->a.f.
static void boundaries(PhasedState *s)
{
    const Dims &d = s->d;
    const int ls = G, le0 = G + d.nx, le1 = G + d.ny, le2 = G + d.nz;
    if (d.case_id == IMEXLB_CASE_CAVITY) {
        if (d.mode == IMEXLB_MODE_FAITHFUL) {
            // lid_driven_cavity/src/lbm.cpp:233-301: bcl bcr bcf bcb bcd bcu
            wall(s, 0, ls - 1, ls + 1, +1, 2, 2);
            wall(s, 0, le0, le0 - 2, -1, 2, 2);
            wall(s, 1, ls - 1, ls + 1, +1, 2, 2);
            wall(s, 1, le1, le1 - 2, -1, 2, 2);
            wall(s, 2, ls - 1, ls + 1, +1, 2, 2);
            wall(s, 2, le2, le2 - 2, -1, 2, 1);  // top wall reads j + e_y (:298)
        } else {
            hipLaunchKernelGGL(k_cavity_halfway, grid_for(d.ncell), dim3(256), 0, s->st, d, s->a, s->lidcorr);
        }
    } else if (d.case_id == IMEXLB_CASE_SPHERE) {
        if (d.mode == IMEXLB_MODE_FAITHFUL) {
            // sphere/src/lbm.cpp:277-356: bcf bcb bcd bcu, inlet, outlet, usrbc
            wall(s, 1, ls - 1, ls + 1, +1, 2, 2);
            wall(s, 1, le1, le1 - 2, -1, 2, 2);
            wall(s, 2, ls - 1, ls + 1, +1, 2, 2);
            wall(s, 2, le2, le2 - 2, -1, 2, 2);
            const uint64_t n = (uint64_t)(d.ly - 2 * (G - 1)) * (d.lz - 2 * (G - 1));
            dim3 gio = grid_for(n);
            gio.y = 9;
            hipLaunchKernelGGL(k_sphere_inlet_outlet<0>, gio, dim3(256), 0, s->st, d, s->a);
            hipLaunchKernelGGL(k_sphere_inlet_outlet<1>, gio, dim3(256), 0, s->st, d, s->a);
            // the ball lies within [c - r, c + r] of its centre c on every axis
            const int cc = (d.nx - 1) / 8, rr = (int)std::ceil((d.nx - 1) / 64.0);
            const int lo = cc - rr < 0 ? 0 : cc - rr, nb = cc + rr - lo + 1;
            dim3 gu = grid_for((uint64_t)nb * nb * nb);
            gu.y = Q;
            hipLaunchKernelGGL(k_sphere_usr, gu, dim3(256), 0, s->st, d, s->a, lo, nb);
        } else {
            hipLaunchKernelGGL(k_channel_planes, grid_for((uint64_t)d.ly * d.lz), dim3(256), 0, s->st, d, s->a,
                               s->feq_in);
            hipLaunchKernelGGL(k_channel_solid, grid_for(d.ncell), dim3(256), 0, s->st, d, s->a);
        }
    }
}

int phased_stream(PhasedState *s, std::string &err)
{
    const Dims &d = s->d;
    boundaries(s);
    hipLaunchKernelGGL(k_stream1, grid_for(d.ncell), dim3(256), 0, s->st, d, s->a);
    keep(s);
    PH_CHK(hipGetLastError());
    std::swap(s->a.f, s->a.ft);
    return IMEXLB_OK;
}

int phased_update(PhasedState *s, std::string &err)
{
    hipLaunchKernelGGL(k_update, grid_for(s->d.ncell), dim3(256), 0, s->st, s->d, s->a);
    PH_CHK(hipGetLastError());
    return IMEXLB_OK;
}

int phased_full_step(PhasedState *s, std::string &err)
{
    int rc;
    if ((rc = phased_collide(s, err)) || (rc = phased_pack(s, err)) || (rc = phased_exchange(s, err)) ||
        (rc = phased_unpack(s, err)) || (rc = phased_stream(s, err)) || (rc = phased_update(s, err)))
        return rc;
    return IMEXLB_OK;
}

// n steps of Collision -> pack/exchange/unpack -> Streaming -> Update with
// the same results as n calls of phased_full_step: the first Collision on its
// own, then per step the halo and boundary kernels on the post-collision f,
// one k_stream_update_collide pass (with the next step's Collision, except
// after the last step), which also carries over what keep() would.
int phased_steps(PhasedState *s, int n, std::string &err)
{
    if (n <= 0) return IMEXLB_OK;
    int rc = phased_collide(s, err);
    if (rc) return rc;
    const Dims &d = s->d;
    for (int step = 1; step <= n; ++step) {
        if ((rc = phased_unpack(s, err))) return rc;
        boundaries(s);
        hipLaunchKernelGGL(k_stream_update_collide, grid_for(d.ncell), dim3(256), 0, s->st, d, s->a, s->a.f,
                           s->a.ft, step == n ? 1 : 0);
        PH_CHK(hipGetLastError());
        std::swap(s->a.f, s->a.ft);
    }
    return IMEXLB_OK;
}

static int interior_copy(PhasedState *s, double *dev_src, double *host, int nq, bool get, std::string &err)
{
    const Dims &d = s->d;
    const uint64_t n = (uint64_t)d.nx * d.ny * d.nz * nq;
    double *buf = nullptr;
    PH_CHK(hipMalloc(&buf, n * 8));
    hipError_t e = hipSuccess;
    if (get) {
        hipLaunchKernelGGL(k_interior_io, grid_for(n), dim3(256), 0, s->st, d, dev_src, buf, nq, 1);
        e = hipMemcpyAsync(host, buf, n * 8, hipMemcpyDeviceToHost, s->st);
    } else {
        e = hipMemcpyAsync(buf, host, n * 8, hipMemcpyHostToDevice, s->st);
        if (e == hipSuccess)
            hipLaunchKernelGGL(k_interior_io, grid_for(n), dim3(256), 0, s->st, d, dev_src, buf, nq, 0);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s->st);
    hipFree(buf);
    if (e != hipSuccess) {
        err = std::string("phased copy: ") + hipGetErrorString(e);
        return IMEXLB_EHIP;
    }
    return IMEXLB_OK;
}

int phased_get_moments(PhasedState *s, double *p, double *u, double *v, double *w, std::string &err)
{
    double *src[4] = {s->a.p, s->a.ua, s->a.va, s->a.wa};
    double *dst[4] = {p, u, v, w};
    for (int k = 0; k < 4; ++k) {
        if (!dst[k]) continue;
        // one-population view of a cell field: same copy kernel, nq = 1
        int rc = interior_copy(s, src[k], dst[k], 1, true, err);
        if (rc) return rc;
    }
    return IMEXLB_OK;
}

int phased_get_f(PhasedState *s, double *f, std::string &err) { return interior_copy(s, s->a.f, f, Q, true, err); }

int phased_set_f(PhasedState *s, const double *f, std::string &err)
{
    return interior_copy(s, s->a.f, const_cast<double *>(f), Q, false, err);
}

// The whole state between steps, ghost cells included (the faithful wall
// rules read ghost values written in earlier steps): f, then p, ua, va, wa,
// each over the ghosted box.
int64_t phased_state_doubles(const PhasedState *s) { return (int64_t)(Q + 4) * (int64_t)s->d.ncell; }

int phased_get_state(PhasedState *s, double *host, std::string &err)
{
    const size_t n = s->d.ncell;
    PH_CHK(hipMemcpyAsync(host, s->a.f, Q * n * 8, hipMemcpyDeviceToHost, s->st));
    double *mom[4] = {s->a.p, s->a.ua, s->a.va, s->a.wa};
    for (int k = 0; k < 4; ++k)
        PH_CHK(hipMemcpyAsync(host + (Q + k) * n, mom[k], n * 8, hipMemcpyDeviceToHost, s->st));
    PH_CHK(hipStreamSynchronize(s->st));
    return IMEXLB_OK;
}

// f and ft both take the saved f: ghost cells deeper than the carried-over
// layer must agree in the two arrays, as after Initialize.
int phased_set_state(PhasedState *s, const double *host, std::string &err)
{
    const size_t n = s->d.ncell;
    PH_CHK(hipMemcpyAsync(s->a.f, host, Q * n * 8, hipMemcpyHostToDevice, s->st));
    PH_CHK(hipMemcpyAsync(s->a.ft, s->a.f, Q * n * 8, hipMemcpyDeviceToDevice, s->st));
    double *mom[4] = {s->a.p, s->a.ua, s->a.va, s->a.wa};
    for (int k = 0; k < 4; ++k)
        PH_CHK(hipMemcpyAsync(mom[k], host + (Q + k) * n, n * 8, hipMemcpyHostToDevice, s->st));
    PH_CHK(hipStreamSynchronize(s->st));
    return IMEXLB_OK;
}

int phased_diagnostics(PhasedState *s, double *ke, double *mass, std::string &err)
{
    hipLaunchKernelGGL(k_diag1, dim3(s->nparts), dim3(256), 0, s->st, s->d, s->a, s->partials);
    hipLaunchKernelGGL(k_diag2, dim3(1), dim3(256), 0, s->st, s->partials, s->nparts, s->result);
    double out[2];
    PH_CHK(hipMemcpyAsync(out, s->result, sizeof out, hipMemcpyDeviceToHost, s->st));
    PH_CHK(hipStreamSynchronize(s->st));
    *ke = out[0];
    *mass = out[1];
    return IMEXLB_OK;
}

}  // namespace imx
