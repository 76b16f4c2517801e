// The fused D3Q27 BGK step: stream (pull or AA), moments, collide, in one
// pass over HBM.  Included by fused_f64.hip (compiled with -ffp-contract=off,
// bit-for-bit the oracle's FP64 arithmetic) and fused_lp.hip (FP32/FP16
// storage, FP32 collision).
//
// What one launch computes, per fluid cell x, equals one iteration of the
// reference's driver loop (tgv/main.cpp:70-119): the post-collision
// populations of the previous step are streamed (Streaming, tgv/src/lbm.cpp:
// 239-256), their moments taken (Update, tgv/src/lbm.cpp:288-333) and relaxed
// towards equilibrium (Collision, tgv/src/lbm.cpp:222-237) for the next step.
// The one-rank halo (tgv/src/mpi_view_transfer.cpp:117-383) is the periodic
// index wrap below.
//
// Layout: structure of arrays, population i at lattice + i * pstride, cells in
// x-fastest order, 64 consecutive x cells per wavefront.
#pragma once

#include <hip/hip_runtime.h>
#include <type_traits>

#include "kernels.h"

namespace imx {

// Device memory through the global address space.  The single-domain
// kernels get their pointers as kernel arguments, where the compiler infers
// this by itself; the ensemble kernel reads them from its domain table, and
// without the cast it would emit flat accesses, which also count against the
// scalar-load counter and so are waited for at every parameter load.
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T *glb(T *p)
{
    return (__attribute__((address_space(1))) T *)p;
}

template <typename S> struct RealOf { using type = float; };
template <> struct RealOf<double> { using type = double; };

// Load the stored value at p, as a population of weight class `cls` (FP16
// storage is scaled per class).
template <typename S, bool NT = false>
__device__ __forceinline__ typename RealOf<S>::type ld_ptr(const S *ptr, int cls)
{
    const auto *p = glb(ptr);
    S x;
    if constexpr (NT) x = __builtin_nontemporal_load(p);
    else x = *p;
    if constexpr (std::is_same<S, _Float16>::value)
        return (float)x * H_INV_SCALE[cls];
    else
        return x;
}

// The value at element `at` of the lattice.
template <typename S, bool NT = false>
__device__ __forceinline__ typename RealOf<S>::type ld_at(const S *lat, uint64_t at, int cls)
{
    return ld_ptr<S, NT>(lat + at, cls);
}

// How the kernels address a population (the launch picks the most compact
// mode the population stride allows, capped by the handle's configuration;
// every mode gives the same results):
// - AM_WIDE: 64-bit element offsets.
// - AM_W32: population i's array as a uniform base plus the cell's byte
//   offset as a 32-bit value.  base + zext(offset) is the global saddr form,
//   one VGPR per address instead of a 64-bit pair, which keeps the AA odd step
//   (27 neighbour addresses loaded and stored) at the register count of the
//   other steps.  Needs one population within 4 GiB.
// - AM_GROUP: as AM_W32, and where every population is addressed at the same
//   cell (the AA even step, the pull step's stores) the populations are taken
//   in groups of PopGroup<S>::n (16 bytes of population stride per cell): one
//   uniform base per group, and the population's place in the group added to
//   the 32-bit offset.  With a base per population there the compiler forms
//   lattice + offset once in VGPRs and adds each population's 64-bit offset
//   in VALU; with the groups the step keeps a few scalar bases (FP16 even
//   step: 72 instead of 103 VGPRs).  Needs a group within 4 GiB: at most 2^28
//   cells per population.
enum AddrMode : int { AM_WIDE = 0, AM_W32 = 1, AM_GROUP = 2 };

template <typename S>
struct PopGroup { static constexpr int n = 16 / (int)sizeof(S); };

template <typename S, int AM, bool SAME_CELL = false>
__device__ __forceinline__ S *pop_ptr(S *lat, uint64_t pstride, int i, uint32_t c)
{
    using B = typename std::conditional<std::is_const<S>::value, const char, char>::type;
    if constexpr (AM == AM_GROUP && SAME_CELL) {
        constexpr int G = PopGroup<S>::n;
        // the group's base through readfirstlane: a uniform value the
        // compiler cannot fold back into the per-lane address arithmetic
        const uint64_t a = (uint64_t)(lat + (uint64_t)(i / G * G) * pstride);
        const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)a);
        const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(a >> 32));
        const uint32_t o = (uint32_t)(i % G) * (uint32_t)(pstride * sizeof(S)) + c * (uint32_t)sizeof(S);
        return (S *)((B *)(((uint64_t)hi << 32) | lo) + o);
    } else if constexpr (AM != AM_WIDE) {
        return (S *)((B *)(lat + (uint64_t)i * pstride) + (uint32_t)(c * (uint32_t)sizeof(S)));
    } else {
        return lat + (uint64_t)i * pstride + c;
    }
}

template <typename S, bool NT = false, int AM = AM_WIDE>
__device__ __forceinline__ typename RealOf<S>::type ld_pop(const S *lat, uint64_t pstride, int i, uint32_t c)
{
    return ld_ptr<S, NT>(pop_ptr<const S, AM>(lat, pstride, i, c), WCLASS[i]);
}

// NTL (non-temporal loads) variant of the gather: lines that no other
// wavefront reads -- every line of the AA even (in-place) step, the x-aligned
// lines of the pull and AA odd steps -- are loaded without being kept in the
// caches.  Chosen per device by the autotuner; results are identical.

// Store policy: 0 non-temporal everywhere, 1 plain everywhere, 2 (default)
// non-temporal for aligned writes and plain for the AA odd step's x-shifted
// writes, whose partial lines at segment ends then merge in L2 before they
// reach HBM (policy 0 measured slower on the FP64 AA 512^3 step; the odd
// step's write traffic is 1.043 x its algorithmic bytes with policy 2).
#ifndef IMX_STORE_POLICY
#define IMX_STORE_POLICY 2
#endif
template <typename S, bool NT = true>
__device__ __forceinline__ void st_ptr(S *ptr, int cls, typename RealOf<S>::type v)
{
    auto *p = glb(ptr);
    S x;
    if constexpr (std::is_same<S, _Float16>::value) x = (_Float16)(v * H_SCALE[cls]);
    else x = v;
    if constexpr (NT && IMX_STORE_POLICY != 1) __builtin_nontemporal_store(x, p);
    else *p = x;
}

template <typename S, bool NT = true, int AM = AM_WIDE, bool SAME_CELL = false>
__device__ __forceinline__ void st_pop(S *lat, uint64_t pstride, int i, uint32_t c, typename RealOf<S>::type v)
{
    st_ptr<S, NT>(pop_ptr<S, AM, SAME_CELL>(lat, pstride, i, c), WCLASS[i], v);
}

// XCD-aware tile order: the hardware deals workgroups to the 8 XCDs round
// robin; this gives XCD r one contiguous run of tiles so that the partially
// used cache lines at tile edges stay in one L2.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t nb)
{
    const uint32_t xcd = b & 7u, idx = b >> 3, per = nb >> 3, rem = nb & 7u;
    return xcd * per + (xcd < rem ? xcd : rem) + idx;
}

// e_i . u with the zero components left out, summed x, y, z in that order --
// the value of ((e_x u + e_y v) + e_z w) for e in {-1, 0, 1}^3.
template <typename R>
__device__ __forceinline__ R edu_of(int i, R u, R v, R w)
{
    R acc = R(0);
    bool has = false;
    if (EX[i] != 0) {
        acc = EX[i] > 0 ? u : -u;
        has = true;
    }
    if (EY[i] != 0) {
        R t = EY[i] > 0 ? v : -v;
        acc = has ? acc + t : t;
        has = true;
    }
    if (EZ[i] != 0) {
        R t = EZ[i] > 0 ? w : -w;
        acc = has ? acc + t : t;
    }
    return acc;
}

// Moments in direction order 0..26 (lid_driven_cavity/src/lbm.cpp:329-347):
// p = sum f/3, u = sum f e.  FP64 storage: exactly the oracle's sums.
__device__ __forceinline__ void moments_f64(const double (&f)[Q], double &pm, double &um, double &vm, double &wm)
{
#pragma clang fp contract(off)
    pm = 0.0;
    um = 0.0;
    vm = 0.0;
    wm = 0.0;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        pm = pm + f[i] / 3.0;
        if (EX[i] > 0) um = um + f[i];
        if (EX[i] < 0) um = um - f[i];
        if (EY[i] > 0) vm = vm + f[i];
        if (EY[i] < 0) vm = vm - f[i];
        if (EZ[i] > 0) wm = wm + f[i];
        if (EZ[i] < 0) wm = wm - f[i];
    }
}

// BGK relaxation in the reference's operation order (tgv/src/lbm.cpp:228-234):
// feq = t p 3 + t (3 e.u + 4.5 (e.u)^2 - 1.5 u.u); f -= (f - feq) / (tau0 + 0.5).
__device__ __forceinline__ void collide_f64(double (&f)[Q], double pm, double u, double v, double w, double tau0)
{
    // every product and sum rounds separately, whatever the translation
    // unit's contraction setting: the FP64 path is the oracle's arithmetic
#pragma clang fp contract(off)
    const double den = tau0 + 0.5;
    const double udu = u * u + v * v + w * w;
    const double u15 = 1.5 * udu;
    double tp3[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) tp3[k] = WVAL[k] * pm * 3.0;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        const double edu = edu_of<double>(i, u, v, w);
        const double eu2 = edu * edu;
        const double feq = tp3[WCLASS[i]] + WVAL[WCLASS[i]] * (3.0 * edu + 4.5 * eu2 - u15);
        f[i] = f[i] - (f[i] - feq) / den;
    }
}

// Reduced-precision storage holds g = f - w S.  The incompressible BGK update
// is linear in p, so g obeys the same update with p' = p - S/3; the density
// sum is kept in FP64, the rest in FP32.
//
// Both functions work on the 13 bounce-back pairs (i, bb(i)) = (2k+1, 2k+2),
// whose velocities are opposite and whose weights are equal: the moments take
// the pair sums and differences, and the equilibria of a pair share
// 3 p' - 1.5 u.u + 4.5 (e.u)^2 and differ only in the sign of 3 e.u.  That
// saves about a quarter of the vector operations per cell, which matters
// where 108 bytes per cell leave little time for arithmetic (FP16).
// Contraction is off and the fused multiply-adds are explicit, so every
// kernel variant that inlines these rounds the same way and the launch shape
// never changes a bit.
constexpr int NPAIR = 13;
constexpr bool pairs_adjacent()
{
    for (int k = 0; k < NPAIR; ++k) {
        const int i = 2 * k + 1;
        if (BB[i] != i + 1 || EX[i + 1] != -EX[i] || EY[i + 1] != -EY[i] || EZ[i + 1] != -EZ[i] ||
            WCLASS[i + 1] != WCLASS[i])
            return false;
    }
    return true;
}
static_assert(pairs_adjacent(), "the pair form needs bb(2k+1) = 2k+2 with opposite velocity and equal weight");

__device__ __forceinline__ void moments_lp(const float (&g)[Q], double &pd, float &u, float &v, float &w)
{
#pragma clang fp contract(off)
    pd = (double)g[0];
    u = 0.f;
    v = 0.f;
    w = 0.f;
#pragma unroll
    for (int k = 0; k < NPAIR; ++k) {
        const int i = 2 * k + 1;  // bb(i) = i + 1, e_bb(i) = -e_i
        const float sum = g[i] + g[i + 1];
        const float dif = g[i] - g[i + 1];
        pd += (double)sum;
        if (EX[i] > 0) u += dif;
        if (EX[i] < 0) u -= dif;
        if (EY[i] > 0) v += dif;
        if (EY[i] < 0) v -= dif;
        if (EZ[i] > 0) w += dif;
        if (EZ[i] < 0) w -= dif;
    }
}

// g <- g + omega (feq - g), feq = w (3 p' + 3 e.u + 4.5 (e.u)^2 - 1.5 u.u).
// The non-equilibrium part feq - g is formed first, as in the reference's
// update (tgv/src/lbm.cpp:228-234); (1 - omega) g + omega feq would round
// at the size of g and lose most of it at omega near 2.
__device__ __forceinline__ void collide_lp(float (&g)[Q], double pd, float u, float v, float w, float omega)
{
#pragma clang fp contract(off)
    const float a = (float)pd - 1.5f * (u * u + v * v + w * w);  // 3 p' - 1.5 u.u
    g[0] = __builtin_fmaf(omega, (float)WVAL[0] * a - g[0], g[0]);
#pragma unroll
    for (int k = 0; k < NPAIR; ++k) {
        const int i = 2 * k + 1;
        const float edu = edu_of<float>(i, u, v, w);
        const float even = __builtin_fmaf(4.5f * edu, edu, a);  // a + 4.5 (e.u)^2
        const float odd = 3.f * edu;
        const float wi = (float)WVAL[WCLASS[i]];
        g[i] = __builtin_fmaf(omega, wi * (even + odd) - g[i], g[i]);
        g[i + 1] = __builtin_fmaf(omega, wi * (even - odd) - g[i + 1], g[i + 1]);
    }
}

// collide_lp for populations stored relative to the equilibrium of a free
// stream (U, 0, 0) at p = S/3 (the open channel's inflow):
// g = f - feq(S/3, U), and (pd, du, dv, dw) = (sum g, sum g e) are the
// departures 3 p - S and u - U.  feq(p, u) - feq(S/3, U) =
// w (pd - 1.5 (u.u - U^2) + 4.5 ((e.u)^2 - (e_x U)^2) + 3 e.(u - U)), written
// in the departures so that FP32 resolves them whatever U is:
// u.u - U^2 = du (du + 2U) + dv^2 + dw^2, (e.u)^2 - (e_x U)^2 =
// (e.du) (e.du + 2 e_x U).  With U = 0 it is collide_lp.
__device__ __forceinline__ void collide_lp_stream(float (&g)[Q], double pd, float du, float dv, float dw,
                                                  float omega, float U)
{
#pragma clang fp contract(off)
    const float uu = __builtin_fmaf(du, du + 2.f * U, __builtin_fmaf(dv, dv, dw * dw));
    const float a = (float)pd - 1.5f * uu;
    g[0] = __builtin_fmaf(omega, (float)WVAL[0] * a - g[0], g[0]);
    const float U2 = 2.f * U;
#pragma unroll
    for (int k = 0; k < NPAIR; ++k) {
        const int i = 2 * k + 1;
        const float edu = edu_of<float>(i, du, dv, dw);
        float ex = edu;
        if (EX[i] > 0) ex = edu + U2;
        if (EX[i] < 0) ex = edu - U2;
        const float even = __builtin_fmaf(4.5f * edu, ex, a);
        const float odd = 3.f * edu;
        const float wi = (float)WVAL[WCLASS[i]];
        g[i] = __builtin_fmaf(omega, wi * (even + odd) - g[i], g[i]);
        g[i + 1] = __builtin_fmaf(omega, wi * (even - odd) - g[i + 1], g[i + 1]);
    }
}

// Where a cell's populations come from and go to.
struct CellCtx {
    uint32_t c;
    uint32_t xv[3];       // x - 1, x, x + 1 (wrapped or clamped per boundary)
    uint32_t rowb[3][3];  // row offsets of (z + a - 1, y + b - 1)
    bool x0, x1, y0, y1, z0, z1;
    uint8_t m;            // channel mask of the cell
    bool live;
    uint32_t rx0, ry0, rz0;  // BC_FAITHFUL: ring offsets of (y, z), (x, z), (x, y) on the x, y, z faces
    uint32_t z, y;           // BC_FAITHFUL: the cell's z and y
};

template <int BC>
__device__ __forceinline__ void cell_setup(const FusedParams &P, uint32_t c, CellCtx &k)
{
    const uint32_t nx = P.nx, ny = P.ny, nz = P.nz;
    k.c = c;
    k.live = c < P.ncell;
    if (!k.live) c = 0;  // keep the arithmetic in range; nothing is loaded
    const uint32_t t = P.divx.div(c);
    const uint32_t x = c - t * nx;
    const uint32_t z = P.divy.div(t);
    const uint32_t y = t - z * ny;
    k.m = 0;
    if constexpr (BC == BC_CHANNEL) {
        if (k.live) {
            k.m = *glb(P.mask + c);
            if (k.m == 1) k.live = false;  // solid: nothing streams from or to it
        }
    }
    // periodic axes wrap, walls clamp (the clamped value is never used as a
    // source), the channel outlet clamps (zero-gradient ghost)
    uint32_t yv[3], zv[3];
    if constexpr (BC == BC_PERIODIC) {
        k.xv[0] = x == 0 ? nx - 1 : x - 1;
        k.xv[2] = x == nx - 1 ? 0 : x + 1;
    } else {
        k.xv[0] = x == 0 ? 0 : x - 1;
        k.xv[2] = x == nx - 1 ? nx - 1 : x + 1;
    }
    if constexpr (BC == BC_CAVITY || BC == BC_FAITHFUL) {
        yv[0] = y == 0 ? 0 : y - 1;
        yv[2] = y == ny - 1 ? ny - 1 : y + 1;
        zv[0] = z == 0 ? 0 : z - 1;
        zv[2] = z == nz - 1 ? nz - 1 : z + 1;
    } else {
        yv[0] = y == 0 ? ny - 1 : y - 1;
        yv[2] = y == ny - 1 ? 0 : y + 1;
        zv[0] = z == 0 ? nz - 1 : z - 1;
        zv[2] = z == nz - 1 ? 0 : z + 1;
    }
    k.xv[1] = x;
    yv[1] = y;
    zv[1] = z;
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) k.rowb[a][b] = (zv[a] * ny + yv[b]) * nx;
    k.x0 = x == 0;
    k.x1 = x == nx - 1;
    k.y0 = y == 0;
    k.y1 = y == ny - 1;
    k.z0 = z == 0;
    k.z1 = z == nz - 1;
    if constexpr (BC == BC_FAITHFUL) {
        k.z = z;
        k.y = y;
        k.rx0 = (z + 1) * (ny + 2) + y + 1;
        k.ry0 = (z + 1) * nx + x;
        k.rz0 = y * nx + x;
    }
}

// is the cell x - e_i outside the cavity?
__device__ __forceinline__ bool src_out(const CellCtx &k, int i)
{
    return (EX[i] > 0 && k.x0) || (EX[i] < 0 && k.x1) || (EY[i] > 0 && k.y0) || (EY[i] < 0 && k.y1) ||
           (EZ[i] > 0 && k.z0) || (EZ[i] < 0 && k.z1);
}
// offset of the cell x - e_i
__device__ __forceinline__ uint32_t src_off(const CellCtx &k, int i)
{
    return k.rowb[1 - EZ[i]][1 - EY[i]] + k.xv[1 - EX[i]];
}

// AA storage reads and writes the same 27 addresses per cell: direction i is
// stored where direction bb(i) was loaded from (even step: slot bb(i) of the
// cell; odd step: slot i of x + e_i, which the load of bb(i) read from x -
// e_bb(i)).  The gather hands its addresses to the scatter through `at`, so
// the odd step keeps one set of addresses -- a scalar base and a 32-bit
// vector offset each with AM_W32 -- instead of forming a second set for the
// stores.
template <typename S, int BC, int PAT, bool NTL = false, int AM = AM_WIDE>
__device__ __forceinline__ void cell_gather(const FusedParams &P, const CellCtx &k,
                                            typename RealOf<S>::type (&f)[Q], const S *(&at)[Q])
{
    using R = typename RealOf<S>::type;
    constexpr bool F64 = std::is_same<S, double>::value;
    const S *src = (const S *)P.src;
    const uint64_t ps = P.pstride;
    const uint32_t c = k.c;
    if constexpr (PAT == PAT_AA_EVEN) {
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            at[i] = pop_ptr<const S, AM, true>(src, ps, i, c);
            f[i] = ld_ptr<S, NTL>(at[i], WCLASS[i]);
            if constexpr (BC == BC_CAVITY) {
                // the odd step stored the bounced population raw; the lid
                // term is added on arrival, as in the pull form (a select:
                // see the pull gather below)
                if (EZ[i] < 0) {
                    R lid;
                    if constexpr (F64) lid = f[i] + P.corr[i];
                    else lid = f[i] + (float)P.corr[i];
                    f[i] = k.z1 ? lid : f[i];
                }
            }
        }
    } else {
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            if constexpr (BC == BC_PERIODIC) {
                at[i] = pop_ptr<const S, AM>(src, ps, PAT == PAT_PULL ? i : BB[i], src_off(k, i));
                if (NTL && EX[i] == 0) f[i] = ld_ptr<S, true>(at[i], WCLASS[i]);
                else f[i] = ld_ptr<S, false>(at[i], WCLASS[i]);
            } else if constexpr (BC == BC_CAVITY) {
                // Boundary choices are selects, not branches: a value loaded
                // inside a divergent branch, or used by one, is waited for
                // there, which would serialise the 27 loads of the wavefront.
                // Halfway bounce-back where x - e_i is outside: f_i(x) = f*_bb(i)(x),
                // plus the lid term on the moving wall.
                const bool out = src_out(k, i);
                const int s_in = PAT == PAT_PULL ? i : BB[i], s_out = PAT == PAT_PULL ? BB[i] : i;
                at[i] = src + (out ? (uint64_t)s_out * ps + c : (uint64_t)s_in * ps + src_off(k, i));
                R v;
                if (NTL && EX[i] == 0) v = ld_ptr<S, true>(at[i], WCLASS[i]);
                else v = ld_ptr<S>(at[i], WCLASS[i]);
                if (EZ[i] < 0) {  // x - e_i outside whenever z is the top layer
                    R lid;
                    if constexpr (F64) lid = v + P.corr[i];
                    else lid = v + (float)P.corr[i];
                    v = k.z1 ? lid : v;
                }
                f[i] = v;
            } else if constexpr (BC == BC_FAITHFUL) {  // pull only
                // a source outside the cavity is a ghost cell of the first
                // layer, which the reference's wall kernels filled this step:
                // read it from the ring (selects, as above)
                const bool ox = (EX[i] > 0 && k.x0) || (EX[i] < 0 && k.x1);
                const bool oy = (EY[i] > 0 && k.y0) || (EY[i] < 0 && k.y1);
                const bool oz = (EZ[i] > 0 && k.z0) || (EZ[i] < 0 && k.z1);
                const uint32_t nx = P.nx, ny = P.ny;
                const uint32_t rxi = (EX[i] < 0 ? P.ring_a : 0u) + k.rx0 - EZ[i] * (ny + 2) - EY[i];
                const uint32_t ryi = 2 * P.ring_a + (EY[i] < 0 ? P.ring_b : 0u) + k.ry0 - EZ[i] * nx - EX[i];
                const uint32_t rzi =
                    2 * (P.ring_a + P.ring_b) + (EZ[i] < 0 ? P.ring_c : 0u) + k.rz0 - EY[i] * nx - EX[i];
                const uint32_t r = ox ? rxi : (oy ? ryi : rzi);
                // one base, two offsets: the ring's offset from the lattice
                // may be negative, so the address is formed modulo 2^64
                const uint64_t at = (ox || oy || oz) ? (uint64_t)P.ring_rel + (uint64_t)i * P.ring_n + r
                                                     : (uint64_t)i * ps + src_off(k, i);
                const S *p = (const S *)((uintptr_t)src + at * sizeof(S));
                if (NTL && EX[i] == 0) f[i] = __builtin_nontemporal_load(glb(p));
                else f[i] = *glb(p);
            } else {  // channel, pull only
                // selects as above; only the mask byte of a neighbour of a
                // solid cell is loaded under a branch (its value picks the address)
                const uint32_t so = src_off(k, i);
                bool solid = false;
                if (i != 0 && k.m == 2) solid = *glb(P.mask + so) == 1;
                const uint64_t at = solid ? (uint64_t)BB[i] * ps + c : (uint64_t)i * ps + so;
                R v = ld_at<S>(src, at, WCLASS[i]);
                if constexpr (!F64) {
                    // shifted storage: a link bounced off the ball arrives as
                    // g_i = g*_bb(i) + s_bb(i) - s_i = g*_bb(i) - 6 w_i e_x U (the
                    // inflow equilibrium's stream U runs along x: nothing to add
                    // where e_x = 0)
                    if (EX[i] != 0) {
                        const R b = v - (float)(WVAL[WCLASS[i]] * EX[i]) * (6.f * (float)P.ush);
                        v = solid ? b : v;
                    }
                }
                if (EX[i] > 0) {
                    R in;
                    if constexpr (F64) in = P.feq_in[i];
                    else in = (float)P.feq_in[i];
                    v = k.x0 ? in : v;
                }
                f[i] = v;
            }
        }
    }
}

template <typename S, int BC>
__device__ __forceinline__ void cell_collide(const FusedParams &P, const CellCtx &k,
                                             typename RealOf<S>::type (&f)[Q], bool write_moments)
{
    constexpr bool F64 = std::is_same<S, double>::value;
    const uint32_t c = k.c;
    if constexpr (F64) {
        double pm, um, vm, wm;
        moments_f64(f, pm, um, vm, wm);
        if constexpr (BC == BC_FAITHFUL) {
            if (P.faithful_sphere) {
                // sphere/src/lbm.cpp:399-412: u, p *= usr; the inlet plane
                // takes the parabolic profile in z, the outlet plane p = 0
                const double m = P.mask[c] == 1 ? 0.0 : 1.0;
                um = um * m;
                vm = vm * m;
                wm = wm * m;
                pm = pm * m;
                const int glz = P.nz, kz = (int)k.z;
                const double gm = (double)(glz - 1);
                const double uin = P.u0 * 4.0 * (double)kz * (double)(glz - 1 - kz) / (gm * gm);
                um = k.x0 ? uin : um;
                vm = k.x0 ? 0.0 : vm;
                wm = k.x0 ? 0.0 : wm;
                pm = k.x1 ? 0.0 : pm;
            } else {
                // the reference's lid: Update sets ua = 0.1 on the top fluid
                // layer (lid_driven_cavity/src/lbm.cpp:343-344)
                um = k.z1 ? 0.1 : um;
            }
        }
        if (write_moments) {
            *glb(P.mom_p + c) = pm;
            *glb(P.mom_u + c) = um;
            *glb(P.mom_v + c) = vm;
            *glb(P.mom_w + c) = wm;
        }
        collide_f64(f, pm, um, vm, wm, P.tau0);
    } else {
        double pd;
        float u, v, w;
        moments_lp(f, pd, u, v, w);
        if constexpr (BC == BC_CHANNEL) {
            // populations stored relative to the inflow equilibrium: the
            // moments above are the departure from the free stream (U, 0, 0)
            if (write_moments) {
                *glb(P.mom_p + c) = (pd + P.shift) / 3.0;
                *glb(P.mom_u + c) = (double)u + P.ush;
                *glb(P.mom_v + c) = v;
                *glb(P.mom_w + c) = w;
            }
            collide_lp_stream(f, pd, u, v, w, P.omega_f, (float)P.ush);
        } else {
            if (write_moments) {
                *glb(P.mom_p + c) = (pd + P.shift) / 3.0;
                *glb(P.mom_u + c) = u;
                *glb(P.mom_v + c) = v;
                *glb(P.mom_w + c) = w;
            }
            collide_lp(f, pd, u, v, w, P.omega_f);
        }
    }
}

// REUSE: the AA odd step stores through the gather's addresses (see
// cell_gather); otherwise the store addresses are formed again from the
// cell, which is cheaper in registers when the addresses of two cells per
// thread, or the even step's in-place addresses, would be held.
template <typename S, int BC, int PAT, int AM = AM_WIDE, bool REUSE = false>
__device__ __forceinline__ void cell_scatter(const FusedParams &P, const CellCtx &k,
                                             const typename RealOf<S>::type (&f)[Q], const S *const (&at)[Q])
{
    S *dst = (S *)P.dst;
    const uint64_t ps = P.pstride;
    const uint32_t c = k.c;
    if constexpr (PAT == PAT_PULL) {
        if constexpr (BC == BC_FAITHFUL) {
            // what the next step's x-face boundary kernels read, each row's
            // nine together (FusedParams::stage): the cavity's walls x = 1
            // (e_x = -1) and x = nx - 2 (e_x = +1), the sphere's inlet x = 0
            // (e_x = +1) and outlet x = nx - 1 (e_x = -1)
            // The register array f is indexed with compile-time constants
            // only (XDIR_P / XDIR_M), the two cases chosen by selects: an index
            // computed at run time would move f to private memory (64 more
            // bytes written per cell).
            if (P.stage != nullptr) {
                const uint64_t plane = (uint64_t)P.ny * P.nz, yz = (uint64_t)k.z * P.ny + k.y;
                const uint32_t x = k.xv[1];
                const bool sph = P.faithful_sphere != 0;
                if (x == (sph ? 0u : 1u)) {
                    auto *p0 = glb(P.stage + yz * STAGE_ROW);
#pragma unroll
                    for (int r = 0; r < 9; ++r) p0[r] = sph ? f[XDIR_P[r]] : f[XDIR_M[r]];
                }
                if (x == (sph ? P.nx - 1u : P.nx - 2u)) {
                    auto *p1 = glb(P.stage + (plane + yz) * STAGE_ROW);
#pragma unroll
                    for (int r = 0; r < 9; ++r) p1[r] = sph ? f[XDIR_M[r]] : f[XDIR_P[r]];
                }
            }
        }
#pragma unroll
        for (int i = 0; i < Q; ++i) st_pop<S, true, AM, true>(dst, ps, i, c, f[i]);
    } else if constexpr (PAT == PAT_AA_EVEN) {
#pragma unroll
        for (int i = 0; i < Q; ++i) st_pop<S, true, AM, true>(dst, ps, BB[i], c, f[i]);
    } else {
        // to the natural slot of x + e_i, or bounced back into slot bb(i) at
        // x: in both cases where direction bb(i) was loaded from.  Shifted
        // writes (e_x != 0, not bounced) cover partial lines at row segment
        // ends and stay plain stores, which merge in L2.
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            bool bounced = false;
            if constexpr (BC == BC_CAVITY) bounced = src_out(k, BB[i]);
            const bool plain = IMX_STORE_POLICY == 2 && EX[i] != 0 && !bounced;
            S *p;
            if constexpr (REUSE) p = (S *)at[BB[i]];
            else if constexpr (BC == BC_CAVITY)
                p = bounced ? pop_ptr<S, AM>(dst, ps, BB[i], c) : pop_ptr<S, AM>(dst, ps, i, src_off(k, BB[i]));
            else p = pop_ptr<S, AM>(dst, ps, i, src_off(k, BB[i]));
            if (plain) st_ptr<S, false>(p, WCLASS[i], f[i]);
            else st_ptr<S>(p, WCLASS[i], f[i]);
        }
    }
}

// CPT cells per thread, blockDim.x apart; the loads of all CPT cells are
// issued before any of them is collided, to keep more bytes in flight.
// AM: the address mode (AddrMode).
template <typename S, int BC, int PAT, int CPT, bool NTL, int AM>
__global__ void __launch_bounds__(512) k_fused(const FusedParams P)
{
    using R = typename RealOf<S>::type;
    uint32_t b = blockIdx.x;
    if (P.xcd_remap) b = xcd_tile(b, P.nblocks);
    const uint32_t base = b * blockDim.x * CPT + threadIdx.x;
    CellCtx k[CPT];
    R f[CPT][Q];
    const S *at[CPT][Q];
#pragma unroll
    for (int r = 0; r < CPT; ++r) {
        cell_setup<BC>(P, base + (uint32_t)r * blockDim.x, k[r]);
        if (k[r].live) cell_gather<S, BC, PAT, NTL, AM>(P, k[r], f[r], at[r]);
    }
#pragma unroll
    for (int r = 0; r < CPT; ++r) {
        if (!k[r].live) continue;
        cell_collide<S, BC>(P, k[r], f[r], P.mom_p != nullptr);
        cell_scatter<S, BC, PAT, AM, PAT == PAT_AA_ODD && CPT == 1>(P, k[r], f[r], at[r]);
    }
}

// Two adjacent cells per thread: cells c and c + 1 with c = 2 t, x even, on
// a periodic lattice with nx even (the pair never leaves its row).  Every
// population of the pair moves as one access of 2 sizeof(S) bytes: a
// wavefront covers 128 cells of a row, so a shifted row segment splits fewer
// cache lines at its ends than two segments of 64 cells.  Only the pair at a
// row's start or end, whose shifted neighbours wrap to (nx - 1, 0), needs the
// element x = 0 separately: it is loaded in a block that only wavefronts
// holding such a pair enter, and stored by those lanes alone.
template <typename S>
struct PairOf {
    // two elements at the element's own alignment (shifted pairs start at
    // odd x)
    typedef S type __attribute__((ext_vector_type(2), aligned(sizeof(S))));
};

template <typename S, bool NT>
__device__ __forceinline__ void ld_pair(const S *ptr, int cls, typename RealOf<S>::type &a,
                                        typename RealOf<S>::type &b)
{
    using V = typename PairOf<S>::type;
    const auto *p = (const __attribute__((address_space(1))) V *)ptr;
    V v;
    if constexpr (NT) v = __builtin_nontemporal_load(p);
    else v = *p;
    if constexpr (std::is_same<S, _Float16>::value) {
        a = (float)v.x * H_INV_SCALE[cls];
        b = (float)v.y * H_INV_SCALE[cls];
    } else {
        a = v.x;
        b = v.y;
    }
}

template <typename S, bool NT>
__device__ __forceinline__ void st_pair(S *ptr, int cls, typename RealOf<S>::type a, typename RealOf<S>::type b)
{
    using V = typename PairOf<S>::type;
    auto *p = (__attribute__((address_space(1))) V *)ptr;
    V v;
    if constexpr (std::is_same<S, _Float16>::value) {
        v.x = (_Float16)(a * H_SCALE[cls]);
        v.y = (_Float16)(b * H_SCALE[cls]);
    } else {
        v.x = a;
        v.y = b;
    }
    if constexpr (NT && IMX_STORE_POLICY != 1) __builtin_nontemporal_store(v, p);
    else *p = v;
}

__device__ __forceinline__ bool wave_any(bool v) { return __builtin_amdgcn_ballot_w64(v) != 0; }

template <typename S, int PAT, bool NTL>
__global__ void __launch_bounds__(512) k_fused_pair(const FusedParams P)
{
    using R = typename RealOf<S>::type;
    uint32_t blk = blockIdx.x;
    if (P.xcd_remap) blk = xcd_tile(blk, P.nblocks);
    const uint32_t c = 2u * (blk * blockDim.x + threadIdx.x);
    if (c >= P.ncell) return;  // ncell is even: both cells of a started pair exist
    const uint32_t nx = P.nx, ny = P.ny, nz = P.nz;
    const uint32_t t = P.divx.div(c);
    const uint32_t x = c - t * nx;
    const uint32_t z = P.divy.div(t);
    const uint32_t y = t - z * ny;
    const uint32_t yv[3] = {y == 0 ? ny - 1 : y - 1, y, y == ny - 1 ? 0 : y + 1};
    const uint32_t zv[3] = {z == 0 ? nz - 1 : z - 1, z, z == nz - 1 ? 0 : z + 1};
    uint32_t rowb[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) rowb[a][b] = (zv[a] * ny + yv[b]) * nx;
    // a pair at the row's start: its x - 1 neighbours are (nx - 1, 0); at the
    // row's end its x + 1 neighbours are (nx - 1, 0)
    const bool lo = x == 0, hi = x + 2 == nx;
    const S *src = (const S *)P.src;
    S *dst = (S *)P.dst;
    const uint64_t ps = P.pstride;

    R f0[Q], f1[Q];
    if constexpr (PAT == PAT_AA_EVEN) {
#pragma unroll
        for (int i = 0; i < Q; ++i) ld_pair<S, NTL>(pop_ptr<const S, AM_GROUP, true>(src, ps, i, c), WCLASS[i], f0[i], f1[i]);
    } else {
        // slot s of the pair (x, x + 1) - e_i (pull: s = i; AA odd: s = bb(i)).
        // The element x = 0 of split pairs is loaded first, raw, in blocks
        // that only wavefronts holding such a pair enter; the selects and
        // conversions come after the pairs' loads, so nothing waits inside
        // the blocks.
        using V = typename PairOf<S>::type;
        S fix[Q];
        if (wave_any(lo)) {
#pragma unroll
            for (int i = 0; i < Q; ++i)
                if (EX[i] > 0)
                    fix[i] = *glb(pop_ptr<const S, AM_GROUP, true>(src, ps, PAT == PAT_PULL ? i : BB[i],
                                                                   rowb[1 - EZ[i]][1 - EY[i]]));
        }
        if (wave_any(hi)) {
#pragma unroll
            for (int i = 0; i < Q; ++i)
                if (EX[i] < 0)
                    fix[i] = *glb(pop_ptr<const S, AM_GROUP, true>(src, ps, PAT == PAT_PULL ? i : BB[i],
                                                                   rowb[1 - EZ[i]][1 - EY[i]]));
        }
        V raw[Q];
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            const int s = PAT == PAT_PULL ? i : BB[i];
            const uint32_t row = rowb[1 - EZ[i]][1 - EY[i]];
            const bool split = EX[i] > 0 ? lo : (EX[i] < 0 ? hi : false);
            // a split pair loads (nx - 2, nx - 1) and keeps the second
            const uint32_t o = split ? row + nx - 2 : row + x - EX[i];
            const auto *p = (const __attribute__((address_space(1))) V *)pop_ptr<const S, AM_GROUP, true>(src, ps, s, o);
            if (NTL && EX[i] == 0) raw[i] = __builtin_nontemporal_load(p);
            else raw[i] = *p;
        }
        // every load is issued before the first value is used
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            S a = raw[i].x, b = raw[i].y;
            if (EX[i] != 0) {
                const bool split = EX[i] > 0 ? lo : hi;
                a = split ? raw[i].y : a;
                b = split ? fix[i] : b;
            }
            if constexpr (std::is_same<S, _Float16>::value) {
                f0[i] = (float)a * H_INV_SCALE[WCLASS[i]];
                f1[i] = (float)b * H_INV_SCALE[WCLASS[i]];
            } else {
                f0[i] = a;
                f1[i] = b;
            }
        }
    }

    CellCtx k;
    k.c = c;
    cell_collide<S, BC_PERIODIC>(P, k, f0, P.mom_p != nullptr);
    k.c = c + 1;
    cell_collide<S, BC_PERIODIC>(P, k, f1, P.mom_p != nullptr);

    if constexpr (PAT == PAT_PULL) {
#pragma unroll
        for (int i = 0; i < Q; ++i) st_pair<S, true>(pop_ptr<S, AM_GROUP, true>(dst, ps, i, c), WCLASS[i], f0[i], f1[i]);
    } else if constexpr (PAT == PAT_AA_EVEN) {
#pragma unroll
        for (int i = 0; i < Q; ++i) st_pair<S, true>(pop_ptr<S, AM_GROUP, true>(dst, ps, BB[i], c), WCLASS[i], f0[i], f1[i]);
    } else {
        // direction i to slot i of the pair (x, x + 1) + e_i.  The split
        // pairs of one x direction are stored in one block per side, so the
        // lanes' masks change four times per thread, not twice per direction.
        constexpr bool NTS = IMX_STORE_POLICY != 2;  // shifted stores: see st_ptr
#pragma unroll
        for (int i = 0; i < Q; ++i)
            if (EX[i] == 0)
                st_pair<S, true>(pop_ptr<S, AM_GROUP, true>(dst, ps, i, rowb[1 + EZ[i]][1 + EY[i]] + x), WCLASS[i],
                                 f0[i], f1[i]);
#pragma unroll
        for (int side = 0; side < 2; ++side) {
            const int ex = side == 0 ? 1 : -1;
            const bool split = ex > 0 ? hi : lo;
            if (!split) {
#pragma unroll
                for (int i = 0; i < Q; ++i)
                    if (EX[i] == ex)
                        st_pair<S, NTS>(pop_ptr<S, AM_GROUP, true>(dst, ps, i, rowb[1 + EZ[i]][1 + EY[i]] + x + ex),
                                        WCLASS[i], f0[i], f1[i]);
            } else {
#pragma unroll
                for (int i = 0; i < Q; ++i)
                    if (EX[i] == ex) {
                        const uint32_t row = rowb[1 + EZ[i]][1 + EY[i]];
                        st_ptr<S, NTS>(pop_ptr<S, AM_GROUP, true>(dst, ps, i, row + nx - 1), WCLASS[i], f0[i]);
                        st_ptr<S, NTS>(pop_ptr<S, AM_GROUP, true>(dst, ps, i, row), WCLASS[i], f1[i]);
                    }
            }
        }
    }
}

template <typename S, int CPT, bool NTL, int AM>
hipError_t launch_fused_cpt(int bc, int pattern, const FusedParams &P, int block, hipStream_t st)
{
    dim3 grid(P.nblocks), blk(block);
#define IMX_L(B, PT) hipLaunchKernelGGL((k_fused<S, B, PT, CPT, NTL, AM>), grid, blk, 0, st, P)
    if (bc == BC_PERIODIC) {
        if (pattern == PAT_PULL) IMX_L(BC_PERIODIC, PAT_PULL);
        else if (pattern == PAT_AA_EVEN) IMX_L(BC_PERIODIC, PAT_AA_EVEN);
        else IMX_L(BC_PERIODIC, PAT_AA_ODD);
    } else if (bc == BC_CAVITY) {
        if (pattern == PAT_PULL) IMX_L(BC_CAVITY, PAT_PULL);
        else if (pattern == PAT_AA_EVEN) IMX_L(BC_CAVITY, PAT_AA_EVEN);
        else IMX_L(BC_CAVITY, PAT_AA_ODD);
    } else if (bc == BC_CHANNEL) {
        if (pattern != PAT_PULL) return hipErrorInvalidValue;
        IMX_L(BC_CHANNEL, PAT_PULL);
    } else {
        // the faithful cavity: FP64, pull storage
        if constexpr (std::is_same<S, double>::value) {
            if (pattern != PAT_PULL) return hipErrorInvalidValue;
            IMX_L(BC_FAITHFUL, PAT_PULL);
        } else {
            return hipErrorInvalidValue;
        }
    }
#undef IMX_L
    return hipGetLastError();
}

template <typename S>
hipError_t launch_fused_t(int bc, int pattern, const FusedParams &P, int block, hipStream_t st)
{
    static_assert(16 / sizeof(S) == (size_t)PopGroup<S>::n, "group of 16 bytes of stride per cell");
    const int am = fused_addr_mode(sizeof(S), P.pstride, P.am_max);
    const int cpt = fused_cells_per_thread(P.cpt, bc, P.nx, am);
    if (cpt == 3) {
        FusedParams P2 = P;
        P2.nblocks = (P.ncell / 2 + (uint32_t)block - 1) / (uint32_t)block;
        dim3 grid(P2.nblocks), blk(block);
#define IMX_L(PT)                                                                 \
    do {                                                                          \
        if (P.ntl) hipLaunchKernelGGL((k_fused_pair<S, PT, true>), grid, blk, 0, st, P2); \
        else hipLaunchKernelGGL((k_fused_pair<S, PT, false>), grid, blk, 0, st, P2);      \
    } while (0)
        if (pattern == PAT_PULL) IMX_L(PAT_PULL);
        else if (pattern == PAT_AA_EVEN) IMX_L(PAT_AA_EVEN);
        else IMX_L(PAT_AA_ODD);
#undef IMX_L
        return hipGetLastError();
    }
    // larger lattices, and pairs where the lattice does not allow them: one
    // cell per thread (the same results), so the grid covers block cells per
    // workgroup whatever cells per thread was asked
    FusedParams P1 = P;
    P1.cpt = 1;
    P1.nblocks = (P.ncell + (uint32_t)block - 1) / (uint32_t)block;
    if (am == AM_GROUP) {
        if (cpt == 2) {
            if (P.ntl) return launch_fused_cpt<S, 2, true, AM_GROUP>(bc, pattern, P, block, st);
            return launch_fused_cpt<S, 2, false, AM_GROUP>(bc, pattern, P, block, st);
        }
        if (P.ntl) return launch_fused_cpt<S, 1, true, AM_GROUP>(bc, pattern, P1, block, st);
        return launch_fused_cpt<S, 1, false, AM_GROUP>(bc, pattern, P1, block, st);
    }
    if (am == AM_W32) {
        if (P.ntl) return launch_fused_cpt<S, 1, true, AM_W32>(bc, pattern, P1, block, st);
        return launch_fused_cpt<S, 1, false, AM_W32>(bc, pattern, P1, block, st);
    }
    // beyond 4 GiB per population: 64-bit addresses, plain loads
    return launch_fused_cpt<S, 1, false, AM_WIDE>(bc, pattern, P1, block, st);
}

}  // namespace imx
