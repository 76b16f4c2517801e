// Kernel parameter blocks and host-side launchers shared by the HIP
// translation units and the C-ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lattice.h"

// Bytes added to each population stride of a fused-engine lattice (after
// rounding the cell count up to 256 bytes): 9 x 256 keeps the 27 streams off a
// common power-of-two alignment.
#ifndef IMEXLB_DEFAULT_PAD
#define IMEXLB_DEFAULT_PAD 2304
#endif

namespace imx {

// BC_FAITHFUL: the cavity with the reference's own wall, lid and halo rules;
// the first ghost layer lives in a ring of face arrays (FusedParams::ring).
enum Bc : int { BC_PERIODIC = 0, BC_CAVITY = 1, BC_CHANNEL = 2, BC_FAITHFUL = 3 };
enum Pattern : int { PAT_PULL = 0, PAT_AA_EVEN = 1, PAT_AA_ODD = 2 };
enum Dtype : int { DT_F64 = 0, DT_F32 = 1, DT_F16 = 2 };

// The faithful sphere's inlet perturbation replaces Kokkos' random pool by a
// seeded counter hash: the splitmix64 finaliser (Steele, Lea and Flood, "Fast
// splittable pseudorandom number generators", OOPSLA 2014).
__host__ __device__ __forceinline__ uint64_t mix64(uint64_t x)
{
    x ^= x >> 30;
    x *= 0xBF58476D1CE4E5B9ULL;
    x ^= x >> 27;
    x *= 0x94D049BB133111EBULL;
    x ^= x >> 31;
    return x;
}

// Division by a run-time invariant 32-bit divisor (Granlund & Montgomery,
// "Division by invariant integers using multiplication", PLDI 1994, fig. 4.1):
// exact for every 32-bit dividend.
struct FastDiv {
    uint32_t d, m, s1, s2;
    void init(uint32_t div)
    {
        d = div;
        uint32_t l = 0;
        while ((uint64_t(1) << l) < div) ++l;
        m = uint32_t(((uint64_t(1) << 32) * ((uint64_t(1) << l) - div)) / div + 1);
        s1 = l < 1 ? l : 1;
        s2 = l > 0 ? l - 1 : 0;
    }
#if defined(__HIPCC__)
    __device__ __forceinline__ uint32_t div(uint32_t n) const
    {
        uint32_t t = __umulhi(m, n);
        return (t + ((n - t) >> s1)) >> s2;
    }
#endif
};

// One fused stream + moments + collide launch over a single domain.
// The address mode the fused kernels run with (fused.h AddrMode): 2 (a
// uniform base per group of 16 / elem populations, 32-bit offsets) while such
// a group spans at most 4 GiB, else 1 (a uniform base per population) while a
// population does, else 0 (64-bit offsets); capped at am_max.
inline int fused_addr_mode(uint64_t elem, uint64_t pstride, int am_max)
{
    const uint64_t lim = uint64_t(1) << 32;
    if (am_max >= 2 && pstride * 16 <= lim) return 2;
    if (am_max >= 1 && pstride * elem <= lim) return 1;
    return 0;
}

// Doubles per row of FusedParams::stage: the nine populations of one row of
// the lattice, together.
constexpr int STAGE_ROW = 9;

// The cells per thread the fused kernels run with for a requested `cpt`:
// two adjacent cells (3) only on periodic lattices with nx even and grouped
// bases (address mode 2), two cells blockDim apart (2) only with grouped
// bases, else one.
inline int fused_cells_per_thread(int cpt, int bc, uint32_t nx, int am)
{
    if (am != 2) return 1;
    if (cpt == 3) return bc == BC_PERIODIC && nx % 2 == 0 ? 3 : 1;
    return cpt == 2 ? 2 : 1;
}

struct FusedParams {
    int nx, ny, nz;
    uint32_t ncell;          // nx * ny * nz
    uint64_t pstride;        // elements between consecutive populations
    const void *src;         // pull: lattice read; AA: the single lattice
    void *dst;               // pull: lattice written; AA: same as src
    double *mom_p, *mom_u, *mom_v, *mom_w;  // written when mom_p != nullptr
    const uint8_t *mask;     // channel: 0 fluid, 1 solid, 2 fluid next to solid
    double tau0;             // f <- f - (f - feq) / (tau0 + 0.5)
    float omega_f;           // 1 / (tau0 + 0.5) for the FP32 collision
    double shift;            // reduced precision: S, with sum_i s_i = S (p = (sum g + S) / 3)
    double ush;              // reduced-precision channel: the shift's free-stream x velocity U
                             // (u = sum g e + U); 0 otherwise
    double sh[Q];            // s_i (storage_shift), for the transfer and force kernels
    double corr[Q];          // cavity: moving-lid term for a link crossing z = nz
    double feq_in[Q];        // channel: inlet equilibrium, stored convention
    uint32_t nblocks;        // grid size (for the XCD-aware block order)
    int xcd_remap;           // 1: give each XCD one contiguous run of tiles
    int cpt;                 // cells per thread: 1, 2 (blockDim apart) or 3 (two adjacent
                             // x cells; periodic lattices with nx even, else 1)
    int ntl;                 // 1: non-temporal loads of unshared lines
    int am_max;              // widest-saving address mode allowed (AddrMode in fused.h)
    int bc;                  // boundary kind (read by the ensemble kernel)
    FastDiv divx, divy;
    // BC_FAITHFUL: the first ghost layer, population-major (q * ring_n + r).
    // Cell r of the ring: the two x faces (x = -1, nx; y in [-1, ny], z in
    // [-1, nz]; ring_a cells each) first, then the two y faces (y = -1, ny;
    // x in [0, nx), z in [-1, nz]; ring_b each), then the two z faces (z =
    // -1, nz; x in [0, nx), y in [0, ny); ring_c each), each x fastest.
    double *ring;
    int64_t ring_rel;        // (ring - src) in elements: ring loads address from the lattice's base
    uint32_t ring_n, ring_a, ring_b, ring_c;
    int faithful_sphere;     // BC_FAITHFUL: 0 the cavity, 1 the sphere channel
    // faithful mode: the populations the x-face boundary kernels read,
    // written by each fused pass: stage[((side * nz + z) * ny + y) * STAGE_ROW
    // + r] holds, for the cavity's walls, the r-th population with e_x = -1
    // at x = 1 (side 0) or with e_x = +1 at x = nx - 2 (side 1); for the
    // sphere's inlet and outlet copies, the r-th with e_x = +1 at x = 0 (side
    // 0) or with e_x = -1 at x = nx - 1 (side 1); each row's nine together;
    // null elsewhere
    double *stage;
    double u0;               // faithful sphere: inlet speed of the parabolic profile
    uint64_t seed;           // faithful sphere: the initial perturbation's seed
};

// The ring index of ghost cell (i, j, k), -2 for an interior cell and -1 for
// a ghost cell deeper than the first layer (those stay 0 in the faithful
// cavity: nothing writes them and Initialize makes f = 0).
// (selects rather than branches: a kernel's loads stay free of waits)
__host__ __device__ __forceinline__ int64_t ring_index(const FusedParams &P, int i, int j, int k)
{
    const bool deep = i < -1 || i > P.nx || j < -1 || j > P.ny || k < -1 || k > P.nz;
    const bool xf = i == -1 || i == P.nx, yf = j == -1 || j == P.ny, zf = k == -1 || k == P.nz;
    const int64_t rx = (i == P.nx ? (int64_t)P.ring_a : 0) + (int64_t)(k + 1) * (P.ny + 2) + (j + 1);
    const int64_t ry = 2 * (int64_t)P.ring_a + (j == P.ny ? (int64_t)P.ring_b : 0) + (int64_t)(k + 1) * P.nx + i;
    const int64_t rz =
        2 * (int64_t)(P.ring_a + P.ring_b) + (k == P.nz ? (int64_t)P.ring_c : 0) + (int64_t)j * P.nx + i;
    return deep ? -1 : (xf ? rx : (yf ? ry : (zf ? rz : -2)));
}

// The faithful cavity's unpack and wall kernels on the ring (aux.hip):
// Streaming's boundary phase, before the fused pass.
hipError_t launch_faithful_boundaries(const FusedParams &P, hipStream_t st);
// Initialize's values for the ring (0 in the cavity)
hipError_t launch_faithful_ring_init(const FusedParams &P, hipStream_t st);
// fills FusedParams::stage from the lattice P.src (after Initialize, set_f)
hipError_t launch_faithful_stage_fill(const FusedParams &P, hipStream_t st);

// The faithful sphere: solid iff (x-c)^2 + (y-c)^2 + (z-c)^2 <= r^2 with the
// integer centre c = (glx - 1) / 8 on every axis and r = (glx - 1) / 64
// (sphere/src/lbm.cpp:238).
__host__ __device__ __forceinline__ bool faithful_sphere_solid(int glx, int x, int y, int z)
{
    const int c = (glx - 1) / 8;
    const int di = x - c, dj = y - c, dk = z - c;
    const double r = (glx - 1) / 64.0;
    return (double)di * (double)di + (double)dj * (double)dj + (double)dk * (double)dk <= r * r;
}

// Storage shift of direction i: FP32/FP16 lattices store g_i = f_i - s_i.
// The TGV and the cavity (U = 0) subtract s_i = w_i S, S = rho0; the open
// channel subtracts its inflow equilibrium feq_i(S/3, (U, 0, 0)) in the
// oracle's evaluation order, so that the stored values are the flow's
// departure from the free stream and the inflow itself is stored as 0.
inline void storage_shift(double S, double U, double out[Q])
{
    const double pin = S * (1.0 / 3.0);
    for (int i = 0; i < Q; ++i) {
        if (U == 0.0) {
            out[i] = w(i) * S;
        } else {
            const double edu = (double)EX[i] * U, udu = U * U, eu2 = edu * edu;
            out[i] = w(i) * pin * 3.0 + w(i) * (3.0 * edu + 4.5 * eu2 - 1.5 * udu);
        }
    }
}

// Launchers (defined next to their kernels).
hipError_t launch_fused(int dtype, int bc, int pattern, const FusedParams &P, int block, hipStream_t st);

struct InitParams {
    int nx, ny, nz;
    uint32_t ncell;
    uint64_t pstride;
    void *dst;               // lattice written (stored convention)
    int layout;              // 0 natural (pull), 1 swapped (AA after an even step)
    int kind;                // 0 TGV from the per-axis tables, 1 uniform (u0,0,0)/p0, masked,
                             // 2 the faithful sphere (parabolic inlet profile, seeded perturbation)
    const double *tab;       // TGV: 9 tables [sx cx c2x | sy cy c2y | sz cz c2z]
    double u0, rho0, cs2, p0;
    double sh[Q];            // storage shift s_i
    const uint8_t *mask;
    uint64_t seed;           // kind 2
    double *mom[4];          // p, u, v, w of the initial fields (solid cells: 0);
                             // null: not written
    FastDiv divx, divy;
};
hipError_t launch_init(int dtype, const InitParams &P, hipStream_t st);

// Copy between the stored lattice (any dtype, pull or AA layout after step
// count parity) and a natural-order FP64 array of post-collision populations.
struct XferParams {
    int nx, ny, nz;
    uint32_t ncell;
    uint64_t pstride;
    void *lat;
    double *buf;             // [Q][nz][ny][nx]
    int layout;              // 0 natural, 1 swapped (AA even), 2 streamed (AA odd)
    int bc;                  // periodic or cavity (AA streamed layout needs the rule)
    double sh[Q];            // storage shift s_i
    double corr[Q];
    FastDiv divx, divy;
};
hipError_t launch_get_f(int dtype, const XferParams &P, hipStream_t st);
hipError_t launch_set_f(int dtype, const XferParams &P, hipStream_t st);

// Deterministic FP64 reductions.
hipError_t launch_diag(const double *p, const double *u, const double *v, const double *w,
                       const uint8_t *mask, uint32_t n, double *partials, int nparts,
                       double *out2, hipStream_t st);
hipError_t launch_force(int dtype, const FusedParams &P, double *partials, int nparts,
                        double *out3, hipStream_t st);

hipError_t launch_copy(const double *src, double *dst, size_t n, hipStream_t st);

// Sets the message imexlb_last_error(NULL) returns on this thread (entry
// points without a handle).
void set_call_error(const char *fmt, ...) __attribute__((format(printf, 1, 2)));
// Bandwidth probe `kind` (imexlb_bandwidth_probe) over n doubles; returns the
// number of streams it moves (2 for copies, 1 for read- or write-only), 0 for
// an unknown kind.
int bandwidth_probe_streams(int kind);
hipError_t launch_bandwidth_probe(int kind, const double *src, double *dst, size_t n, int cus, hipStream_t st);

// flags[d] = 1 if the moments mom[4d .. 4d+3] (ncell[d] cells each) are all
// finite, else 0.
hipError_t launch_finite_flags(const double *const *mom, const uint32_t *ncell, int ndom, int32_t *flags,
                               hipStream_t st);

}  // namespace imx

#include <string>

namespace imx {

// Autotuner cache (JSON lines), shared by the single-domain and ensemble tuners.
struct TuneChoice {
    int block, cpt, grid_or_xcd, ntl;
    float ms;
};
std::string tune_key(int device, const char *kind, int dtype, int storage, const std::string &shape);
bool tune_cache_read(const char *path, const std::string &key, TuneChoice &out);
void tune_cache_write(const char *path, const std::string &key, const TuneChoice &c);

}  // namespace imx
