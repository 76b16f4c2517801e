/*
 * imexlb_oracle.c -- CPU oracle for imexlb's D3Q27 BGK time step (FP64).
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker, not the product: only
 * the tests, the smoke() entry point and the CPU-baseline leg of bench.py load
 * it.  Nothing under imexlb_amd/ links, loads or calls it.
 *
 * What it restates (paths relative to the reference checkout):
 *   lattice tables        tgv/src/lbm.cpp:27-190 (same in all three cases)
 *   Initialize            tgv/src/lbm.cpp:195-218
 *                         lid_driven_cavity/src/lbm.cpp:191-210
 *                         sphere/src/lbm.cpp:221-254
 *   Collision             tgv/src/lbm.cpp:222-237 (= lid_driven_cavity/src/lbm.cpp:214-229
 *                         = sphere/src/lbm.cpp:257-272)
 *   pack/exchange/unpack  tgv/src/mpi_view_transfer.cpp:117-383 (one rank: periodic)
 *                         lid_driven_cavity/src/mpi_view_transfer.cpp:84-166,598-690
 *                         (one rank: four ghost planes receive never-written buffers)
 *                         sphere/src/mpi_view_transfer.cpp:116-439 (one rank: no effect)
 *   Streaming             tgv/src/lbm.cpp:239-256
 *                         lid_driven_cavity/src/lbm.cpp:231-317
 *                         sphere/src/lbm.cpp:274-371
 *   Update                tgv/src/lbm.cpp:288-333
 *                         lid_driven_cavity/src/lbm.cpp:319-349
 *                         sphere/src/lbm.cpp:374-417
 *
 * Data layout follows the reference: ghost depth 3, every field includes the
 * ghosts (lx = nx + 6), populations f(q, lx, ly, lz).  Index order is
 * Kokkos-OpenMP's LayoutRight (last index fastest), the layout the reference
 * takes when its Views live in a host space.
 *
 * Two boundary modes:
 *   ORC_FAITHFUL  the reference's own boundary code, quirks included:
 *                 wall bounce-back f_i(g) = f*_bb(i)(g + 2 e_i); the cavity's
 *                 top wall reads j + e_y (lid_driven_cavity/src/lbm.cpp:298);
 *                 the lid is imposed by overwriting ua = 0.1 in Update; the
 *                 cavity's one-rank unpack writes receive buffers that were
 *                 never filled -- defined here as zeros; the sphere's Kokkos
 *                 random pool is replaced by a seeded counter hash.
 *   ORC_CONFIG    textbook boundaries of the benchmark configurations, written
 *                 in the same ghost-fill-then-pull structure: halfway
 *                 bounce-back f_i(g) = f*_bb(i)(g + e_i), moving lid by the
 *                 momentum term 6 w_i rho0 (e_i . u_lid), sphere mask with
 *                 halfway bounce-back, equilibrium inlet ghost, zero-gradient
 *                 outlet ghost, periodic y/z.
 * The Taylor-Green case is identical in both modes.
 *
 * Arithmetic is written operation by operation in the reference's evaluation
 * order and compiled with -ffp-contract=off.  Two deliberate restatements:
 * pow(x, 2) is written x * x (the exact square, correctly rounded), and the
 * uninitialised LBM::cs2 used by the TGV pressure (tgv/src/lbm.hpp:135,
 * tgv/src/lbm.cpp:202) is defined as 1/3, the value System computes
 * (tgv/src/System.cpp:8).  Moments are summed in direction order 0..26, the
 * order of the cavity and sphere host loops (lid_driven_cavity/src/lbm.cpp:
 * 329-347); the TGV Update reduces with Kokkos' ThreadVectorRange, whose order
 * is not specified.
 *
 * Parity pins (tests/test_oracle_*.py): lattice isotropy identities; f = feq
 * reproduces p and u; collision is a fixed point at equilibrium and conserves
 * sum f and sum f e; the one-rank halo equals modular indexing; streaming is a
 * permutation; an independent NumPy restatement (oracle/oracle_np.py) agrees
 * on small grids for every case and mode; TGV early-time decay matches the
 * analytic viscous decay; mass is conserved in the periodic box.  The sphere's
 * random perturbation and the cavity's never-written receive buffers have no
 * reference values to pin against (parity unpinned for those two inputs; both
 * are defined here).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define Q 27
#define GH 3 /* tgv/src/lbm.hpp:22 */
#define PI_R 3.1415926 /* tgv/src/lbm.cpp:5 */

enum { ORC_TGV = 0, ORC_LDC = 1, ORC_SPHERE = 2 };
enum { ORC_FAITHFUL = 0, ORC_CONFIG = 1 };

/* Discrete velocities, weights and bounce-back partners in the reference's
 * order: tgv/src/lbm.cpp:27-53 (t), 55-81 (bb), 84-190 (e). */
static const int E[Q][3] = {
    {0, 0, 0},   {1, 0, 0},   {-1, 0, 0},  {0, 1, 0},   {0, -1, 0},  {0, 0, 1},
    {0, 0, -1},  {1, 1, 0},   {-1, -1, 0}, {1, -1, 0},  {-1, 1, 0},  {1, 0, 1},
    {-1, 0, -1}, {1, 0, -1},  {-1, 0, 1},  {0, 1, 1},   {0, -1, -1}, {0, 1, -1},
    {0, -1, 1},  {1, 1, 1},   {-1, -1, -1}, {1, -1, 1}, {-1, 1, -1}, {1, 1, -1},
    {-1, -1, 1}, {1, -1, -1}, {-1, 1, 1}};
static const int BB[Q] = {0,  2,  1,  4,  3,  6,  5,  8,  7,  10, 9,  12, 11, 14,
                          13, 16, 15, 18, 17, 20, 19, 22, 21, 24, 23, 26, 25};
static double T[Q];

static void init_tables(void)
{
    T[0] = 8.0 / 27.0;
    for (int i = 1; i <= 6; i++) T[i] = 2.0 / 27.0;
    for (int i = 7; i <= 18; i++) T[i] = 1.0 / 54.0;
    for (int i = 19; i <= 26; i++) T[i] = 1.0 / 216.0;
}

typedef struct {
    int case_id;
    int mode;
    int nx, ny, nz;
    double tau0, rho0, u0;
    double lid_u;                      /* config-mode cavity lid speed */
    double sph_cx, sph_cy, sph_cz, sph_r; /* config-mode sphere */
    unsigned long long seed;           /* faithful-mode sphere perturbation */
    int threads;                       /* OpenMP threads, 0 = runtime default */
} orc_config;

typedef struct {
    orc_config c;
    int glx, gly, glz;
    int lx, ly, lz;
    int l_s[3], l_e[3], l_l[3];
    int x_lo, x_hi, y_lo, y_hi, z_lo, z_hi;
    double cs2;
    double *f, *ft;
    double *ua, *va, *wa, *p;
    int *usr, *ran;
    double *buf[26]; /* halo staging, one per message tag */
    double feq_in[Q];
    double lidcorr[Q];
} orc_t;

static inline size_t F(const orc_t *s, int ii, int i, int j, int k)
{
    return (((size_t)ii * s->lx + i) * s->ly + j) * s->lz + k;
}
static inline size_t C(const orc_t *s, int i, int j, int k)
{
    return ((size_t)i * s->ly + j) * s->lz + k;
}

/* feq in the reference's evaluation order (tgv/src/lbm.cpp:228-232). */
static inline double feq_of(int ii, double u, double v, double w, double p)
{
    double edu = (double)E[ii][0] * u + (double)E[ii][1] * v + (double)E[ii][2] * w;
    double udu = u * u + v * v + w * w;
    double eu2 = edu * edu;
    return T[ii] * p * 3.0 + T[ii] * (3.0 * edu + 4.5 * eu2 - 1.5 * udu);
}

static uint64_t mix64(uint64_t x)
{
    x ^= x >> 30;
    x *= 0xBF58476D1CE4E5B9ULL;
    x ^= x >> 27;
    x *= 0x94D049BB133111EBULL;
    x ^= x >> 31;
    return x;
}

/* Seeded stand-in for Random_XorShift64_Pool(5374857).urand(100)
 * (sphere/src/lbm.cpp:221-229): an integer in [0, 100) per cell. */
int orc_sphere_ran(unsigned long long seed, int i, int j, int k, int ly, int lz)
{
    uint64_t idx = ((uint64_t)i * (uint64_t)ly + (uint64_t)j) * (uint64_t)lz + (uint64_t)k;
    return (int)(mix64((uint64_t)seed + idx * 0x9E3779B97F4A7C15ULL) % 100ULL);
}

static int is_solid_config(const orc_t *s, int i, int j, int k)
{
    double dx = (double)(i - GH) - s->c.sph_cx;
    double dy = (double)(j - GH) - s->c.sph_cy;
    double dz = (double)(k - GH) - s->c.sph_cz;
    return dx * dx + dy * dy + dz * dz <= s->c.sph_r * s->c.sph_r;
}

void orc_free(orc_t *s)
{
    if (!s) return;
    free(s->f);
    free(s->ft);
    free(s->ua);
    free(s->va);
    free(s->wa);
    free(s->p);
    free(s->usr);
    free(s->ran);
    for (int m = 0; m < 26; m++) free(s->buf[m]);
    free(s);
}

/* The one-rank LBM constructor (tgv/src/lbm.hpp:161-236): rx = ry = rz = 1,
 * l_l = global extent, x_lo = 0, x_hi = glx - 1. */
orc_t *orc_create(const orc_config *cfg)
{
    init_tables();
    if (cfg->nx < 1 || cfg->ny < 1 || cfg->nz < 1) return NULL;
    orc_t *s = (orc_t *)calloc(1, sizeof(orc_t));
    if (!s) return NULL;
    s->c = *cfg;
    s->glx = cfg->nx;
    s->gly = cfg->ny;
    s->glz = cfg->nz;
    s->l_l[0] = s->glx;
    s->l_l[1] = s->gly;
    s->l_l[2] = s->glz;
    s->lx = s->l_l[0] + 2 * GH;
    s->ly = s->l_l[1] + 2 * GH;
    s->lz = s->l_l[2] + 2 * GH;
    for (int d = 0; d < 3; d++) {
        s->l_s[d] = GH;
        s->l_e[d] = GH + s->l_l[d];
    }
    s->x_lo = 0;
    s->y_lo = 0;
    s->z_lo = 0;
    s->x_hi = s->glx - 1;
    s->y_hi = s->gly - 1;
    s->z_hi = s->glz - 1;
    s->cs2 = 1.0 / 3.0;
    size_t ncell = (size_t)s->lx * s->ly * s->lz;
    s->f = (double *)malloc(sizeof(double) * Q * ncell);
    s->ft = (double *)malloc(sizeof(double) * Q * ncell);
    s->ua = (double *)malloc(sizeof(double) * ncell);
    s->va = (double *)malloc(sizeof(double) * ncell);
    s->wa = (double *)malloc(sizeof(double) * ncell);
    s->p = (double *)malloc(sizeof(double) * ncell);
    s->usr = (int *)malloc(sizeof(int) * ncell);
    s->ran = (int *)malloc(sizeof(int) * ncell);
    if (!s->f || !s->ft || !s->ua || !s->va || !s->wa || !s->p || !s->usr || !s->ran) {
        orc_free(s);
        return NULL;
    }
    /* halo staging sizes: 6 faces q*3*(full)*(full), 12 edges q*3*3*l_l,
     * 8 corners q*27 (tgv/src/mpi_view_transfer.cpp:8-59) */
    size_t fx = (size_t)Q * GH * s->ly * s->lz, fy = (size_t)Q * s->lx * GH * s->lz,
           fz = (size_t)Q * s->lx * s->ly * GH;
    size_t sz[26] = {fx, fx, fz, fz, fy, fy};
    for (int m = 6; m < 18; m++) sz[m] = (size_t)Q * GH * GH * (s->lx > s->ly ? (s->lx > s->lz ? s->lx : s->lz) : (s->ly > s->lz ? s->ly : s->lz));
    for (int m = 18; m < 26; m++) sz[m] = (size_t)Q * GH * GH * GH;
    for (int m = 0; m < 26; m++) {
        s->buf[m] = (double *)calloc(sz[m], sizeof(double));
        if (!s->buf[m]) {
            orc_free(s);
            return NULL;
        }
    }
#ifdef _OPENMP
    if (cfg->threads > 0) omp_set_num_threads(cfg->threads);
#endif
    /* config-mode boundary constants, computed once in a fixed order */
    double pin = s->c.rho0 * s->cs2;
    for (int ii = 0; ii < Q; ii++) {
        s->feq_in[ii] = feq_of(ii, s->c.u0, 0.0, 0.0, pin);
        s->lidcorr[ii] = 6.0 * T[ii] * s->c.rho0 * ((double)E[ii][0] * s->c.lid_u);
    }
    return s;
}

int orc_dims(const orc_t *s, int *out)
{
    out[0] = s->lx;
    out[1] = s->ly;
    out[2] = s->lz;
    out[3] = GH;
    return 0;
}

/* ---------------------------------------------------------------- Initialize */

static void init_f_from_fields(orc_t *s)
{
    /* "initf" (tgv/src/lbm.cpp:209-218): f = feq(p, u) everywhere, ft = 0 */
#pragma omp parallel for collapse(2) schedule(static)
    for (int ii = 0; ii < Q; ii++)
        for (int i = 0; i < s->lx; i++)
            for (int j = 0; j < s->ly; j++)
                for (int k = 0; k < s->lz; k++) {
                    size_t c = C(s, i, j, k);
                    s->f[F(s, ii, i, j, k)] = feq_of(ii, s->ua[c], s->va[c], s->wa[c], s->p[c]);
                    s->ft[F(s, ii, i, j, k)] = 0.0;
                }
}

void orc_initialize(orc_t *s)
{
    const int lx = s->lx, ly = s->ly, lz = s->lz;
    const double u0 = s->c.u0, rho0 = s->c.rho0;
    if (s->c.case_id == ORC_TGV) {
        /* tgv/src/lbm.cpp:195-206 */
#pragma omp parallel for collapse(2) schedule(static)
        for (int i = 0; i < lx; i++)
            for (int j = 0; j < ly; j++)
                for (int k = 0; k < lz; k++) {
                    double ax = (double)(i - GH + s->x_lo) / (double)s->glx * 2.0 * PI_R;
                    double ay = (double)(j - GH + s->y_lo) / (double)s->gly * 2.0 * PI_R;
                    double az = (double)(k - GH + s->z_lo) / (double)s->glz * 2.0 * PI_R;
                    size_t c = C(s, i, j, k);
                    s->ua[c] = u0 * sin(ax) * cos(ay) * cos(az);
                    s->va[c] = -u0 * cos(ax) * sin(ay) * cos(az);
                    s->wa[c] = 0;
                    s->p[c] = rho0 * s->cs2 +
                              rho0 * u0 * u0 / 16.0 * (cos(ax * 2.0) + cos(ay * 2.0)) * (cos(az * 2.0) + 2.0);
                    s->usr[c] = 1;
                    s->ran[c] = 0;
                }
    } else if (s->c.case_id == ORC_LDC) {
        /* faithful: lid_driven_cavity/src/lbm.cpp:191-199 (rest, p = 0).
         * config: rest at uniform density rho0, i.e. p = rho0 cs2. */
        double p0 = s->c.mode == ORC_FAITHFUL ? 0.0 : rho0 * s->cs2;
        for (size_t c = 0; c < (size_t)lx * ly * lz; c++) {
            s->ua[c] = 0;
            s->va[c] = 0;
            s->wa[c] = 0;
            s->p[c] = p0;
            s->usr[c] = 1;
            s->ran[c] = 0;
        }
    } else {
        if (s->c.mode == ORC_FAITHFUL) {
            /* sphere/src/lbm.cpp:221-242 */
            const int glx = s->glx, glz = s->glz;
#pragma omp parallel for collapse(2) schedule(static)
            for (int i = 0; i < lx; i++)
                for (int j = 0; j < ly; j++)
                    for (int k = 0; k < lz; k++) {
                        size_t c = C(s, i, j, k);
                        s->ran[c] = orc_sphere_ran(s->c.seed, i, j, k, ly, lz);
                        int di = s->x_lo + i - GH - (glx - 1) / 8;
                        int dj = s->y_lo + j - GH - (glx - 1) / 8;
                        int dk = s->z_lo + k - GH - (glx - 1) / 8;
                        double r = (glx - 1) / 64.0;
                        s->usr[c] = ((double)di * (double)di + (double)dj * (double)dj +
                                         (double)dk * (double)dk <=
                                     r * r)
                                        ? 0
                                        : 1;
                        int kz = s->z_lo + k - GH;
                        double gm = (double)(glz - 1);
                        s->ua[c] = (u0 * 4.0 * (double)kz * (double)(glz - 1 - kz) / (gm * gm)) *
                                   (double)s->usr[c] * (1.0 + (double)s->ran[c] * 0.001);
                        s->va[c] = 0.0;
                        s->wa[c] = 0.0;
                        s->p[c] = 0;
                    }
        } else {
            /* config: uniform u = (u0, 0, 0) in the fluid, rest in the solid */
            double p0 = rho0 * s->cs2;
#pragma omp parallel for collapse(2) schedule(static)
            for (int i = 0; i < lx; i++)
                for (int j = 0; j < ly; j++)
                    for (int k = 0; k < lz; k++) {
                        size_t c = C(s, i, j, k);
                        s->ran[c] = 0;
                        s->usr[c] = is_solid_config(s, i, j, k) ? 0 : 1;
                        s->ua[c] = s->usr[c] ? u0 : 0.0;
                        s->va[c] = 0.0;
                        s->wa[c] = 0.0;
                        s->p[c] = p0;
                    }
        }
    }
    init_f_from_fields(s);
}

/* ---------------------------------------------------------------- Collision */

void orc_collision(orc_t *s)
{
    /* tgv/src/lbm.cpp:226-235 over the interior; config-mode sphere skips the
     * solid cells */
    const double tau0 = s->c.tau0;
    const int skip_solid = (s->c.case_id == ORC_SPHERE && s->c.mode == ORC_CONFIG);
#pragma omp parallel for collapse(2) schedule(static)
    for (int ii = 0; ii < Q; ii++)
        for (int i = s->l_s[0]; i < s->l_e[0]; i++)
            for (int j = s->l_s[1]; j < s->l_e[1]; j++)
                for (int k = s->l_s[2]; k < s->l_e[2]; k++) {
                    size_t c = C(s, i, j, k);
                    if (skip_solid && s->usr[c] == 0) continue;
                    double feq = feq_of(ii, s->ua[c], s->va[c], s->wa[c], s->p[c]);
                    size_t a = F(s, ii, i, j, k);
                    s->f[a] -= (s->f[a] - feq) / (tau0 + 0.5);
                }
}

/* ------------------------------------------------- pack / exchange / unpack */

/* A box [i0,i1) x [j0,j1) x [k0,k1) of f, copied to or from a staging buffer
 * in the subview's own index order (Kokkos::deep_copy of equal-extent views). */
typedef struct {
    int i0, i1, j0, j1, k0, k1;
} box;

/* Offset of (ii, i, j, k) in a box's staging buffer: the subview's index
 * order, so every (ii, i) slab is a contiguous run and the slabs can be
 * copied by separate threads (Kokkos-OpenMP runs deep_copy in parallel). */
static size_t box_off(box b, int ii, int i)
{
    return ((size_t)ii * (size_t)(b.i1 - b.i0) + (size_t)(i - b.i0)) * (size_t)(b.j1 - b.j0) * (size_t)(b.k1 - b.k0);
}
static void box_to_buf(const orc_t *s, box b, double *buf)
{
#pragma omp parallel for collapse(2) schedule(static)
    for (int ii = 0; ii < Q; ii++)
        for (int i = b.i0; i < b.i1; i++) {
            size_t n = box_off(b, ii, i);
            for (int j = b.j0; j < b.j1; j++)
                for (int k = b.k0; k < b.k1; k++) buf[n++] = s->f[F(s, ii, i, j, k)];
        }
}
static void buf_to_box(orc_t *s, box b, const double *buf)
{
#pragma omp parallel for collapse(2) schedule(static)
    for (int ii = 0; ii < Q; ii++)
        for (int i = b.i0; i < b.i1; i++) {
            size_t n = box_off(b, ii, i);
            for (int j = b.j0; j < b.j1; j++)
                for (int k = b.k0; k < b.k1; k++) s->f[F(s, ii, i, j, k)] = buf[n++];
        }
}

/* Pack boxes and unpack boxes of the periodic TGV halo, indexed by message tag
 * (mar = 1..26, tgv/src/mpi_view_transfer.cpp:176-312).  send[m] is what tag
 * m+1 sends (pack, :117-174); recv[m] is where tag m+1 is unpacked (:314-383).
 * On one rank every neighbour is the rank itself, so tag m's payload lands in
 * recv[m]. */
static void tgv_halo_boxes(const orc_t *s, box send[26], box recv[26])
{
    const int lx = s->lx, ly = s->ly, lz = s->lz, g = GH;
    const int *ls = s->l_s, *le = s->l_e;
    /* x: left = [g,2g) out, [0,g) in; right = [lx-2g,lx-g) out, [lx-g,lx) in.
     * z: down/up.  y: front/back.  Whole extent = [0, l). */
#define XL_OUT g, 2 * g
#define XR_OUT lx - 2 * g, lx - g
#define YF_OUT g, 2 * g
#define YB_OUT ly - 2 * g, ly - g
#define ZD_OUT g, 2 * g
#define ZU_OUT lz - 2 * g, lz - g
#define XL_IN 0, g
#define XR_IN lx - g, lx
#define YF_IN 0, g
#define YB_IN ly - g, ly
#define ZD_IN 0, g
#define ZU_IN lz - g, lz
#define XA 0, lx
#define YA 0, ly
#define ZA 0, lz
#define XI ls[0], le[0]
#define YI ls[1], le[1]
#define ZI ls[2], le[2]
    box snd[26] = {
        {XL_OUT, YA, ZA},         /* 1 leftout  -> right */
        {XR_OUT, YA, ZA},         /* 2 rightout -> left */
        {XA, YA, ZD_OUT},         /* 3 downout  -> up */
        {XA, YA, ZU_OUT},         /* 4 upout    -> down */
        {XA, YF_OUT, ZA},         /* 5 frontout -> back */
        {XA, YB_OUT, ZA},         /* 6 backout  -> front */
        {XL_OUT, YF_OUT, ZI},     /* 7 frontleftout  -> backright */
        {XR_OUT, YB_OUT, ZI},     /* 8 backrightout  -> frontleft */
        {XR_OUT, YF_OUT, ZI},     /* 9 frontrightout -> backleft */
        {XL_OUT, YB_OUT, ZI},     /* 10 backleftout  -> frontright */
        {XL_OUT, YI, ZU_OUT},     /* 11 leftupout    -> rightdown */
        {XR_OUT, YI, ZD_OUT},     /* 12 rightdownout -> leftup */
        {XL_OUT, YI, ZD_OUT},     /* 13 leftdownout  -> rightup */
        {XR_OUT, YI, ZU_OUT},     /* 14 rightupout   -> leftdown */
        {XI, YF_OUT, ZU_OUT},     /* 15 frontupout   -> backdown */
        {XI, YB_OUT, ZD_OUT},     /* 16 backdownout  -> frontup */
        {XI, YF_OUT, ZD_OUT},     /* 17 frontdownout -> backup */
        {XI, YB_OUT, ZU_OUT},     /* 18 backupout    -> frontdown */
        {XL_OUT, YF_OUT, ZD_OUT}, /* 19 fld -> bru */
        {XR_OUT, YB_OUT, ZU_OUT}, /* 20 bru -> fld */
        {XL_OUT, YB_OUT, ZU_OUT}, /* 21 blu -> frd */
        {XR_OUT, YF_OUT, ZD_OUT}, /* 22 frd -> blu */
        {XL_OUT, YB_OUT, ZD_OUT}, /* 23 bld -> fru */
        {XR_OUT, YF_OUT, ZU_OUT}, /* 24 fru -> bld */
        {XL_OUT, YF_OUT, ZU_OUT}, /* 25 flu -> brd */
        {XR_OUT, YB_OUT, ZD_OUT}, /* 26 brd -> flu */
    };
    box rcv[26] = {
        {XR_IN, YA, ZA},         /* 1 m_right */
        {XL_IN, YA, ZA},         /* 2 m_left */
        {XA, YA, ZU_IN},         /* 3 m_up */
        {XA, YA, ZD_IN},         /* 4 m_down */
        {XA, YB_IN, ZA},         /* 5 m_back */
        {XA, YF_IN, ZA},         /* 6 m_front */
        {XR_IN, YB_IN, ZI},      /* 7 m_backright */
        {XL_IN, YF_IN, ZI},      /* 8 m_frontleft */
        {XL_IN, YB_IN, ZI},      /* 9 m_backleft */
        {XR_IN, YF_IN, ZI},      /* 10 m_frontright */
        {XR_IN, YI, ZD_IN},      /* 11 m_rightdown */
        {XL_IN, YI, ZU_IN},      /* 12 m_leftup */
        {XR_IN, YI, ZU_IN},      /* 13 m_rightup */
        {XL_IN, YI, ZD_IN},      /* 14 m_leftdown */
        {XI, YB_IN, ZD_IN},      /* 15 m_backdown */
        {XI, YF_IN, ZU_IN},      /* 16 m_frontup */
        {XI, YB_IN, ZU_IN},      /* 17 m_backup */
        {XI, YF_IN, ZD_IN},      /* 18 m_frontdown */
        {XR_IN, YB_IN, ZU_IN},   /* 19 m_backrightup */
        {XL_IN, YF_IN, ZD_IN},   /* 20 m_frontleftdown */
        {XR_IN, YF_IN, ZD_IN},   /* 21 m_frontrightdown */
        {XL_IN, YB_IN, ZU_IN},   /* 22 m_backleftup */
        {XR_IN, YF_IN, ZU_IN},   /* 23 m_frontrightup */
        {XL_IN, YB_IN, ZD_IN},   /* 24 m_backleftdown */
        {XR_IN, YB_IN, ZD_IN},   /* 25 m_backrightdown */
        {XL_IN, YF_IN, ZU_IN},   /* 26 m_frontleftup */
    };
    memcpy(send, snd, sizeof(snd));
    memcpy(recv, rcv, sizeof(rcv));
}

/* Unpack order of the reference (tgv/src/mpi_view_transfer.cpp:318-380):
 * faces left,right,down,up,front,back; edges leftup,rightup,frontleft,
 * frontright,leftdown,rightdown,backleft,backright,frontup,frontdown,backup,
 * backdown; corners fld,frd,bld,brd,flu,fru,blu,bru -- as tag numbers. */
static const int TGV_UNPACK_ORDER[26] = {2,  1,  4,  3,  6,  5,  12, 13, 8,  10, 14, 11, 9,
                                         7,  16, 18, 17, 15, 20, 21, 24, 25, 26, 23, 22, 19};

void orc_pack(orc_t *s)
{
    if (s->c.case_id == ORC_TGV) {
        box snd[26], rcv[26];
        tgv_halo_boxes(s, snd, rcv);
        for (int m = 0; m < 26; m++) box_to_buf(s, snd[m], s->buf[m]);
    }
    /* The cavity and sphere packs only fill send buffers that one rank never
     * delivers: no effect on f. */
}

void orc_exchange(orc_t *s)
{
    (void)s; /* one rank: every message is to self; payload stays in buf[tag] */
}

static void fill_plane_zero_x(orc_t *s, int i)
{
    for (int ii = 0; ii < Q; ii++)
        for (int j = 0; j < s->ly; j++)
            for (int k = 0; k < s->lz; k++) s->f[F(s, ii, i, j, k)] = 0.0;
}
static void fill_plane_zero_y(orc_t *s, int j)
{
    for (int ii = 0; ii < Q; ii++)
        for (int i = 0; i < s->lx; i++)
            for (int k = 0; k < s->lz; k++) s->f[F(s, ii, i, j, k)] = 0.0;
}
static void fill_plane_zero_z(orc_t *s, int k)
{
    for (int ii = 0; ii < Q; ii++)
        for (int i = 0; i < s->lx; i++)
            for (int j = 0; j < s->ly; j++) s->f[F(s, ii, i, j, k)] = 0.0;
}

/* Periodic image fill of every ghost cell along the periodic axes (used by the
 * config-mode sphere, periodic in y and z). */
static void periodic_fill_yz(orc_t *s)
{
    const int ly = s->ly, lz = s->lz, Ny = s->l_l[1], Nz = s->l_l[2];
#pragma omp parallel for collapse(2) schedule(static)
    for (int ii = 0; ii < Q; ii++)
        for (int i = 0; i < s->lx; i++)
            for (int j = 0; j < ly; j++)
                for (int k = 0; k < lz; k++) {
                    int jj = j, kk = k;
                    if (jj < GH) jj += Ny;
                    else if (jj >= GH + Ny) jj -= Ny;
                    if (kk < GH) kk += Nz;
                    else if (kk >= GH + Nz) kk -= Nz;
                    if (jj != j || kk != k) s->f[F(s, ii, i, j, k)] = s->f[F(s, ii, i, jj, kk)];
                }
}

void orc_unpack(orc_t *s)
{
    if (s->c.case_id == ORC_TGV) {
        box snd[26], rcv[26];
        tgv_halo_boxes(s, snd, rcv);
        for (int n = 0; n < 26; n++) {
            int m = TGV_UNPACK_ORDER[n] - 1;
            buf_to_box(s, rcv[m], s->buf[m]);
        }
    } else if (s->c.case_id == ORC_LDC && s->c.mode == ORC_FAITHFUL) {
        /* lid_driven_cavity/src/mpi_view_transfer.cpp:602-690 on one rank:
         * y_lo == 0 -> back plane y = l_e1; z_lo == 0 -> up plane z = l_e2;
         * x_hi == glx-1 -> x = l_s0 - 1; x_lo == 0 -> x = l_e0; each from a
         * receive buffer nothing wrote (defined as zeros). */
        fill_plane_zero_y(s, s->l_e[1]);
        fill_plane_zero_z(s, s->l_e[2]);
        fill_plane_zero_x(s, s->l_s[0] - 1);
        fill_plane_zero_x(s, s->l_e[0]);
    } else if (s->c.case_id == ORC_SPHERE && s->c.mode == ORC_CONFIG) {
        periodic_fill_yz(s);
    }
}

/* ---------------------------------------------------------------- Streaming */

static void stream_pull(orc_t *s)
{
    /* "stream1" then "stream2" (tgv/src/lbm.cpp:243-253) */
    const int skip_solid = (s->c.case_id == ORC_SPHERE && s->c.mode == ORC_CONFIG);
#pragma omp parallel for collapse(2) schedule(static)
    for (int ii = 0; ii < Q; ii++)
        for (int i = GH; i < s->lx - GH; i++)
            for (int j = GH; j < s->ly - GH; j++)
                for (int k = GH; k < s->lz - GH; k++)
                    s->ft[F(s, ii, i, j, k)] = s->f[F(s, ii, i - E[ii][0], j - E[ii][1], k - E[ii][2])];
#pragma omp parallel for collapse(2) schedule(static)
    for (int ii = 0; ii < Q; ii++)
        for (int i = GH; i < s->lx - GH; i++)
            for (int j = GH; j < s->ly - GH; j++)
                for (int k = GH; k < s->lz - GH; k++) {
                    if (skip_solid && s->usr[C(s, i, j, k)] == 0) continue;
                    s->f[F(s, ii, i, j, k)] = s->ft[F(s, ii, i, j, k)];
                }
}

/* Faithful wall kernels (lid_driven_cavity/src/lbm.cpp:233-301,
 * sphere/src/lbm.cpp:277-324).  Each reads one plane and writes another, so
 * the in-kernel order does not matter; the kernel order does. */
static void bc_x_lo(orc_t *s)
{
    const int ls0 = s->l_s[0];
    for (int ii = 0; ii < Q; ii++)
        if (E[ii][0] > 0)
            for (int j = GH - 1; j < s->ly - GH + 1; j++)
                for (int k = GH - 1; k < s->lz - GH + 1; k++)
                    s->f[F(s, ii, ls0 - 1, j, k)] = s->f[F(s, BB[ii], ls0 + 1, j + 2 * E[ii][1], k + 2 * E[ii][2])];
}
static void bc_x_hi(orc_t *s)
{
    const int le0 = s->l_e[0];
    for (int ii = 0; ii < Q; ii++)
        if (E[ii][0] < 0)
            for (int j = GH - 1; j < s->ly - GH + 1; j++)
                for (int k = GH - 1; k < s->lz - GH + 1; k++)
                    s->f[F(s, ii, le0, j, k)] = s->f[F(s, BB[ii], le0 - 2, j + 2 * E[ii][1], k + 2 * E[ii][2])];
}
static void bc_y_lo(orc_t *s)
{
    const int ls1 = s->l_s[1];
    for (int ii = 0; ii < Q; ii++)
        if (E[ii][1] > 0)
            for (int i = GH - 1; i < s->lx - GH + 1; i++)
                for (int k = GH - 1; k < s->lz - GH + 1; k++)
                    s->f[F(s, ii, i, ls1 - 1, k)] = s->f[F(s, BB[ii], i + 2 * E[ii][0], ls1 + 1, k + 2 * E[ii][2])];
}
static void bc_y_hi(orc_t *s)
{
    const int le1 = s->l_e[1];
    for (int ii = 0; ii < Q; ii++)
        if (E[ii][1] < 0)
            for (int i = GH - 1; i < s->lx - GH + 1; i++)
                for (int k = GH - 1; k < s->lz - GH + 1; k++)
                    s->f[F(s, ii, i, le1, k)] = s->f[F(s, BB[ii], i + 2 * E[ii][0], le1 - 2, k + 2 * E[ii][2])];
}
static void bc_z_lo(orc_t *s)
{
    const int ls2 = s->l_s[2];
    for (int ii = 0; ii < Q; ii++)
        if (E[ii][2] > 0)
            for (int i = GH - 1; i < s->lx - GH + 1; i++)
                for (int j = GH - 1; j < s->ly - GH + 1; j++)
                    s->f[F(s, ii, i, j, ls2 - 1)] = s->f[F(s, BB[ii], i + 2 * E[ii][0], j + 2 * E[ii][1], ls2 + 1)];
}
/* ystep = 1 for the cavity (lid_driven_cavity/src/lbm.cpp:298), 2 for the
 * sphere (sphere/src/lbm.cpp:321). */
static void bc_z_hi(orc_t *s, int ystep)
{
    const int le2 = s->l_e[2];
    for (int ii = 0; ii < Q; ii++)
        if (E[ii][2] < 0)
            for (int i = GH - 1; i < s->lx - GH + 1; i++)
                for (int j = GH - 1; j < s->ly - GH + 1; j++)
                    s->f[F(s, ii, i, j, le2)] = s->f[F(s, BB[ii], i + 2 * E[ii][0], j + ystep * E[ii][1], le2 - 2)];
}

/* Sphere inlet and outlet ghost copies (sphere/src/lbm.cpp:326-347). */
static void bc_sphere_inlet(orc_t *s)
{
    for (int ii = 0; ii < Q; ii++)
        if (E[ii][0] > 0)
            for (int j = GH - 1; j < s->ly - GH + 1; j++)
                for (int k = GH - 1; k < s->lz - GH + 1; k++)
                    s->f[F(s, ii, GH - 1, j, k)] = s->f[F(s, ii, GH, j, k)];
}
static void bc_sphere_outlet(orc_t *s)
{
    const int le0 = s->l_e[0];
    for (int ii = 0; ii < Q; ii++)
        if (E[ii][0] < 0)
            for (int j = GH - 1; j < s->ly - GH + 1; j++)
                for (int k = GH - 1; k < s->lz - GH + 1; k++)
                    s->f[F(s, ii, le0, j, k)] = s->f[F(s, ii, le0 + E[ii][0], j + E[ii][1], k + E[ii][2])];
}
/* "usrbc" (sphere/src/lbm.cpp:350-356): solid s with fluid s + e_i takes
 * f_i(s) = f_bb(i)(s + 2 e_i).  No location written is read (for a convex
 * mask, s and s + 2e solid imply s + e solid), so the order is immaterial. */
static void bc_sphere_usr(orc_t *s)
{
    for (int ii = 0; ii < Q; ii++)
        for (int i = GH; i < s->lx - GH; i++)
            for (int j = GH; j < s->ly - GH; j++)
                for (int k = GH; k < s->lz - GH; k++)
                    if (s->usr[C(s, i, j, k)] == 0 &&
                        s->usr[C(s, i + E[ii][0], j + E[ii][1], k + E[ii][2])] == 1)
                        s->f[F(s, ii, i, j, k)] =
                            s->f[F(s, BB[ii], i + 2 * E[ii][0], j + 2 * E[ii][1], k + 2 * E[ii][2])];
}

static int inside(const orc_t *s, int i, int j, int k)
{
    return i >= s->l_s[0] && i < s->l_e[0] && j >= s->l_s[1] && j < s->l_e[1] && k >= s->l_s[2] &&
           k < s->l_e[2];
}

/* Config-mode cavity: halfway bounce-back written as a ghost fill.  Every
 * depth-1 ghost g and direction i with g + e_i inside takes
 * f_i(g) = f*_bb(i)(g + e_i), plus 6 w_i rho0 (e_i . u_lid) when g lies above
 * the lid plane. */
static void bc_cavity_halfway(orc_t *s)
{
    for (int i = GH - 1; i <= s->l_e[0]; i++)
        for (int j = GH - 1; j <= s->l_e[1]; j++)
            for (int k = GH - 1; k <= s->l_e[2]; k++) {
                if (inside(s, i, j, k)) continue;
                for (int ii = 0; ii < Q; ii++) {
                    int xi = i + E[ii][0], xj = j + E[ii][1], xk = k + E[ii][2];
                    if (!inside(s, xi, xj, xk)) continue;
                    double v = s->f[F(s, BB[ii], xi, xj, xk)];
                    if (k == s->l_e[2]) v = v + s->lidcorr[ii];
                    s->f[F(s, ii, i, j, k)] = v;
                }
            }
}

/* Config-mode sphere: equilibrium inlet ghost plane, zero-gradient outlet
 * ghost plane, halfway bounce-back on the mask as a fill of the solid cells
 * next to fluid.  y/z ghosts were filled periodically by unpack. */
static void bc_sphere_config(orc_t *s)
{
    const int ls0 = s->l_s[0], le0 = s->l_e[0];
    for (int ii = 0; ii < Q; ii++)
        for (int j = 0; j < s->ly; j++)
            for (int k = 0; k < s->lz; k++) {
                s->f[F(s, ii, ls0 - 1, j, k)] = s->feq_in[ii];
                s->f[F(s, ii, le0, j, k)] = s->f[F(s, ii, le0 - 1, j, k)];
            }
    for (int i = GH; i < s->lx - GH; i++)
        for (int j = GH; j < s->ly - GH; j++)
            for (int k = GH; k < s->lz - GH; k++) {
                if (s->usr[C(s, i, j, k)] != 0) continue;
                for (int ii = 0; ii < Q; ii++) {
                    int xi = i + E[ii][0], xj = j + E[ii][1], xk = k + E[ii][2];
                    if (s->usr[C(s, xi, xj, xk)] == 1)
                        s->f[F(s, ii, i, j, k)] = s->f[F(s, BB[ii], xi, xj, xk)];
                }
            }
}

void orc_streaming(orc_t *s)
{
    if (s->c.case_id == ORC_LDC) {
        if (s->c.mode == ORC_FAITHFUL) {
            /* order of lid_driven_cavity/src/lbm.cpp:233-301 */
            bc_x_lo(s);
            bc_x_hi(s);
            bc_y_lo(s);
            bc_y_hi(s);
            bc_z_lo(s);
            bc_z_hi(s, 1);
        } else {
            bc_cavity_halfway(s);
        }
    } else if (s->c.case_id == ORC_SPHERE) {
        if (s->c.mode == ORC_FAITHFUL) {
            /* order of sphere/src/lbm.cpp:277-356 */
            bc_y_lo(s);
            bc_y_hi(s);
            bc_z_lo(s);
            bc_z_hi(s, 2);
            bc_sphere_inlet(s);
            bc_sphere_outlet(s);
            bc_sphere_usr(s);
        } else {
            bc_sphere_config(s);
        }
    }
    stream_pull(s);
}

/* ---------------------------------------------------------------- Update */

static inline void moments_at(orc_t *s, int i, int j, int k)
{
    size_t c = C(s, i, j, k);
    double pm = 0.0, um = 0.0, vm = 0.0, wm = 0.0;
    for (int ii = 0; ii < Q; ii++) {
        double fv = s->f[F(s, ii, i, j, k)];
        pm = pm + fv / 3.0;
        um = um + fv * (double)E[ii][0];
        vm = vm + fv * (double)E[ii][1];
        wm = wm + fv * (double)E[ii][2];
    }
    s->p[c] = pm;
    s->ua[c] = um;
    s->va[c] = vm;
    s->wa[c] = wm;
}

void orc_update(orc_t *s)
{
    if (s->c.case_id == ORC_TGV) {
        /* tgv/src/lbm.cpp:292-330: every cell, ghosts included */
#pragma omp parallel for collapse(2) schedule(static)
        for (int i = 0; i < s->lx; i++)
            for (int j = 0; j < s->ly; j++)
                for (int k = 0; k < s->lz; k++) moments_at(s, i, j, k);
        return;
    }
    /* cavity and sphere: zero everything, then the interior loop
     * (lid_driven_cavity/src/lbm.cpp:321-347, sphere/src/lbm.cpp:376-415) */
    size_t ncell = (size_t)s->lx * s->ly * s->lz;
    for (size_t c = 0; c < ncell; c++) s->ua[c] = s->va[c] = s->wa[c] = s->p[c] = 0.0;
#pragma omp parallel for collapse(2) schedule(static)
    for (int i = GH; i < s->lx - GH; i++)
        for (int j = GH; j < s->ly - GH; j++)
            for (int k = GH; k < s->lz - GH; k++) {
                moments_at(s, i, j, k);
                if (s->c.case_id == ORC_SPHERE) {
                    size_t c = C(s, i, j, k);
                    double m = (double)s->usr[c];
                    s->ua[c] = s->ua[c] * m;
                    s->va[c] = s->va[c] * m;
                    s->wa[c] = s->wa[c] * m;
                    s->p[c] = s->p[c] * m;
                }
            }
    if (s->c.mode != ORC_FAITHFUL) return;
    if (s->c.case_id == ORC_LDC) {
        /* lid_driven_cavity/src/lbm.cpp:343-344: top fluid layer ua = 0.1 */
        for (int i = GH; i < s->lx - GH; i++)
            for (int j = GH; j < s->ly - GH; j++) s->ua[C(s, i, j, s->l_e[2] - 1)] = 0.1;
    } else if (s->c.case_id == ORC_SPHERE) {
        /* sphere/src/lbm.cpp:399-412: parabolic inlet at i = ghost, p = 0 at
         * the last interior x layer */
        const int glz = s->glz;
        for (int j = GH; j < s->ly - GH; j++)
            for (int k = GH; k < s->lz - GH; k++) {
                int kz = s->z_lo + k - GH;
                double gm = (double)(glz - 1);
                s->ua[C(s, GH, j, k)] = s->c.u0 * 4.0 * (double)kz * (double)(glz - 1 - kz) / (gm * gm);
                s->va[C(s, GH, j, k)] = 0.0;
                s->wa[C(s, GH, j, k)] = 0.0;
                s->p[C(s, s->lx - GH - 1, j, k)] = 0.0;
            }
    }
}

/* One time step in the driver's order (tgv/main.cpp:70-119). */
void orc_run(orc_t *s, int nsteps)
{
    for (int it = 0; it < nsteps; it++) {
        orc_collision(s);
        orc_pack(s);
        orc_exchange(s);
        orc_unpack(s);
        orc_streaming(s);
        orc_update(s);
    }
}

/* ------------------------------------------------------------ data access */

/* Interior moments in [z][y][x] order (x fastest). */
void orc_get_moments(const orc_t *s, double *p, double *u, double *v, double *w)
{
    const int nx = s->l_l[0], ny = s->l_l[1], nz = s->l_l[2];
    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                size_t c = C(s, x + GH, y + GH, z + GH);
                size_t o = ((size_t)z * ny + y) * nx + x;
                if (p) p[o] = s->p[c];
                if (u) u[o] = s->ua[c];
                if (v) v[o] = s->va[c];
                if (w) w[o] = s->wa[c];
            }
}

/* Interior populations as f[ii][z][y][x]. */
void orc_get_f(const orc_t *s, double *out)
{
    const int nx = s->l_l[0], ny = s->l_l[1], nz = s->l_l[2];
    for (int ii = 0; ii < Q; ii++)
        for (int z = 0; z < nz; z++)
            for (int y = 0; y < ny; y++)
                for (int x = 0; x < nx; x++)
                    out[(((size_t)ii * nz + z) * ny + y) * nx + x] = s->f[F(s, ii, x + GH, y + GH, z + GH)];
}

void orc_set_f(orc_t *s, const double *in)
{
    const int nx = s->l_l[0], ny = s->l_l[1], nz = s->l_l[2];
    for (int ii = 0; ii < Q; ii++)
        for (int z = 0; z < nz; z++)
            for (int y = 0; y < ny; y++)
                for (int x = 0; x < nx; x++)
                    s->f[F(s, ii, x + GH, y + GH, z + GH)] = in[(((size_t)ii * nz + z) * ny + y) * nx + x];
}

/* Whole arrays in the oracle's own layout (ghosts included). */
void orc_get_f_full(const orc_t *s, double *out)
{
    memcpy(out, s->f, sizeof(double) * Q * (size_t)s->lx * s->ly * s->lz);
}
void orc_set_f_full(orc_t *s, const double *in)
{
    memcpy(s->f, in, sizeof(double) * Q * (size_t)s->lx * s->ly * s->lz);
}
void orc_get_fields_full(const orc_t *s, double *p, double *u, double *v, double *w, int *usr)
{
    size_t n = (size_t)s->lx * s->ly * s->lz;
    if (p) memcpy(p, s->p, n * sizeof(double));
    if (u) memcpy(u, s->ua, n * sizeof(double));
    if (v) memcpy(v, s->va, n * sizeof(double));
    if (w) memcpy(w, s->wa, n * sizeof(double));
    if (usr) memcpy(usr, s->usr, n * sizeof(int));
}
void orc_set_fields_full(orc_t *s, const double *p, const double *u, const double *v, const double *w)
{
    size_t n = (size_t)s->lx * s->ly * s->lz;
    if (p) memcpy(s->p, p, n * sizeof(double));
    if (u) memcpy(s->ua, u, n * sizeof(double));
    if (v) memcpy(s->va, v, n * sizeof(double));
    if (w) memcpy(s->wa, w, n * sizeof(double));
}

/* FP64 diagnostics over interior fluid cells: kinetic energy 0.5 sum |u|^2 and
 * mass sum 3p (= sum f).  Sequential sum in z, y, x order. */
void orc_diagnostics(const orc_t *s, double *ke, double *mass)
{
    double k2 = 0.0, m = 0.0;
    for (int z = GH; z < s->l_e[2]; z++)
        for (int y = GH; y < s->l_e[1]; y++)
            for (int x = GH; x < s->l_e[0]; x++) {
                size_t c = C(s, x, y, z);
                if (s->usr[c] == 0) continue;
                k2 += s->ua[c] * s->ua[c] + s->va[c] * s->va[c] + s->wa[c] * s->wa[c];
                m += 3.0 * s->p[c];
            }
    *ke = 0.5 * k2;
    *mass = m;
}

/* Config-mode sphere drag proxy: momentum handed to the body by halfway
 * bounce-back, F = sum over fluid x and directions i with x + e_i solid of
 * 2 f*_i(x) e_i, taken from the post-collision populations (call after
 * orc_collision). */
void orc_sphere_force(const orc_t *s, double *fout)
{
    double fx = 0, fy = 0, fz = 0;
    for (int i = GH; i < s->l_e[0]; i++)
        for (int j = GH; j < s->l_e[1]; j++)
            for (int k = GH; k < s->l_e[2]; k++) {
                if (s->usr[C(s, i, j, k)] == 0) continue;
                for (int ii = 1; ii < Q; ii++) {
                    int xi = i + E[ii][0], xj = j + E[ii][1], xk = k + E[ii][2];
                    if (s->usr[C(s, xi, xj, xk)] != 0) continue;
                    double v = 2.0 * s->f[F(s, ii, i, j, k)];
                    fx += v * (double)E[ii][0];
                    fy += v * (double)E[ii][1];
                    fz += v * (double)E[ii][2];
                }
            }
    fout[0] = fx;
    fout[1] = fy;
    fout[2] = fz;
}

/* Lattice tables for the tests. */
void orc_tables(int *e, int *bb, double *w)
{
    init_tables();
    for (int ii = 0; ii < Q; ii++) {
        e[3 * ii] = E[ii][0];
        e[3 * ii + 1] = E[ii][1];
        e[3 * ii + 2] = E[ii][2];
        bb[ii] = BB[ii];
        w[ii] = T[ii];
    }
}
