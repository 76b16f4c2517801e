/*
 * imexlb.h -- C-ABI of the MI355X D3Q27 BGK time step.
 *
 * One handle = one rank's subdomain of the reference's struct LBM
 * (tgv/src/lbm.hpp:119-252).  The entry points replace the reference's
 * methods one for one; the reference's driver loop
 *   Collision -> pack -> exchange -> unpack -> Streaming -> Update
 * (tgv/main.cpp:70-119) maps onto them unchanged.  Plain C types only: device
 * memory is owned by the handle, host data moves through caller-owned arrays.
 *
 * Every function returns an int status (IMEXLB_OK = 0, negative on error);
 * imexlb_last_error() gives the message of the last failure on a handle.
 * A handle is not thread safe; calls on one handle enqueue in order on the
 * handle's own HIP stream.
 */
#ifndef IMEXLB_H
#define IMEXLB_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IMEXLB_Q 27
#define IMEXLB_VERSION 1

/* status codes */
#define IMEXLB_OK 0
#define IMEXLB_EINVAL (-1)       /* bad argument or configuration */
#define IMEXLB_EHIP (-2)         /* HIP runtime error */
#define IMEXLB_ENOMEM (-3)       /* device or host allocation failed */
#define IMEXLB_EUNSUPPORTED (-4) /* valid, but not offered by this engine */
#define IMEXLB_ESTATE (-5)       /* call out of order (e.g. before Initialize) */
#define IMEXLB_EIO (-6)          /* input.in missing or malformed */
#define IMEXLB_ENODEV (-7)       /* no HIP device */

/* flow cases: the reference's three applications */
#define IMEXLB_CASE_TGV 0    /* tgv/ : periodic Taylor-Green vortex */
#define IMEXLB_CASE_CAVITY 1 /* lid_driven_cavity/ */
#define IMEXLB_CASE_SPHERE 2 /* sphere/ : channel with a voxel sphere */

/* boundary semantics */
#define IMEXLB_MODE_FAITHFUL 0 /* the reference's own boundary code */
#define IMEXLB_MODE_CONFIG 1   /* halfway bounce-back, lid momentum term,
                                  equilibrium inlet, zero-gradient outlet */

/* engines */
#define IMEXLB_ENGINE_FUSED 0  /* one kernel per step: stream+moments+collide
                                  (TGV, config modes, and the faithful cavity and
                                  sphere in FP64 pull storage, whose first ghost
                                  layer the reference's boundary kernels fill
                                  first) */
#define IMEXLB_ENGINE_PHASED 1 /* the reference's phases as separate kernels on
                                  its ghost-3 layout (FP64, all cases, faithful) */

/* population storage */
#define IMEXLB_F64 0 /* FP64, FP64 collision: drop-in */
#define IMEXLB_F32 1 /* FP32 of f - w_i S, FP32 collision */
#define IMEXLB_F16 2 /* FP16 of (f - w_i S) * 2^k_i, FP32 collision */

#define IMEXLB_PULL 0 /* two lattices, pull streaming */
#define IMEXLB_AA 1   /* one lattice, AA pattern in place */

typedef struct imexlb_config {
    int32_t case_id;  /* IMEXLB_CASE_* */
    int32_t mode;     /* IMEXLB_MODE_* */
    int32_t engine;   /* IMEXLB_ENGINE_* */
    int32_t dtype;    /* IMEXLB_F64 / F32 / F16 */
    int32_t storage;  /* IMEXLB_PULL / IMEXLB_AA */
    int32_t nx, ny, nz; /* cells of this subdomain (the reference's glx, gly, glz) */
    double tau0;      /* System::tau; BGK rate 1 / (tau0 + 0.5) */
    double rho0;      /* System::rho0 */
    double u0;        /* System::u0: TGV amplitude, channel inlet speed */
    double lid_u;     /* config-mode cavity lid speed (faithful uses 0.1) */
    double sphere_cx, sphere_cy, sphere_cz, sphere_r; /* config-mode sphere */
    uint64_t seed;    /* faithful-mode sphere inlet perturbation */
    int32_t block;    /* threads per workgroup, 0 = default */
    int32_t cells_per_thread; /* 0 = default; 1, 2 (blockDim apart) or 3 (two
                                 adjacent x cells; periodic lattices with nx
                                 even, one cell per thread elsewhere) */
    int32_t xcd_remap;        /* -1 = default, 0 off, 1 on */
    int32_t device;           /* HIP device ordinal */
    int32_t stride_pad;       /* bytes added between populations, multiple of
                                 256; -1 = default.  Keeps the 27 population
                                 streams off a common power-of-two alignment */
    int32_t nt_loads;         /* -1 = default, 0/1: non-temporal loads of the
                                 lines no other wavefront reads */
    int32_t addr_mode;        /* fused kernels' addressing, the same results in
                                 every mode: -1 = default (the most compact the
                                 lattice allows), or at most 2 (a uniform base per
                                 group of populations; up to 2^28 cells per
                                 population), 1 (a base per population; up to
                                 4 GiB per population), 0 (64-bit offsets) */
} imexlb_config;

typedef struct imexlb_handle imexlb_handle;

typedef struct imexlb_info {
    int64_t cells;          /* interior cells */
    int64_t steps;          /* steps taken since Initialize */
    double bytes_per_lup;   /* algorithmic HBM bytes per cell update */
    int64_t device_bytes;   /* device memory held by the handle */
    int32_t block, cells_per_thread, xcd_remap, nt_loads; /* launch shape; cells per
                               thread as the fused kernels run it (1 where the
                               lattice does not allow the requested 2 or 3) */
    int32_t addr_mode;      /* the fused kernels' address mode; -1 phased engine */
} imexlb_info;

/* Defaults: TGV, faithful, fused, FP64, pull, 64^3, tau0 = 0.0153, rho0 = 1,
 * u0 = 0.05, lid 0.1, seed 5374857 (sphere/src/lbm.cpp:223). */
int imexlb_config_init(imexlb_config *cfg);

/* LBM::LBM(MPI_Comm, sx, sy, sz, tau, rho0, u0) on one rank
 * (tgv/src/lbm.hpp:161-236): validates and allocates. */
int imexlb_create(const imexlb_config *cfg, imexlb_handle **out);
int imexlb_destroy(imexlb_handle *h);

/* LBM::Initialize (tgv/src/lbm.cpp:7-221; lid_driven_cavity/src/lbm.cpp:
 * 3-213; sphere/src/lbm.cpp:36-255).  May be called again to reset. */
int imexlb_initialize(imexlb_handle *h);
/* LBM::setup_subdomain (tgv/src/mpi_view_transfer.cpp:2-115): nothing to
 * stage on one GPU; kept for the call sequence. */
int imexlb_setup_subdomain(imexlb_handle *h);

/* The per-phase surface (tgv/src/lbm.hpp:238-246).
 *   fused engine : Collision, pack, exchange, unpack are ordering markers;
 *                  Streaming runs the fused step and writes the moments;
 *                  Update publishes them.
 *   phased engine: each is its own kernel(s), as in the reference. */
int imexlb_collide(imexlb_handle *h);  /* LBM::Collision, tgv/src/lbm.cpp:222 */
int imexlb_pack(imexlb_handle *h);     /* LBM::pack, tgv/src/mpi_view_transfer.cpp:117 */
int imexlb_exchange(imexlb_handle *h); /* LBM::exchange, :176 */
int imexlb_unpack(imexlb_handle *h);   /* LBM::unpack, :314 */
int imexlb_stream(imexlb_handle *h);   /* LBM::Streaming, tgv/src/lbm.cpp:239 */
int imexlb_update(imexlb_handle *h);   /* LBM::Update, tgv/src/lbm.cpp:288 */

/* n whole steps (the six calls above, n times).  Moments are written on every
 * step whose index is a multiple of moments_every (> 0) and on the last step
 * of the call; moments_every = 0 writes none.  Fused engine: runs of 16
 * moment-free steps replay a captured HIP graph (same results).  Phased
 * engine: per step, the halo and boundary kernels, then one pass of
 * streaming, Update and the next Collision (same results as the six calls);
 * it writes the moments of the last step whatever moments_every is, as the
 * reference's Update leaves them after every step. */
int imexlb_step(imexlb_handle *h, int nsteps, int moments_every);
/* As imexlb_step, bracketed by HIP events on the handle's stream; *ms is the
 * device time of the n launches. */
int imexlb_step_timed(imexlb_handle *h, int nsteps, int moments_every, float *ms);

/* Moments p, u, v, w (the reference's p, ua, va, wa) of the last step that
 * wrote them, interior only, [z][y][x] with x fastest.  NULL skips a field. */
int imexlb_get_moments(imexlb_handle *h, double *p, double *u, double *v, double *w);
/* The same fields as density and velocity: rho = 3 p = sum f (the model is
 * incompressible, tgv/src/lbm.cpp:307-325 sums p = sum f / 3), u, v, w. */
int imexlb_get_macro(imexlb_handle *h, double *rho, double *u, double *v, double *w);
/* Populations, interior, [i][z][y][x], FP64.  Fused engine: post-collision
 * populations of the last step (the engine's state; set_f restores it).
 * Phased engine: the reference's f array at this point of the step. */
int imexlb_get_f(imexlb_handle *h, double *f);
int imexlb_set_f(imexlb_handle *h, const double *f);

/* Restart file (no counterpart in the reference, which writes only the
 * macroscopic fields).  Fused engine: "IMXLBCK1", the creating
 * configuration, the step count, then the post-collision populations
 * [i][z][y][x] as FP64; a handle of the same case, mode and extents reads it
 * (storage type and precision may differ).  Phased engine: "IMXLBPH1", the
 * configuration, the step count, then the engine's whole state between steps
 * as FP64 -- f, p, ua, va, wa over the ghosted box, since the faithful wall
 * rules read ghost values written in earlier steps; a phased handle of the
 * same case, mode and extents reads it.  The fused engine's faithful
 * cavity and sphere: "IMXLBCR1", as "IMXLBCK1" followed by the ghost ring
 * (27 populations over the six first-layer ghost faces, FP64).  Every
 * reader must also have the writer's tau0 and rho0, and where the case uses
 * them its lid speed (config cavity), inlet speed u0 (sphere), ball centre and
 * radius (config sphere) and seed (faithful sphere); a file written with other
 * physics is refused with IMEXLB_EINVAL.  Each kind of handle refuses the
 * other kinds' files.  All resume at the saved step. */
int imexlb_checkpoint_write(imexlb_handle *h, const char *path);
int imexlb_checkpoint_read(imexlb_handle *h, const char *path);

/* FP64 kinetic energy 0.5 sum |u|^2 and mass sum 3p over the fluid cells of
 * the last moments written. */
int imexlb_diagnostics(imexlb_handle *h, double *ke, double *mass);
/* Config-mode sphere: momentum-exchange force on the solid from the current
 * post-collision populations (drag proxy), FP64. */
int imexlb_sphere_force(imexlb_handle *h, double *force3);

/* Allocation-free form of imexlb_diagnostics.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory -- code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc.  It takes no handle: p, u, v, w are caller-owned device arrays
 * of ncell FP64 values each (as the moments of imexlb_get_moments), and
 * ke_mass is a caller-owned device array of 2 doubles that receives
 * 0.5 sum |u|^2 and sum 3p -- the same bits imexlb_diagnostics returns for
 * the same fields.  It enqueues two kernels on `stream` (a hipStream_t; NULL
 * = the null stream) of the current device and returns without waiting;
 * nothing it does allocates, synchronises or clears memory, so it may be
 * captured into a HIP graph.  Scratch: a fixed 16 KiB device array held by
 * the library on every GPU it runs on, until the process ends; calls that
 * overlap on one GPU (on different streams, or replays of graphs holding
 * it) race on it and must be ordered by the caller.  The first call on a
 * device looks the array's address up (make it outside a capture if the
 * capture must hold kernels only).  ncell in [0, IMEXLB_NOALLOC_MAX_CELLS]
 * (2^31), else IMEXLB_EINVAL; imexlb_last_error(NULL) gives the message. */
#define IMEXLB_NOALLOC_MAX_CELLS 2147483648LL
int imexlb_diagnostics_noalloc(const double *p, const double *u, const double *v, const double *w,
                               int64_t ncell, double *ke_mass, void *stream);

int imexlb_sync(imexlb_handle *h);
int imexlb_get_info(imexlb_handle *h, imexlb_info *info);
/* Message of the last failure on h; with h = NULL, of the last failed
 * imexlb_create or handle-free call on this thread. */
const char *imexlb_last_error(const imexlb_handle *h);
const char *imexlb_strerror(int status);

/* Device copy of `bytes` bytes (read + write), `reps` times: measured
 * bandwidth ceiling (GB/s counted as read + write bytes / time). */
int imexlb_copy_bandwidth(int device, size_t bytes, int reps, double *gbps, float *ms_per_launch);
/* Bandwidth ceiling probes (no counterpart in the reference), over `bytes`
 * bytes per stream, `reps` timed launches after one warm-up:
 *   0 copy, one 16-byte vector per lane    1 the same, non-temporal
 *   2 copy, four 16-byte vectors per lane  3 the same, non-temporal
 *   4 copy, grid-stride over 8 workgroups per CU   5 the same, non-temporal
 *   6 read only, four vectors per lane     7 the same, non-temporal
 *   8 write only, four vectors per lane    9 the same, non-temporal
 * *gbps counts every byte read and written. */
int imexlb_bandwidth_probe(int device, int kind, size_t bytes, int reps, double *gbps, float *ms_per_launch);

/* Launch-shape autotuner (no counterpart in the reference).  Times every
 * workgroup size x cells per thread (1, 2, and 3 where the lattice allows
 * cell pairs) x XCD tile order x load variant on this
 * handle for `steps` steps each, times the four fastest again (twice each,
 * interleaved, 2 x steps steps), keeps the fastest sum and appends it to the
 * JSON-lines cache at cache_path (NULL: no cache; a cache hit skips the
 * timing).  All shapes give bit-identical results.  The handle is
 * re-initialised after tuning.  chosen[4] (may be NULL) receives block, cells
 * per thread, XCD order, non-temporal loads.  Fused engine only. */
int imexlb_autotune(imexlb_handle *h, const char *cache_path, int steps, int32_t *chosen);
/* Set the launch shape by hand (block 64..512 step 64, cells per thread 1|2|3
 * as imexlb_config.cells_per_thread, XCD order 0|1, non-temporal loads 0|1);
 * results do not depend on it. */
int imexlb_set_launch(imexlb_handle *h, int block, int cells_per_thread, int xcd_order, int nt_loads);
/* Lattice placement search (no counterpart in the reference).  The step rate
 * of a lattice depends by several percent on where its pages land in HBM,
 * and stays fixed for the life of the allocation.  This allocates up to
 * trials - 1 further lattices while device memory allows (2 GiB stay free),
 * keeping each candidate until the end so that every one has its own pages,
 * times `steps` steps on each with the current launch shape, keeps the
 * fastest and frees the others.  ms_per_step (may be NULL; `trials`
 * entries) receives each candidate's time per step, 0 where none was
 * allocated; *chosen (may be NULL) the index kept, 0 being the original.
 * The state is re-initialised.  Results do not depend on the placement.
 * Fused engine only; trials in [1, 16]. */
int imexlb_tune_placement(imexlb_handle *h, int trials, int steps, float *ms_per_step, int32_t *chosen);
/* Placement search inside one address range (no counterpart in the
 * reference).  The step rate follows the lattice's base address smoothly
 * (rising by up to 20 % over a few GiB), so this maps the lattice(s) plus
 * slack_bytes once, in chunks of physical memory of about step_bytes, times
 * `steps` steps with the lattice at each offset k * step_bytes (k = 0 ..
 * slack / step; step a multiple of 2 MiB dividing the slack), keeps the
 * fastest offset and frees the handle's previous lattice.  The chunks the
 * kept lattice does not touch are then given back: the handle holds at most
 * two chunks beyond its lattice(s).  ms_per_pos (npos_max entries, may be NULL)
 * receives each offset's time per step, *chosen (may be NULL) the index
 * kept.  The state is re-initialised; results do not depend on the offset.
 * Fused engine only; at most 256 offsets. */
int imexlb_tune_placement_window(imexlb_handle *h, int64_t slack_bytes, int64_t step_bytes, int steps,
                                 float *ms_per_pos, int32_t npos_max, int32_t *chosen);

/* ---- batched ensemble (no counterpart in the reference) ----------------
 * Many independent domains -- TGV boxes and config-mode cavities of uneven
 * size and Reynolds number -- advanced in lockstep by one persistent launch
 * per step over a table of tiles, optionally replayed from a captured HIP
 * graph.  Each member steps bit for bit as the fused engine steps it alone. */
typedef struct imexlb_domain {
    int32_t case_id;          /* IMEXLB_CASE_TGV or IMEXLB_CASE_CAVITY (config mode) */
    int32_t nx, ny, nz;
    double tau0, rho0, u0, lid_u;
} imexlb_domain;
typedef struct imexlb_ensemble imexlb_ensemble;

#define IMEXLB_ENS_PERSISTENT 0 /* one persistent launch per step */
#define IMEXLB_ENS_GRAPH 1      /* persistent launches replayed from a HIP graph */
#define IMEXLB_ENS_PER_DOMAIN 2 /* one single-domain launch per domain per step */

/* n domains drawn from seed: alternating TGV / cavity, edge in
 * {32,48,64,80,96}, Re in [50,1000], lid speed in [0.05,0.1], TGV U0 0.05. */
int imexlb_ensemble_generate(uint64_t seed, int n, imexlb_domain *out);
/* All members' lattices share one device allocation.  IMEXLB_EINVAL for a
 * member of more than 2^28 cells (the member kernels' 32-bit offsets). */
int imexlb_ensemble_create(const imexlb_domain *doms, int n, int dtype, int storage, int device,
                           imexlb_ensemble **out);
int imexlb_ensemble_destroy(imexlb_ensemble *e);
int imexlb_ensemble_initialize(imexlb_ensemble *e);
int imexlb_ensemble_step(imexlb_ensemble *e, int nsteps, int mode, int moments_every);
int imexlb_ensemble_step_timed(imexlb_ensemble *e, int nsteps, int mode, int moments_every, float *ms);
int imexlb_ensemble_get_moments(imexlb_ensemble *e, int domain, double *p, double *u, double *v, double *w);
/* Per-domain health after a step that wrote moments: finite[d] (n entries,
 * may be NULL) is 1 while every moment of domain d is finite and 0 once the
 * domain has blown up -- BGK has a stability limit at small tau0 (high Re on
 * a coarse grid), and the seeded draw of imexlb_ensemble_generate reaches
 * past it for some small domains.  *unstable (may be NULL) receives the
 * number of zeros.  Other domains are unaffected: each steps independently.
 * IMEXLB_ESTATE if the current step wrote no moments. */
int imexlb_ensemble_status(imexlb_ensemble *e, int32_t *finite, int32_t *unstable);
/* out[7]: cells, tiles, grid, block, cells per thread, device bytes, steps */
int imexlb_ensemble_info(imexlb_ensemble *e, int64_t *out);
/* grid: workgroups of each persistent launch (each walks every grid-th tile
 * of its run); 0 = one workgroup per tile. */
int imexlb_ensemble_set_launch(imexlb_ensemble *e, int block, int cells_per_thread, int grid);
/* As imexlb_autotune, over block x cells per thread x grid (4 to 64
 * workgroups per CU, or one per tile), timed with graph replay; chosen[3] =
 * block, cells per thread, grid. */
int imexlb_ensemble_autotune(imexlb_ensemble *e, const char *cache_path, int steps, int32_t *chosen);
/* As imexlb_tune_placement, for the ensemble's lattices, which share one
 * allocation: up to trials - 1 fresh allocations of the whole set are timed
 * with graph-replayed steps, the fastest is kept and the ensemble is reset
 * to its initial state.  trials in [1, 16]. */
int imexlb_ensemble_tune_placement(imexlb_ensemble *e, int trials, int steps, float *ms_per_step, int32_t *chosen);
/* As imexlb_tune_placement_window, for the pool that holds every member's
 * lattices: the pool slides through one address range of its size plus
 * slack_bytes (graph-replayed steps timed at each offset), the fastest offset
 * is kept and the chunks the pool does not touch are given back. */
int imexlb_ensemble_tune_placement_window(imexlb_ensemble *e, int64_t slack_bytes, int64_t step_bytes, int steps,
                                          float *ms_per_pos, int32_t npos_max, int32_t *chosen);
const char *imexlb_ensemble_last_error(const imexlb_ensemble *e);

/* ---- host-side pieces of the reference, no device use ----------------- */

/* System (tgv/src/System.cpp:5-24, lid_driven_cavity/src/System.cpp:5-20,
 * sphere/src/System.cpp:5-23): parse ./input.in, three whitespace-separated
 * lines "rho R Re / u0 Time inter / nx ny nz", and derive tau per case.
 * keep_hardcoded != 0 keeps TGV's Time = 100, inter = 10 override
 * (tgv/src/System.cpp:17-18); 0 honours the file. */
typedef struct imexlb_system {
    double rho0, R, Re, u0;
    int32_t Time, inter;
    int32_t sx, sy, sz;
    double cs2, cs, Ma, miu, tau;
} imexlb_system;
int imexlb_system_read(const char *path, int case_id, int keep_hardcoded, imexlb_system *out);

/* CommHelper (tgv/src/lbm.hpp:34-117; non-periodic variant
 * lid_driven_cavity/src/lbm.hpp:33-103): rank grid and the 26 neighbours.
 * neighbours[] order: left right down up front back | frontup frontright
 * frontdown frontleft backup backright backdown backleft leftup rightup
 * leftdown rightdown | frontleftup frontleftdown frontrightup frontrightdown
 * backleftup backleftdown backrightup backrightdown.  -1 = no neighbour. */
typedef struct imexlb_comm {
    int32_t nranks, me, rx, ry, rz, px, py, pz;
    int32_t neighbours[26];
    int32_t l_l[3];          /* local extents (tgv/src/lbm.hpp:164-166) */
    int32_t lo[3], hi[3];    /* global offsets x_lo..z_hi (:187-231) */
} imexlb_comm;
int imexlb_comm_layout(int nranks, int me, int periodic, int glx, int gly, int glz, imexlb_comm *out);

/* Tecplot binary (#!TDV112) field file: variables x y z u v w p of one
 * ordered nx*ny*nz zone, FP64, x fastest -- the file LBM::MPIoutput writes
 * (tgv/src/lbm.cpp:335-669; lid_driven_cavity/src/lbm.cpp:351-674), here for
 * one rank from caller-owned host arrays (as returned by imexlb_get_moments). */
int imexlb_tecplot_write(const char *path, double solution_time, int nx, int ny, int nz, const double *p,
                         const double *u, const double *v, const double *w);

#ifdef __cplusplus
}
#endif
#endif /* IMEXLB_H */
