/*
 * ekpnp.h -- C ABI of the MI355X (gfx950) EK-PNP coupled time step.
 *
 * One opaque context owns every device buffer, the per-channel descriptor
 * tables, the FFT twiddle tables, one HIP stream and the cached hipGraphs.
 * A context advances one or many independent channels (periodic in x, walls
 * at y = 0 and y = ny-1) in lock-step; each channel has its own size and
 * physics.  All entry points return EKPNP_OK (0) or a negative status; none
 * calls exit().  A context is not thread-safe; calls on one context are
 * ordered on its stream.  Host buffers are owned by the caller.
 *
 * Entry points and the reference interface each one replaces
 * (file:line in gyf135/EK-PNP):
 *   ekpnp_create            main.cu:89-142   (cudaMalloc of populations and
 *                                             fields, cufftPlan2d, kx/ky tables)
 *   ekpnp_destroy           main.cu:278-301
 *   ekpnp_initialization    LBM.cu:95-160    (initialization / gpu_initialization)
 *   ekpnp_set_fields        LBM.cu:1350-1391 (read_data: macroscopic fields in)
 *   ekpnp_init_equilibrium  LBM.cu:183-305   (init_equilibrium)
 *   ekpnp_stream_collide_save LBM.cu:307-329 (collide, boundary, stream, bc_charge)
 *   ekpnp_fast_poisson      poisson.cu:180-208 (+ efield, poisson.cu:132-174)
 *   ekpnp_step              main.cu:190-199  (the time loop body, n times)
 *   ekpnp_get_field         LBM.cu:1272-1312 (the D2H copies of save_data_end)
 *   ekpnp_get_populations / ekpnp_set_populations
 *                           f0/f1, h0/h1, temp0/temp1 device arrays, LBM.h:124-126
 *   ekpnp_save_data_end     LBM.cu:1272-1312 (save_data_end, data_end.dat)
 *   ekpnp_read_data         LBM.cu:1350-1391 (read_data, the flag == 1 restart)
 *   ekpnp_save_tecplot      LBM.cu:1100-1269 (save_data_tecplot, data.dat)
 *   ekpnp_save_charge_data  LBM.cu:1314-1348 (save_data_dmd, charge_data)
 *   ekpnp_current           LBM.cu:1394-1417 (current)
 *   ekpnp_record_umax       LBM.cu:1419-1451 (record_umax, umax.dat)
 *   ekpnp_save_checkpoint / ekpnp_load_checkpoint
 *                           no counterpart: exact binary restart (the
 *                           reference's data_end.dat restart re-equilibrates
 *                           from 6-decimal fields and loses the populations)
 */
#ifndef EKPNP_H
#define EKPNP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
enum {
  EKPNP_OK = 0,
  EKPNP_ERR_INVALID = -1,     /* bad argument or parameter set              */
  EKPNP_ERR_HIP = -2,         /* HIP runtime reported an error              */
  EKPNP_ERR_NOMEM = -3,       /* device or host allocation failed           */
  EKPNP_ERR_STATE = -4,       /* entry point called out of order            */
  EKPNP_ERR_UNSUPPORTED = -5, /* shape outside the kernels' limits          */
  EKPNP_ERR_UNSTABLE = -6,    /* NaN / runaway found by the reductions      */
  EKPNP_ERR_IO = -7           /* file could not be opened, read or written  */
};

/* ---- population storage ------------------------------------------------ */
enum {
  EKPNP_POP_FP32 = 0, /* FP32 deviations from w_i*(rho, c+, c-)_ref(x, y)
                         (one reference per 64-node block of a row)        */
  EKPNP_POP_FP64 = 1  /* FP64 full populations (the reference's storage)   */
};

/* ---- physics / transcription flags (ekpnp_channel_params.flags) -------- */
/* Ions drift in E + E_applied; the reference drifts in the solved E only
 * (LBM.cu:479-488). */
#define EKPNP_FLAG_ION_DRIFT_APPLIED 0x1u
/* Negative-ion rest equilibrium uses the fluid |u|^2, as LBM.cu:493 does;
 * off = the drift velocity, as LBM.cu:242 does at initialisation. */
#define EKPNP_FLAG_G0_FLUID_U 0x2u

/* ---- fields ------------------------------------------------------------ */
enum {
  EKPNP_FIELD_RHO = 0,  /* r      */
  EKPNP_FIELD_UX = 1,   /* u      */
  EKPNP_FIELD_UY = 2,   /* v      */
  EKPNP_FIELD_CPOS = 3, /* c      (positive ion concentration) */
  EKPNP_FIELD_CNEG = 4, /* T/temp (negative ion concentration) */
  EKPNP_FIELD_PSI = 5,  /* phi    */
  EKPNP_FIELD_EX = 6,
  EKPNP_FIELD_EY = 7,
  EKPNP_FIELD_COUNT = 8
};

/* species for ekpnp_get_populations / ekpnp_set_populations */
enum { EKPNP_SPECIES_F = 0, EKPNP_SPECIES_H = 1, EKPNP_SPECIES_G = 2 };

/* initialisation modes for ekpnp_initialization */
enum {
  EKPNP_INIT_COLD_NOISE = 0, /* rho0, u = 0, c+- = c_bulk(1 + noise*U(-1,1))
                                inside, wall values on the walls, then one
                                Poisson solve for psi and E                 */
  EKPNP_INIT_REFERENCE = 1,  /* gpu_initialization values (LBM.cu:147-160):
                                rho0, c+- = 0, psi = psi_lo, u = E = 0      */
  EKPNP_INIT_PB_PICARD = 2   /* the reference's start: gpu_initialization, then
                                pb_iterations x { c+- = c_bulk exp(-+inv_vt psi);
                                Poisson; psi = w psi + (1-w) psi_old }
                                (LBM.cu:95-145), entirely on the device     */
};

/* indices into the per-channel record of ekpnp_reductions */
enum {
  EKPNP_RED_MASS = 0,      /* sum rho over all nodes                        */
  EKPNP_RED_CPOS = 1,      /* sum c+ over all nodes                         */
  EKPNP_RED_CNEG = 2,      /* sum c- over all nodes                         */
  EKPNP_RED_CHARGE = 3,    /* sum (c+ - c-) over interior rows              */
  EKPNP_RED_DUX2 = 4,      /* sum (ux - ux_prev)^2 over interior rows       */
  EKPNP_RED_UX2 = 5,       /* sum ux^2 over interior rows                   */
  EKPNP_RED_NONFINITE = 6, /* count of nodes with a non-finite rho, ux, uy,
                              c+, c- or psi                                  */
  EKPNP_RED_RESIDUAL = 7,  /* sqrt(DUX2/UX2)/interval (0 before a snapshot) */
  EKPNP_RED_COUNT = 8
};

/*
 * Per-channel parameters.  General units, as LBM.h:39-112 writes them; the
 * lattice-unit configurations use dx = dy = dt = cfl = 1, cs2 = 1/3.
 */
typedef struct ekpnp_channel_params {
  int32_t nx, ny;    /* NX (periodic), NY (walls at 0 and NY-1), LBM.h:35-36 */
  uint32_t flags;    /* EKPNP_FLAG_* */
  int32_t reserved0;
  double dx, dy, dt; /* LBM.h:42-43, 51 */
  double cfl;        /* dt/dx, LBM.h:50 */
  double cs2;        /* cs_square, LBM.h:52 */
  double rho0;       /* LBM.h:53 */
  double nu;         /* LBM.h:65 */
  double diff_pos;   /* diffu, LBM.h:64 */
  double diff_neg;   /* D, LBM.h:71 */
  double mob_pos;    /* K, LBM.h:67 */
  double mob_neg;    /* Kn, LBM.h:74 */
  double magic_f;    /* V,  LBM.h:110 */
  double magic_pos;  /* VC, LBM.h:111 */
  double magic_neg;  /* VT, LBM.h:112 */
  double psi_lo;     /* voltage,  LBM.h:58 */
  double psi_hi;     /* voltage2, LBM.h:60 */
  double cpos_lo;    /* charge0,  LBM.h:55 */
  double cpos_hi;    /* charge1,  LBM.h:56 */
  double cneg_lo;    /* charge0n, LBM.h:81 */
  double cneg_hi;    /* charge1n, LBM.h:83 */
  double c_bulk;     /* bulk concentration (chargeinf, LBM.h:54): cold-start
                        value and the FP32 deviation reference              */
  double poisson_coef; /* convertCtoCharge/eps (poisson.cu:229-259)         */
  double force_coef;   /* convertCtoCharge in the body force (LBM.cu:388);
                          0 reproduces the shipped, zeroed force            */
  double e_applied;    /* Ext, LBM.h:62 */
  double exf;          /* LBM.h:47 */
  double uw;           /* top-wall velocity, LBM.h:46 */
  double lx;           /* Lx = nx dx, the x period used by kx (main.cu:120-128); checked */
  double init_noise;   /* relative amplitude of the cold-start noise        */
  double inv_thermal_voltage; /* e/(kB T) in units of 1/psi: the Boltzmann
                          factor exp(-+ inv_vt psi) of the PB start (LBM.cu:143-144) */
  double pb_omega;     /* PB_omega, Picard under-relaxation (LBM.h:91)      */
  double pb_iterations; /* Picard iterations of the PB start (LBM.cu:116: 501) */
} ekpnp_channel_params;

typedef struct ekpnp_options {
  int32_t precision; /* EKPNP_POP_FP32 or EKPNP_POP_FP64 */
  int32_t device;    /* HIP device ordinal */
  int32_t use_graph; /* capture ekpnp_step chunks in hipGraphs */
  int32_t compute_fp64; /* FP32 storage only: collide in FP64 arithmetic (the
                           default, 1).  0 collides in FP32 arithmetic, which
                           does NOT hold the 1e-5 relative-L2 accuracy of the
                           default: u_x drifts ~1e-4 from FP64 over 20 000
                           steps of a 64 x 128 channel (the drift velocity is
                           a small difference of O(1) population sums); for
                           throughput comparisons only */
  int32_t reject_unresolved; /* 1: ekpnp_create fails with
                           EKPNP_ERR_UNSTABLE for a channel that
                           ekpnp_check_params flags (default 0: accepted,
                           as the reference accepts any parameters) */
} ekpnp_options;

/* ---- parameter check ------------------------------------------------------ */
/* Flags of ekpnp_check_params */
#define EKPNP_CHECK_DEBYE_UNRESOLVED 0x1u /* kappa dy > 0.4                      */
#define EKPNP_CHECK_WALL_DRIFT 0x2u       /* Gouy-Chapman wall drift of the ions
                                             > 0.05 lattice speed (dx/dt)        */
#define EKPNP_CHECK_FLOW_FAST 0x4u        /* Helmholtz-Smoluchowski flow of the
                                             applied field > 0.1 lattice speed   */
/* Resolution and stability rule for each channel, without a device:
 * kappa = sqrt(2 poisson_coef c_bulk inv_thermal_voltage) (linearised
 * Poisson-Boltzmann), the Debye length resolved (kappa dy <= 0.4) and the
 * ions' drift in the Gouy-Chapman wall field, max|K| (2 kappa / inv_vt)
 * sinh(inv_vt max|psi_wall| / 2), at most 0.05 dx/dt; and the electro-
 * osmotic flow the applied field drives (tau -> 1/2 or a huge E), the
 * Helmholtz-Smoluchowski velocity |force_coef e_applied max|psi_wall| /
 * (poisson_coef nu rho0)|, at most 0.1 dx/dt.  Channels outside the
 * rule need not fail, but strongly charged thin-layer channels past it have
 * gone non-finite (the batched sweep's sampler redraws them).  flags_out
 * (may be NULL, else n values) receives EKPNP_CHECK_* per channel; returns
 * the number of channels flagged, or EKPNP_ERR_INVALID.  No counterpart in
 * the reference. */
int ekpnp_check_params(const ekpnp_channel_params* channels, int n_channels, uint32_t* flags_out);

typedef struct ekpnp_ctx ekpnp_ctx;

/* ---- parameter helpers --------------------------------------------------- */
/* The reference's compiled-in case (LBM.h:31-112, main.cu:38-44). */
void ekpnp_reference_params(ekpnp_channel_params* p);
/* Lattice units: tau -> nu = (tau-1/2)/3, magic 1/4 for all species,
 * D+- = diff, K+- = +-diff (psi in kT/e), psi walls = zeta,
 * c+ wall = exp(-zeta), c- wall = exp(zeta), c_bulk = 1, rho0 = 1,
 * kappa = kappa_h/h with h = (ny-1)/2, poisson_coef = kappa^2/2,
 * force_coef = poisson_coef, e_applied = e_x, lx = nx, noise 1e-3. */
int ekpnp_lattice_params(ekpnp_channel_params* p, int nx, int ny, double tau,
                         double zeta, double kappa_h, double e_x, double diff);
void ekpnp_default_options(ekpnp_options* o);

/* ---- lifetime ------------------------------------------------------------ */
int ekpnp_create(const ekpnp_channel_params* channels, int n_channels,
                 const ekpnp_options* opt, ekpnp_ctx** out);
void ekpnp_destroy(ekpnp_ctx* ctx);
const char* ekpnp_last_error(const ekpnp_ctx* ctx);
/* last error of a failed ekpnp_create (no context exists then) */
const char* ekpnp_create_error(void);

/* ---- initial state -------------------------------------------------------- */
int ekpnp_initialization(ekpnp_ctx* ctx, int mode, uint64_t seed);
/* Per-x wall profiles of one channel (length nx each); the uniform values of
 * the parameter block are used until this is called.  Any pointer may be
 * NULL to keep that profile. */
int ekpnp_set_wall_profile(ekpnp_ctx* ctx, int channel, const double* psi_lo,
                           const double* psi_hi, const double* cpos_lo,
                           const double* cpos_hi, const double* cneg_lo,
                           const double* cneg_hi);
/* Macroscopic fields of one channel, [ny][nx] each.  psi may be NULL: the
 * potential is then solved from c+ - c-.  E is always -grad(psi). */
int ekpnp_set_fields(ekpnp_ctx* ctx, int channel, const double* rho,
                     const double* ux, const double* uy, const double* cpos,
                     const double* cneg, const double* psi);
int ekpnp_init_equilibrium(ekpnp_ctx* ctx);
/* Restart one channel from its starting fields (EKPNP_INIT_COLD_NOISE with
 * `seed`, or EKPNP_INIT_REFERENCE; the fields ekpnp_initialization gives it)
 * at equilibrium, leaving every other channel's state as it is: the recovery
 * for a channel that ekpnp_run_to_residual reported non-finite.  Its FP32
 * storage references are re-derived from the new fields and its residual
 * snapshot restarts; the step count and the re-centring schedule are the
 * context's.  No counterpart in the reference (its CUDA errors end the
 * process, LBM.cu:44-62).  Needs populations (after init_equilibrium). */
int ekpnp_restart_channel(ekpnp_ctx* ctx, int channel, int mode, uint64_t seed);

#define EKPNP_REBASE_INTERVAL 4096

/* ---- the coupled step ------------------------------------------------------ */
int ekpnp_stream_collide_save(ekpnp_ctx* ctx);
int ekpnp_fast_poisson(ekpnp_ctx* ctx);
/* Iterative Poisson solve of one channel (replaces poisson_phi + gpu_poisson,
 * poisson.cu:35-130): the 5-point finite-difference system in x and y,
 * periodic in x, psi(0) and psi(ny-1) the wall potentials, right-hand side
 * -pc (c+ - c-) from the current charge density.  Red-black SOR sweeps with
 * relaxation omega in (0, 2) (omega = 1: Gauss-Seidel) starting from the
 * current psi, which is overwritten.  Stops when the largest |change| of a
 * sweep drops below tol (checked every check_every sweeps; the reference
 * checks every sweep with tol 1e-9) or after max_iter sweeps (the reference
 * uses 1e6, prints a message and goes on: here *change_out >= tol tells).
 * A cross-check of ekpnp_fast_poisson, whose x-derivative is spectral: the
 * two agree exactly for x-invariant charge and differ by the O(dx^2) x
 * discretisation otherwise.  nx must be even (EKPNP_ERR_UNSUPPORTED).
 * Synchronises every check_every sweeps. */
int ekpnp_poisson_iterative(ekpnp_ctx* ctx, int channel, double omega, double tol, int max_iter,
                            int check_every, int* iters_out, double* change_out);
/* With FP32 storage, ekpnp_step re-centres the storage references (see
 * ekpnp_rebase) 8, 16, 32, ... steps after the last re-centring, the
 * interval doubling up to EKPNP_REBASE_INTERVAL and staying there (counted
 * from init_equilibrium), inside a long call too: the re-centring steps
 * depend on the step count only, not on how the run is cut into calls. */
int ekpnp_step(ekpnp_ctx* ctx, int nsteps);
/* FP32 storage keeps each population as a deviation from w_i times a
 * reference (rho, c+, c-)(x, y), the mean over the node's 64-node block of its
 * row (the whole row when nx < 64), plus w_i c_ix/(cs^2 CFL) times the row's
 * mean x-flux.  This re-centres the references on the current fields; the
 * state it represents is unchanged up to one FP32 rounding.  Channels frozen
 * by a running ekpnp_run_to_residual keep theirs.  No-op for FP64 storage.
 * No counterpart in the reference (its populations are absolute FP64
 * values). */
int ekpnp_rebase(ekpnp_ctx* ctx);
/* Steps until every channel's residual sqrt(sum (ux-ux_prev)^2 / sum ux^2)
 * / check_every over interior rows is <= tol, or max_steps.  check_every
 * must be even.  steps_out[channel] receives the step count at which each
 * channel first met tol (-1 if it never did); may be NULL. */
int ekpnp_run_to_residual(ekpnp_ctx* ctx, double tol, int check_every,
                          int max_steps, int* steps_out);

/* ---- output --------------------------------------------------------------- */
int ekpnp_get_field(ekpnp_ctx* ctx, int channel, int field, double* host_out);
/* Pre-collision populations of the next step, reference layout:
 * [9][ny][nx], index 0 = rest (f0), 1..8 = f1 directions. */
int ekpnp_get_populations(ekpnp_ctx* ctx, int channel, int species,
                          double* host_out);
int ekpnp_set_populations(ekpnp_ctx* ctx, int channel, int species,
                          const double* host_in);
/* Per-channel FP64 reductions, out[channel*EKPNP_RED_COUNT + EKPNP_RED_*];
 * also snapshots ux for the next residual. */
int ekpnp_reductions(ekpnp_ctx* ctx, double* out);
/* ekpnp_reductions without any device allocation or allocating library call:
 * no counterpart in the reference; for callers that cannot allocate (code
 * under graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The same kernels, the same results and the same ux snapshot,
 * with the per-tile partial sums and the per-channel records in a
 * library-wide device array of EKPNP_NOALLOC_SCRATCH_DOUBLES doubles (8 MiB;
 * one copy per GPU, held until the process ends) instead of the context's
 * buffers: contexts whose 8 x (LBM work-groups + channels) exceed it (about
 * 33 million nodes: one work-group per 256 nodes of a channel) get
 * EKPNP_ERR_UNSUPPORTED.  Calls on one GPU share the array: calls that
 * overlap on one GPU (two host threads, two contexts' streams) race and must
 * be serialised by the caller.  Synchronises the context's stream. */
#define EKPNP_NOALLOC_SCRATCH_DOUBLES (1 << 20)
int ekpnp_reductions_noalloc(ekpnp_ctx* ctx, double* out);
int ekpnp_synchronize(ekpnp_ctx* ctx);

/* ---- files and diagnostics (host formatting of one channel's fields) -------- */
/* data_end.dat: one line per node (y outer, x inner), "%10.6f" x 9:
 * t, ux, uy, rho, c+, psi, Ex, Ey, c-. */
int ekpnp_save_data_end(ekpnp_ctx* ctx, int channel, double time, const char* path);
/* Reads a data_end.dat written for this channel's nx, ny into the fields
 * (like ekpnp_set_fields with psi; E is recomputed from psi).  *time_out
 * (may be NULL) receives the file's time.  Follow with init_equilibrium. */
int ekpnp_read_data(ekpnp_ctx* ctx, int channel, const char* path, double* time_out);
/* Tecplot POINT zone of 14 variables; first != 0 truncates the file and
 * writes the VARIABLES line, else the zone is appended. */
int ekpnp_save_tecplot(ekpnp_ctx* ctx, int channel, double time, const char* path, int first);
/* Appends c+ (top row extrapolated), one "%10.6f " value per line. */
int ekpnp_save_charge_data(ekpnp_ctx* ctx, int channel, const char* path);
/* Wall current K dy sum_x c+(x, ny-1) Ey(x, ny-1), c+ extrapolated; reduced
 * on the device (FP64), two numbers copied back. */
int ekpnp_current(ekpnp_ctx* ctx, int channel, double* out);
/* max(0, max u_y) with the top row extrapolated (reduced on the device);
 * appended to path as "%10.6f %10.6f\n" (time, umax) unless path is NULL. */
int ekpnp_record_umax(ekpnp_ctx* ctx, int channel, double time, const char* path, double* umax_out);
/* Exact restart: every channel's stored populations (in the storage
 * precision), storage references, psi, macroscopic fields, wall profiles,
 * residual snapshot and step counters.  load needs a context created with
 * the same channel shapes and precision. */
int ekpnp_save_checkpoint(ekpnp_ctx* ctx, const char* path);
int ekpnp_load_checkpoint(ekpnp_ctx* ctx, const char* path);

/* ---- work decomposition ---------------------------------------------------- */
/* Poisson plan: x-FFT work-groups hold up to fft_elems complex values (whole
 * rows), tridiagonal columns are cut into chunks of about tri_rows rows
 * (default 16); ekpnp_create picks fft_elems = 4096 with the row-pair kernels
 * when the widest power-of-two rows have nx = 4096 or 8192, else 1024 with
 * the persistent kernels; fft_persistent selects the x-FFT kernels:
 * 0 one tile per work-group, 1 the persistent software-pipelined kernels
 * where the tiles allow them (fft_elems <= 2048, nx <= 4096), 2 the row-pair
 * kernels (two rows per work-group transformed one after the other, the
 * second row's loads in flight during the first's transform) where every
 * power-of-two channel has nx = 2048 or 4096 and fft_elems = nx, or every
 * one has nx = 8192 and fft_elems = 4096 (the second row waits in
 * registers); otherwise the one-tile kernels.  Results do not depend on
 * fft_elems or the kernel choice; tri_rows changes the FP64 rounding of psi
 * only (chunk boundaries). */
int ekpnp_set_poisson_plan(ekpnp_ctx* ctx, int fft_elems, int tri_rows, int fft_persistent);
/* Times the Poisson stages over these channels for fft_elems in {1024,
 * 2048, 4096}, the three FFT kernel families and tri_rows in {8, 16} (reps solves
 * each, three passes over the candidates, HIP events on the context's
 * stream), keeps the fastest, and restores
 * psi.  out (may be NULL, else 4 doubles) receives {fft_elems, tri_rows, ms
 * per solve, fft_persistent}.  No counterpart in the reference (cuFFT plans are
 * fixed by main.cu:113). */
int ekpnp_autotune(ekpnp_ctx* ctx, int reps, double* out);
/* Tile shape of the row collide-stream kernel (FP32 storage with FP64
 * arithmetic, channels with nx a multiple of 64; a 256-lane work-group is
 * four waves of 64 nodes of one row each): wrows = 0 (default) gives a
 * work-group four consecutive 64-node blocks in node order, wrows = 2 or 4
 * gives it wrows interior rows x 4 / wrows blocks.  Results do not depend on
 * the shape, bit for bit.  Replaces the reference's fixed tiles of one row
 * x nThreads nodes (LBM.cu:311-314). */
int ekpnp_set_lbm_tiles(ekpnp_ctx* ctx, int wrows);
int ekpnp_get_lbm_tiles(const ekpnp_ctx* ctx, int* wrows);
/* Times the row kernel's interior launch for each tile shape (reps launches
 * each, three passes, HIP events on the context's stream) on the current
 * state without advancing it (it writes only the other lattice and the
 * charge density, which is put back), and keeps the fastest (another shape
 * replaces the default only when at least 1 % faster); needs populations.
 * out (may be NULL, else 2 doubles) receives {wrows, ms per launch}; a
 * context with no channel on the row kernel keeps its shape (ms 0).  No
 * counterpart in the reference. */
int ekpnp_autotune_lbm(ekpnp_ctx* ctx, int reps, double* out);
/* The plan in force: out[0] fft_elems, out[1] tri_rows, out[2] persistent FFT
 * kernels in use, out[3] whether a coupled step updates the collide-stream's
 * wall rows (the rows y = 0 and ny - 1 of channels with nx a multiple of 64,
 * FP32 storage with FP64 arithmetic) inside its tridiagonal launch instead of
 * a launch of their own.  That
 * placement is chosen with the plan, when the wall work-groups fit beside
 * the tridiagonal's without another round; EKPNP_WALLS_IN_TRI=0 in the
 * environment at plan time turns it off.  Results do not depend on it.  No
 * counterpart in the reference. */
int ekpnp_get_poisson_plan(const ekpnp_ctx* ctx, int* out);

/* ---- introspection / timing -------------------------------------------------- */
/* The parameters channel was created with (wall profiles not included). */
int ekpnp_get_params(ekpnp_ctx* ctx, int channel, ekpnp_channel_params* out);
int ekpnp_num_channels(const ekpnp_ctx* ctx);
long long ekpnp_steps_done(const ekpnp_ctx* ctx);
/* Device bytes held by the context. */
size_t ekpnp_device_bytes(const ekpnp_ctx* ctx);
/* Times nsteps eager steps with HIP events on the context's stream, per
 * stage: ms_out[0] LBM, [1] forward x-FFT, [2] wall-normal tridiagonal,
 * [3] inverse x-FFT, [4] whole step (graph replay of nsteps after one
 * untimed replay, per step).  The
 * launches are those of the coupled step: when the plan puts the wall rows
 * on the tridiagonal launch (ekpnp_get_poisson_plan), [0] covers the other
 * rows and [2] includes the wall rows. */
int ekpnp_time_stages(ekpnp_ctx* ctx, int nsteps, double* ms_out);
/* Measurement: a non-temporal float4 device copy of `bytes` bytes (a
 * multiple of 16) on `device`, grid work-groups of 256 lanes striding over
 * the buffer, timed with HIP events over reps copies after one untimed one:
 * *gbps_out = 2 bytes / (time per copy) (read + write), *ms_out (may be NULL)
 * the time per copy.  The streaming ceiling the LBM kernel's HBM rate is set
 * against, and the kernel the FETCH_SIZE / WRITE_SIZE counters are
 * calibrated on.  Allocates and frees its two buffers.  No counterpart in
 * the reference. */
int ekpnp_copy_bandwidth(int device, size_t bytes, int reps, int grid, double* gbps_out, double* ms_out);

#ifdef __cplusplus
}
#endif

#endif /* EKPNP_H */
