// ekpnp_internal.h -- device-side descriptors shared by the gfx950 kernels
// (ekpnp_kernels.hip) and the host side of the C ABI (ekpnp_api.cpp).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/ekpnp.h"

namespace ekpnp {

// Wall arrays are stored as 6 planes of length total_nx (sum of nx over the
// channels); channel c owns [wall_off, wall_off + nx) of each plane.
enum WallPlane { W_PSI_LO = 0, W_PSI_HI = 1, W_CP_LO = 2, W_CP_HI = 3, W_CN_LO = 4, W_CN_HI = 5 };

// Macroscopic output planes (FP64), each of total_nodes elements.
enum MacroPlane { M_RHO = 0, M_UX = 1, M_UY = 2, M_CP = 3, M_CN = 4, M_COUNT = 5 };

constexpr int LBM_WG = 256;       // nodes per LBM work-group (one node per lane)
constexpr int FFT_WG = 256;       // lanes of a persistent x-FFT work-group
constexpr int FFT_WG_ROWS = 512;  // lanes of a one-tile x-FFT work-group (radix-8 passes)
constexpr int FFT_MAX_ELEMS = 4096;  // complex values held in LDS by one FFT work-group
constexpr int FFT_ROW_SLOTS = FFT_MAX_ELEMS / FFT_WG_ROWS;  // values per lane of a one-tile load
// two-level twiddle table of the x-FFTs: 64 fine values W^f, then
// max(1, nx / 64) coarse values W^(64 c); at most 192 (3 KB) for nx <= 8192
constexpr int TW2_LGF = 6, TW2_FINE = 1 << TW2_LGF;
__host__ __device__ constexpr int tw2_count(int nx) { return TW2_FINE + (nx / TW2_FINE > 1 ? nx / TW2_FINE : 1); }
constexpr int FFT_PAIR_WIDE_S = 8;  // values per lane and row of the wide row-pair kernels (nx = 8192)
constexpr int FFT_MAX_ROWS = 256;  // rows per power-of-two FFT tile
// spectrum pairs (X[p], X[N-p]) per lane: rows (N/2 + 1) <= FFT_MAX_ELEMS/2 + FFT_MAX_ROWS
constexpr int FFT_PAIR_SLOTS = (FFT_MAX_ELEMS / 2 + FFT_MAX_ROWS + FFT_WG_ROWS - 1) / FFT_WG_ROWS;
// persistent pipelined x-FFT: tiles of <= FFT_PERSIST_ELEMS complex values,
// 8 prefetched per lane; LDS = rows + quadrant twiddles (nx <= 2 FFT_PERSIST_ELEMS)
constexpr int FFT_PF = 8;
constexpr int FFT_PERSIST_ELEMS = FFT_PF * FFT_WG;  // 2048
constexpr unsigned FFT_LDS_PERSIST = (FFT_PERSIST_ELEMS + FFT_PERSIST_ELEMS / 2) * 16;
constexpr int TRI_WG = 256;       // lanes of a tridiagonal work-group (modes x chunks)
constexpr int MAX_NX = 8192;      // LDS-resident x-rows (64 KiB of complex FP64)
constexpr int MAX_NX_DFT = 2048;  // non-power-of-two rows use a direct DFT
constexpr int TRI_LMAX = 16;              // rows per tridiagonal chunk (one lane)
constexpr int MAX_INTERIOR_ROWS = 2048;  // 128 chunks x 16 rows per tridiagonal column
// FP32 storage references along x: one record of BLK_STRIDE floats (rho,
// c+, c-, unused) per block of REF_BLOCK consecutive nodes of a row (one
// wave of the row kernel); a reference need not be exact, only the same
// wherever it is used, and FP32 values halve the scalar registers a wave
// holds them in
constexpr int REF_BLOCK = 64;
constexpr int BLK_STRIDE = 4;

struct ChanDesc {
  int nx, ny, nk, flags;        // nk = nx/2 + 1 spectral columns
  int nk_stride;                // nk rounded up to a multiple of 8 (spectrum blocks)
  int active;                   // 0: the step kernels skip this channel (frozen)
  int swap;                     // 1: this channel's state is in the step's dst lattice
  int nb;                       // reference blocks per row: ceil(nx / REF_BLOCK)
  long long pop_off;            // first element of this channel's 27 planes
  long long plane;              // population plane stride (nx*ny rounded up to 64)
  long long fld_off;            // first element of this channel's [ny][nx] fields
  long long spec_off;           // first double2 of this channel's spectrum blocks
  long long ref_off;            // first double of this channel's [ny][8] storage references
  long long blk_off;            // first float of this channel's [ny][nb][BLK_STRIDE] block records
  long long cp_off;             // first double of this channel's [tri_C][nk] chunk-entry Thomas coefficients
  int wall_off;                 // offset into each wall plane
  int tw_off;                   // offset of this channel's twiddles e^{-2 pi i j/nx}
  int pow2;                     // nx is a power of two (Stockham path)
  int fft_rows;                 // rows per FFT work-group
  int tri_C, tri_G, tri_L, tile0;  // tridiagonal chunks, modes per WG, rows per chunk
  // collision constants (double; the FP32 kernel rounds them once)
  double inv_cfl, inv_cs2;
  double om_pf, om_mf;          // omega+ dt, omega- dt, fluid
  double om_pp, om_mp;          // positive ion
  double om_pn, om_mn;          // negative ion
  double src_p, src_m;          // 1 - omega+- dt/2 (Guo source split)
  double dt, half_dt;
  double kp, kn;                // mobilities
  double fcoef, e_app, exf;     // body force
  double drift_app;             // e_app if ions drift in the applied field else 0
  double inv2dx, inv2dy;        // central-difference factors
  double multis, multid;        // moving-wall terms of the top wall
  double pc_dy2;                // poisson_coef * dy^2
  double kdy_scale;             // (2 pi / Lx) * dy
  double c_bulk, inv_vt, pb_omega;  // PB-Picard start
};

struct StepBuffers {
  const ChanDesc* chans;
  const int2* lbm_tiles;        // (channel, first node) per LBM work-group
  int n_lbm_tiles;
  // the same tiles split for the coupled step: channels with nx a multiple
  // of 64 go to k_lbm_rows (tiles with interior rows, tiles with wall rows),
  // all the others to k_lbm
  const int2* lbm_row_tiles;
  int n_lbm_row_tiles;
  const int2* lbm_wall_tiles;
  int n_lbm_wall_tiles;
  const int2* lbm_gen_tiles;
  int n_lbm_gen_tiles;
  int lbm_wrows;                // row-kernel interior tiles: 0 consecutive blocks, else rows per tile (2, 4)
  const int4* fft_tiles;        // (channel, first row, rows, 0) per FFT work-group
  int n_fft_tiles;
  int n_fft_pow2;               // the first n_fft_pow2 tiles are power-of-two rows
  int fft_persist;              // power-of-two tiles by the persistent kernels
  int fft_pair_s;               // > 0: by the row-pair kernels, N = fft_pair_s * FFT_WG_ROWS for every one
                                //   (2, 4: two row buffers and the quadrant table; FFT_PAIR_WIDE_S: one row
                                //   buffer, the second row in registers, the two-level table)
  int fft_grid;                 // work-groups of the persistent kernels
  unsigned fft_lds_pow2;        // dynamic LDS bytes of the Stockham kernels
  unsigned fft_lds_dft;         // dynamic LDS bytes of the direct-DFT kernels
  int fft_tw_lds;               // the power-of-two kernels stage twiddles in LDS
  const int2* tri_tiles;        // (channel, first mode) per tridiagonal work-group
  int n_tri_tiles;
  int tri_lmax;                 // longest tridiagonal chunk (<= TRI_LMAX)
  double* rho_e;                // c+ - c- per node (FP64)
  double* psi;                  // potential per node (FP64)
  double* macro;                // M_COUNT planes
  long long macro_stride;       // = total nodes
  double* walls;                // 6 wall planes
  int walls_stride;             // = total nx
  const double2* tw;            // twiddles
  double2* spec;                // spectrum [nk_stride/8][ny][8] per channel
  double* ux_prev;              // residual snapshot
  double* partials;             // reduction partials [n_lbm_tiles][8]
  double* red_out;              // reductions [n_channels][8]
  const int* red_first_tile;    // first LBM tile of each channel (n_channels + 1)
  int n_channels;
  double* refs;                 // storage references, [ny][8] per channel (row means of rho, c+, c-, their x-fluxes, 2 unused)
  float* blocks;                // storage references along x, [ny][nb][BLK_STRIDE] per channel
  const int2* row_tiles;        // (channel, row) per row-means work-group
  int n_row_tiles;
  double* cprime;               // Thomas coefficient entering chunk j of mode k, [tri_C][nk] per channel
};

// ---- launchers (ekpnp_kernels.hip) ----------------------------------------
// precision: 0 FP32 deviations, 1 FP64; compute64: FP64 collision arithmetic
// for FP32 storage
// walls = false: the row kernel's wall tiles are left out (for
// launch_poisson_tri_walls to update them)
hipError_t launch_lbm_rows_interior(const StepBuffers& b, const void* src, void* dst, hipStream_t s);
hipError_t launch_lbm(int precision, bool compute64, bool from_q, bool macro, const StepBuffers& b,
                      const void* src, void* dst, hipStream_t s, bool walls = true);
hipError_t launch_poisson_fwd(const StepBuffers& b, hipStream_t s);
hipError_t launch_poisson_tri(const StepBuffers& b, hipStream_t s);
// the tridiagonal plus the row kernel's wall tiles of the collide-stream that
// read src and write dst (FP32 storage, FP64 arithmetic)
hipError_t launch_poisson_tri_walls(bool macro, const StepBuffers& b, const void* src, void* dst, hipStream_t s);
hipError_t launch_poisson_inv(const StepBuffers& b, hipStream_t s);
// resident one-tile x-FFT work-groups per CU with twiddles in LDS (twl) or
// not, lds bytes of dynamic LDS each (registers and LDS both counted); 0 on error
int fft_one_tile_blocks_per_cu(bool twl, unsigned lds);
int tridiag_walls_blocks_per_cu();
// chunk-entry Thomas coefficients of channel chan for its current tri_C, tri_L
// chunk-entry Thomas coefficients of every channel (max_entries: the
// largest tri_C x nk over the channels)
hipError_t launch_cprime_tables(const StepBuffers& b, int n_channels, long long max_entries, hipStream_t s);
hipError_t launch_init_equilibrium(int precision, const StepBuffers& b, void* dst, hipStream_t s);
// pre-collision populations of the next step, written unshifted into dst
hipError_t launch_to_q(int precision, const StepBuffers& b, const void* src, void* dst,
                       hipStream_t s);
hipError_t launch_reductions(const StepBuffers& b, int snapshot, hipStream_t s);
// PB-Picard start: Boltzmann concentrations from psi, then under-relaxation
// of psi against psi_old (psi_old is updated to the relaxed value)
hipError_t launch_pb_fields(const StepBuffers& b, hipStream_t s);
// row means over x of rho, c+, c- and rho u_x, c+ u_x, c- u_x into out ([ny][8] per channel), and the
// block records (means of rho, c+, c- over each block of REF_BLOCK nodes) into blocks_out; channels that
// are not active keep the references in force (b.refs, b.blocks)
hipError_t launch_row_means(const StepBuffers& b, double* out, float* blocks_out, hipStream_t s);
// re-express every stored population of the active channels against new_refs
// / new_blocks (state in cur, or in other for channels whose swap flag is set)
hipError_t launch_rebase(int precision, const StepBuffers& b, void* cur, void* other, const double* new_refs,
                         const float* new_blocks, hipStream_t s);
hipError_t launch_pb_relax(const StepBuffers& b, double* psi_old, hipStream_t s);
// one red-black SOR half-sweep of channel chan's 5-point Poisson system
// (color 0: x + y even); partials[wg] = largest |change| per work-group,
// ceil((nx/2)(ny-2)/LBM_WG) of them
hipError_t launch_sor_sweep(const StepBuffers& b, int chan, int nx, int ny, int color, double omega, double r,
                            double inv, double* partials, hipStream_t s);
hipError_t launch_max_final(const double* v, int n, double* out, hipStream_t s);
// psi wall rows of channel chan from the wall potentials
hipError_t launch_psi_wall_rows(const StepBuffers& b, int chan, int nx, hipStream_t s);
// Allocation-free reductions: a file-scope device array of
// NOALLOC_SCRATCH_DOUBLES doubles per GPU; noalloc_scratch_init looks up its
// address on `device` (which must be current) once, noalloc_scratch returns
// it (nullptr before the lookup)
constexpr long long NOALLOC_SCRATCH_DOUBLES = EKPNP_NOALLOC_SCRATCH_DOUBLES;  // 8 MiB
hipError_t noalloc_scratch_init(int device);
double* noalloc_scratch(int device);
// Run-time diagnostics of one channel (umax and the top-row current sum, see
// k_diag_partial): grid (1..DIAG_MAX_WG) work-groups of 256 lanes write
// 2 x grid partials, then out[0] = umax, out[1] = sum over the top row of
// c+ E_y (uy, cp, psi: the channel's planes, nx * ny each)
constexpr int DIAG_MAX_WG = 1024;
hipError_t launch_diagnostics(const double* uy, const double* cp, const double* psi, int nx, int ny,
                              double half_inv_dy, double* partials, int grid, double* out, hipStream_t s);
// umax and the top-row current sum of one channel on the device (the
// context's stream, synchronised); either output may be null
int channel_diagnostics(ekpnp_ctx* ctx, int channel, double* umax, double* current_sum);
// dst = src, n_float4 16-byte values, non-temporal, grid work-groups of 256 lanes
hipError_t launch_copy_nt(const void* src, void* dst, long long n_float4, int grid, hipStream_t s);

}  // namespace ekpnp
