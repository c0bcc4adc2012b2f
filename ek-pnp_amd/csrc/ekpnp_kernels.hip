// ekpnp_kernels.hip -- gfx950 kernels of the EK-PNP coupled time step.
//
// One step is an LBM stage and three Poisson launches, each launch covering
// every channel of the context:
//   k_lbm       fused pull collide-stream of f (fluid), h (c+), g (c-) with the
//               wall rules folded in; writes the post-collision state of every
//               node in place of the next lattice and the FP64 charge density
//               (replaces gpu_collide_save, gpu_boundary, gpu_stream,
//               gpu_bc_charge: LBM.cu:331-1024)
//   k_lbm_rows  the same, same bits, with fewer registers for FP32 storage with
//               FP64 arithmetic on channels with nx a multiple of 64 (one
//               launch for their wall rows, one for their interior rows; k_lbm
//               keeps the other channels and the first step after
//               initialisation)
//   k_fft_fwd   per x-row: right-hand side with the wall lifts, real-to-complex
//               FFT in LDS (replaces odd_extension + forward cuFFT,
//               poisson.cu:211-263, 190)
//   k_tridiag   per wavenumber: the wall-normal tridiagonal solve as a chunked
//               Thomas sweep joined by affine scans in LDS (replaces
//               gpu_derivative and the y half of the 2-D cuFFT,
//               poisson.cu:265-283); in a coupled step k_tridiag_walls,
//               which also runs that step's k_lbm_rows wall rows
//   k_fft_inv   per x-row: complex-to-real FFT, writes psi (replaces inverse
//               cuFFT + odd_extract, poisson.cu:196, 285-306)
// The electric field is never stored: k_lbm differentiates psi on the fly
// exactly as gpu_efield/gpu_bc do (poisson.cu:145-174).
//
// Lattice state between steps ("P'"): at every node the post-collision
// populations, except that fluid wall nodes hold the reversed pre-collision
// populations (full-way bounce-back, LBM.cu:816-846).  The next step gathers
// them (pull streaming with the reference's x and y wrap, LBM.cu:869-905); ion
// wall nodes instead anti-bounce-back their own values (LBM.cu:915-1023).
// Right after init_equilibrium or set_populations the lattice holds the
// pre-collision populations themselves, unshifted ("Q"), and the first step
// reads them locally (FROM_Q).
#include <mutex>
#include <type_traits>

#include "ekpnp_internal.h"

namespace ekpnp {

constexpr double W0 = 4.0 / 9.0, WS = 1.0 / 9.0, WD = 1.0 / 36.0;  // LBM.h:105-107

// Build-time tuning of k_lbm (defaults chosen by measurement on MI355X, where
// non-temporal loads and stores of the populations raised the kernel from
// 5.0 to 5.7 TB/s at 4096 x 1024):
//   EKPNP_NT_STORE  1: post-collision populations are written with
//                      non-temporal stores (streamed past the caches)
//   EKPNP_NT_LOAD   1: the gather reads are non-temporal as well
//   EKPNP_LBM_WAVES N: ask for at least N waves per SIMD (register budget)
#ifndef EKPNP_NT_STORE
#define EKPNP_NT_STORE 1
#endif
#ifndef EKPNP_NT_LOAD
#define EKPNP_NT_LOAD 1
#endif
// k_lbm asks for 3 waves per SIMD: its FP32-storage, FP64-arithmetic
// variants otherwise take 169-171 VGPRs (2 waves); at 168 the compiler spills
// 1-4 VGPRs and the kernel is 5-14 % faster (4096x1024: 214 -> 203 us;
// 5000 channels of nx = 4: 214 -> 184 us); variants already within 168 are
// unchanged
#ifndef EKPNP_LBM_WAVES
#define EKPNP_LBM_WAVES 3
#endif
#if EKPNP_LBM_WAVES > 0
#define LBM_WAVES_ATTR __attribute__((amdgpu_waves_per_eu(EKPNP_LBM_WAVES)))
#else
#define LBM_WAVES_ATTR
#endif

template <typename S>
__device__ __forceinline__ S ld_pop(const S* p) {
#if EKPNP_NT_LOAD
  return __builtin_nontemporal_load(p);
#else
  return *p;
#endif
}

template <typename S>
__device__ __forceinline__ void st_pop(S* p, S v) {
#if EKPNP_NT_STORE
  __builtin_nontemporal_store(v, p);
#else
  *p = v;
#endif
}
constexpr unsigned FLAG_G0_FLUID_U = 0x2u;

// Work-groups are handed to the 8 XCDs round-robin; remap so that each XCD
// owns one contiguous range of tiles and neighbouring rows share its L2.
__device__ __forceinline__ int xcd_remap(int wg, int nwg) {
  const int q = nwg >> 3, r = nwg & 7;
  const int xcd = wg & 7, slot = wg >> 3;
  return xcd * q + min(xcd, r) + slot;
}

// ---------------------------------------------------------------------------
// E = -grad(psi) by central differences at node (x, y); wall rows copy Ey of
// the neighbouring row (poisson.cu:145-174).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void efield_at(const double* __restrict__ ps, int nx, int ny, int x,
                                          int xm, int xp, int y, double inv2dx, double inv2dy,
                                          double& Ex, double& Ey) {
  const double* row = ps + (long long)y * nx;
  Ex = (row[xm] - row[xp]) * inv2dx;
  int ya = y - 1, yb = y + 1;
  if (y == 0) { ya = 0; yb = 2; }
  if (y == ny - 1) { ya = ny - 3; yb = ny - 1; }
  Ey = (ps[(long long)ya * nx + x] - ps[(long long)yb * nx + x]) * inv2dy;
}

// Storage references (FP32 storage; FP64 storage keeps absolute populations
// and uses none).  Per channel two tables:
//   row records R[y] (REF_STRIDE doubles): the row means of X = (rho, c+, c-)
//     and of their x-fluxes J = (rho u_x, c+ u_x, c- u_x), two unused slots;
//   block records B[y][k] (BLK_STRIDE floats), k < nb = ceil(nx/64): the
//     means of X over columns 64 k .. 64 k + 63 of the row, rounded to FP32.
// The reference of node (x, y) is X(x, y) = B[y][x/64], and a stored
// population is
//   stored_i = f_i - w_i X(x, y) - w_i c_ix tc J(y),   tc = 1/(cs^2 CFL),
// the departure from the first-order equilibrium of the references.  The
// collision works on the rest deviation q_i = f_i - w_i X(x, y).  Following
// the double layer along x in 64-node blocks (wall potentials that vary
// with x) keeps the stored ion deviations, and their FP32 rounding, small,
// where a row mean alone leaves deviations of the order of the
// concentration itself.  A block is one wave of the row kernel, so its
// references are wave-uniform.
constexpr int REF_ROWS = 6;     // values per row record
constexpr int REF_STRIDE = 8;   // doubles per row record
static_assert(REF_BLOCK == 64, "reference blocks are found by shifts");

// x-fluxes J of a row record
struct RowJ {
  double J[3];
};
__device__ __forceinline__ RowJ load_rowj(const double* __restrict__ R, int y) {
  const double* r = R + (long long)REF_STRIDE * y;
  return RowJ{{r[3], r[4], r[5]}};
}

// X_s at column x of the row whose block records start at Br
__device__ __forceinline__ double block_x(const float* __restrict__ Br, int x, int s) {
  return (double)Br[BLK_STRIDE * (x >> 6) + s];
}

// The references of one node, generic path: block rows of y-1, y, y+1
// (periodic in y), the node's columns x-1, x, x+1 and the three row fluxes.
struct NodeRefs {
  const float* Bm;
  const float* Bo;
  const float* Bp;
  int x, xm, xp;
  RowJ jm, jo, jp;
  __device__ double own(int s) const { return block_x(Bo, x, s); }
  // X(src) - X(node) for the 9 gather sources of D2Q9 (d[0] = 0)
  __device__ void diffs(int s, double* d) const {
    const double o = own(s);
    d[0] = 0.0;
    d[1] = block_x(Bo, xm, s) - o;
    d[2] = block_x(Bm, x, s) - o;
    d[3] = block_x(Bo, xp, s) - o;
    d[4] = block_x(Bp, x, s) - o;
    d[5] = block_x(Bm, xm, s) - o;
    d[6] = block_x(Bm, xp, s) - o;
    d[7] = block_x(Bp, xp, s) - o;
    d[8] = block_x(Bp, xm, s) - o;
  }
};
__device__ __forceinline__ NodeRefs node_refs(const StepBuffers& b, const ChanDesc* __restrict__ cd, int x, int y,
                                              int xm, int xp) {
  const int ny = cd->ny;
  const int ym = (y == 0) ? ny - 1 : y - 1;
  const int yp = (y == ny - 1) ? 0 : y + 1;
  const float* B = b.blocks + cd->blk_off;
  const double* R = b.refs + cd->ref_off;
  const long long bs = (long long)BLK_STRIDE * cd->nb;
  return NodeRefs{B + ym * bs, B + y * bs, B + yp * bs, x, xm, xp,
                  load_rowj(R, ym), load_rowj(R, y), load_rowj(R, yp)};
}

// direction x-components c_ix of D2Q9 (0, E, N, W, S, NE, NW, SW, SE)
__device__ constexpr int CX[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};

// The gathered species' stored values v (FP32 storage) plus the reference
// terms of their sources: q_i = v_i + w_i (X(src) - X(node)) + w_i c_ix tc J(src row)
// with dX = X(src) - X(node) and the row fluxes (times tc) j0 (row y), jm
// (y-1), jp (y+1).
template <typename C>
__device__ __forceinline__ void add_src_refs(C* q, const double* dX, C j0, C jm, C jp) {
  q[1] += C(WS) * (C(dX[1]) + j0);
  q[3] += C(WS) * (C(dX[3]) - j0);
  q[2] += C(WS) * C(dX[2]);
  q[4] += C(WS) * C(dX[4]);
  q[5] += C(WD) * (C(dX[5]) + jm);
  q[6] += C(WD) * (C(dX[6]) - jm);
  q[7] += C(WD) * (C(dX[7]) - jp);
  q[8] += C(WD) * (C(dX[8]) + jp);
}

// rest deviation of node (x, y) from stored values at the node itself
// (FROM_Q) or gathered from the upstream neighbours (pull streaming):
//   q_i = stored_i(src) + w_i (X(src) - X(node)) + w_i c_ix tc J(src).
// 27 values: f 0..8, h 9..17, g 18..26.  REF false (FP64 storage): no
// reference terms.
template <typename S, typename C, bool FROM_Q, bool REF>
__device__ __forceinline__ void load_pre(const S* __restrict__ P, long long plane, int nx, int ny,
                                         int x, int y, int xm, int xp, const NodeRefs& rf, double tc, C* q) {
  const long long r0 = (long long)y * nx;
  if (FROM_Q) {
#pragma unroll
    for (int s = 0; s < 3; ++s) {
      const C j = REF ? C(rf.jo.J[s] * tc) : C(0);
#pragma unroll
      for (int d = 0; d < 9; ++d) {
        q[9 * s + d] = (C)ld_pop(P + (9 * s + d) * plane + r0 + x);
        if (REF && CX[d] != 0) q[9 * s + d] += C(CX[d] * (d < 5 ? WS : WD)) * j;
      }
    }
    return;
  }
  const int ym = (y == 0) ? ny - 1 : y - 1;
  const int yp = (y == ny - 1) ? 0 : y + 1;
  const long long rm = (long long)ym * nx, rp = (long long)yp * nx;
#pragma unroll
  for (int s = 0; s < 3; ++s) {
    const S* B = P + (long long)(9 * s) * plane;
    q[9 * s + 0] = (C)ld_pop(B + 0 * plane + r0 + x);
    q[9 * s + 1] = (C)ld_pop(B + 1 * plane + r0 + xm);
    q[9 * s + 2] = (C)ld_pop(B + 2 * plane + rm + x);
    q[9 * s + 3] = (C)ld_pop(B + 3 * plane + r0 + xp);
    q[9 * s + 4] = (C)ld_pop(B + 4 * plane + rp + x);
    q[9 * s + 5] = (C)ld_pop(B + 5 * plane + rm + xm);
    q[9 * s + 6] = (C)ld_pop(B + 6 * plane + rm + xp);
    q[9 * s + 7] = (C)ld_pop(B + 7 * plane + rp + xp);
    q[9 * s + 8] = (C)ld_pop(B + 8 * plane + rp + xm);
    if (REF) {
      double dX[9];
      rf.diffs(s, dX);
      add_src_refs(q + 9 * s, dX, C(rf.jo.J[s] * tc), C(rf.jm.J[s] * tc), C(rf.jp.J[s] * tc));
    }
  }
}

// stored value of rest deviation v in direction d of a row with flux term j = tc J
template <typename C>
__device__ __forceinline__ C to_stored(C v, int d, C j) {
  return CX[d] == 0 ? v : v - C(CX[d] * (d < 5 ? WS : WD)) * j;
}

// Ion wall node: the next pre-collision state is the anti-bounce-back of the
// node's own post-collision state, q[opp(i)] = -p[i] + 2 w_i c_w
// (LBM.cu:935-943), written here for rest deviations (dcw = c_w - c_ref; the
// stored p[i] gets its flux term j = tc J back first).
template <typename S, typename C>
__device__ __forceinline__ void abb_own(const S* __restrict__ P, long long plane, long long node,
                                        C dcw, C j, C* q) {
  const C two_s = C(2.0 * WS) * dcw, two_d = C(2.0 * WD) * dcw;
  const C js = C(WS) * j, jd = C(WD) * j;
  q[0] = -(C)P[0 * plane + node] + C(2.0 * W0) * dcw;
  q[3] = -((C)P[1 * plane + node] + js) + two_s;
  q[4] = -(C)P[2 * plane + node] + two_s;
  q[1] = -((C)P[3 * plane + node] - js) + two_s;
  q[2] = -(C)P[4 * plane + node] + two_s;
  q[7] = -((C)P[5 * plane + node] + jd) + two_d;
  q[8] = -((C)P[6 * plane + node] - jd) + two_d;
  q[5] = -((C)P[7 * plane + node] - jd) + two_d;
  q[6] = -((C)P[8 * plane + node] + jd) + two_d;
}

// ---------------------------------------------------------------------------
// TRT collision in storage-deviation form.  For a pair (i, opp i) with
// t = (e_i . v)/(cs^2 CFL) the equilibrium splits into
//   eq+ = w (dX + X (om + t^2/2)),  eq- = w X t,  om = -|v|^2/(2 cs^2)
// (LBM.cu:468-552, 599-722); the update is
//   X_i' = X_i - w+ (X+ - eq+) - w- (X- - eq-)   (LBM.cu:775-803).
// ---------------------------------------------------------------------------
template <typename C>
__device__ __forceinline__ void trt_pair(C& a, C& b, C w, C dX, C X, C om, C t, C omp, C omm) {
  const C eqp = w * (dX + X * (om + C(0.5) * t * t));
  const C eqm = w * X * t;
  const C np = C(0.5) * (a + b) - eqp;
  const C nm = C(0.5) * (a - b) - eqm;
  a = a - omp * np - omm * nm;
  b = b - omp * np + omm * nm;
}

template <typename C>
__device__ __forceinline__ void collide_ion(C* q, C dX, C X, C vx, C vy, C omp, C omm, C ics,
                                            C ticf, C rest_om) {
  const C om = C(-0.5) * (vx * vx + vy * vy) * ics;
  const C tx = vx * ticf, ty = vy * ticf;
  const C eq0 = C(W0) * (dX + X * rest_om);
  q[0] = q[0] - omp * (q[0] - eq0);
  trt_pair(q[1], q[3], C(WS), dX, X, om, tx, omp, omm);
  trt_pair(q[2], q[4], C(WS), dX, X, om, ty, omp, omm);
  trt_pair(q[5], q[7], C(WD), dX, X, om, tx + ty, omp, omm);
  trt_pair(q[6], q[8], C(WD), dX, X, om, ty - tx, omp, omm);
}

// Fluid pair with the Guo source split into its even part
// S+ = w/cs^2 (-u.F + t (c.F)) and odd part S- = w/cs^2 (c.F), weighted by
// (1 - w+-/2) dt (LBM.cu:556-572, 745-756).
template <typename C>
__device__ __forceinline__ void trt_pair_src(C& a, C& b, C w, C dX, C X, C om, C t, C omp, C omm,
                                             C uF, C cF, C ics, C sp_dt, C sm_dt) {
  const C eqp = w * (dX + X * (om + C(0.5) * t * t));
  const C eqm = w * X * t;
  const C np = C(0.5) * (a + b) - eqp;
  const C nm = C(0.5) * (a - b) - eqm;
  const C Sp = w * ics * (t * cF - uF);
  const C Sm = w * ics * cF;
  a = a - omp * np - omm * nm + sp_dt * Sp + sm_dt * Sm;
  b = b - omp * np + omm * nm + sp_dt * Sp - sm_dt * Sm;
}

template <typename C>
__device__ __forceinline__ void collide_fluid(C* q, C dr, C rho, C ux, C uy, C Fx, C Fy, C omp,
                                              C omm, C ics, C icfl, C sp_dt, C sm_dt) {
  const C ticf = ics * icfl;
  const C om = C(-0.5) * (ux * ux + uy * uy) * ics;
  const C tx = ux * ticf, ty = uy * ticf;
  const C uF = ux * Fx + uy * Fy;
  const C cFx = Fx * icfl, cFy = Fy * icfl;
  const C eq0 = C(W0) * (dr + rho * om);
  q[0] = q[0] - omp * (q[0] - eq0) + sp_dt * (C(W0) * ics * (-uF));
  trt_pair_src(q[1], q[3], C(WS), dr, rho, om, tx, omp, omm, uF, cFx, ics, sp_dt, sm_dt);
  trt_pair_src(q[2], q[4], C(WS), dr, rho, om, ty, omp, omm, uF, cFy, ics, sp_dt, sm_dt);
  trt_pair_src(q[5], q[7], C(WD), dr, rho, om, tx + ty, omp, omm, uF, cFx + cFy, ics, sp_dt, sm_dt);
  trt_pair_src(q[6], q[8], C(WD), dr, rho, om, ty - tx, omp, omm, uF, cFy - cFx, ics, sp_dt, sm_dt);
}

template <typename S>
struct Compute { using T = double; };
template <>
struct Compute<float> { using T = float; };

// Force-corrected velocity of a pre-collision state (LBM.cu:382-392).
template <typename C>
struct Moments {
  C dr, rho, dcp, dcn, rhoe, Fx, Fy, ux, uy;
};

template <typename C>
__device__ __forceinline__ Moments<C> moments(const C* q, const ChanDesc* __restrict__ cd, const double* Xo,
                                              double Ex, double Ey) {
  Moments<C> m;
  m.dr = ((q[0] + q[1]) + (q[2] + q[3])) + ((q[4] + q[5]) + (q[6] + q[7])) + q[8];
  m.rho = C(Xo[0]) + m.dr;
  m.dcp = ((q[9] + q[10]) + (q[11] + q[12])) + ((q[13] + q[14]) + (q[15] + q[16])) + q[17];
  m.dcn = ((q[18] + q[19]) + (q[20] + q[21])) + ((q[22] + q[23]) + (q[24] + q[25])) + q[26];
  m.rhoe = (m.dcp - m.dcn) + C(Xo[1] - Xo[2]);
  const C jx = (q[1] + q[5] + q[8]) - (q[3] + q[6] + q[7]);
  const C jy = (q[2] + q[5] + q[6]) - (q[4] + q[7] + q[8]);
  m.Fx = C(cd->fcoef) * m.rhoe * C(Ex + cd->e_app) + C(cd->exf);
  m.Fy = C(cd->fcoef) * m.rhoe * C(Ey);
  const C irho = C(1.0) / m.rho;
  const C icfl = C(cd->inv_cfl), hdt = C(cd->half_dt);
  m.ux = (jx * icfl + m.Fx * hdt) * irho;
  m.uy = (jy * icfl + m.Fy * hdt) * irho;
  return m;
}

// ---------------------------------------------------------------------------
// k_lbm: one coupled-step LBM update of every node of every channel.
// ---------------------------------------------------------------------------
// S: storage type of the populations; C: arithmetic type of the collision.
template <typename S, typename C, bool FROM_Q, bool MACRO>
__global__ __launch_bounds__(LBM_WG) LBM_WAVES_ATTR void k_lbm(StepBuffers b, const int2* __restrict__ tiles,
                                                const S* __restrict__ src, S* __restrict__ dst) {
  const int wg = xcd_remap(blockIdx.x, gridDim.x);
  const int2 tile = tiles[wg];
  const ChanDesc* __restrict__ cd = b.chans + tile.x;
  if (!cd->active) return;  // converged channel, frozen by the residual driver
  const int nx = cd->nx, ny = cd->ny;
  const int node = tile.y + (int)threadIdx.x;
  if (node >= nx * ny) return;
  const int y = node / nx;
  const int x = node - y * nx;
  const long long plane = cd->plane;
  // a channel frozen for an odd number of steps holds its state in the other lattice
  const S* __restrict__ P = (cd->swap ? dst : src) + cd->pop_off;
  S* __restrict__ Q = (cd->swap ? const_cast<S*>(src) : dst) + cd->pop_off;
  const double* __restrict__ ps = b.psi + cd->fld_off;
  const int xm = (x == 0) ? nx - 1 : x - 1;
  const int xp = (x == nx - 1) ? 0 : x + 1;
  const bool bottom = (y == 0), top = (y == ny - 1);
  const double tc = cd->inv_cs2 * cd->inv_cfl;
  // FP32 storage: deviations from the block references (FP64 storage: none)
  constexpr bool REF = std::is_same<S, float>::value;
  NodeRefs rf{};
  double Xo[3] = {0.0, 0.0, 0.0};
  if (REF) {
    rf = node_refs(b, cd, x, y, xm, xp);
#pragma unroll
    for (int s = 0; s < 3; ++s) Xo[s] = rf.own(s);
  }

  // y = 0: velocity mirrored from node (x,1)'s pre-collision state, force
  // included (LBM.cu:402-443); evaluated first so its registers are free.
  C mux = 0, muy = 0;
  if (bottom) {
    C m1[27];
    NodeRefs r1{};
    double X1[3] = {0.0, 0.0, 0.0};
    if (REF) {
      r1 = node_refs(b, cd, x, 1, xm, xp);
#pragma unroll
      for (int s = 0; s < 3; ++s) X1[s] = r1.own(s);
    }
    load_pre<S, C, FROM_Q, REF>(P, plane, nx, ny, x, 1, xm, xp, r1, tc, m1);
    double Ex1, Ey1;
    efield_at(ps, nx, ny, x, xm, xp, 1, cd->inv2dx, cd->inv2dy, Ex1, Ey1);
    const Moments<C> mm = moments(m1, cd, X1, Ex1, Ey1);
    mux = -mm.ux;
    muy = -mm.uy;
  }

  C q[27];
  load_pre<S, C, FROM_Q, REF>(P, plane, nx, ny, x, y, xm, xp, rf, tc, q);
  if (!FROM_Q && (bottom || top)) {
    const double* __restrict__ Wl = b.walls + cd->wall_off + x;
    const long long ws = b.walls_stride;
    const double cwp = bottom ? Wl[W_CP_LO * ws] : Wl[W_CP_HI * ws];
    const double cwn = bottom ? Wl[W_CN_LO * ws] : Wl[W_CN_HI * ws];
    abb_own<S, C>(P + 9 * plane, plane, node, C(cwp - Xo[1]), REF ? C(rf.jo.J[1] * tc) : C(0), q + 9);
    abb_own<S, C>(P + 18 * plane, plane, node, C(cwn - Xo[2]), REF ? C(rf.jo.J[2] * tc) : C(0), q + 18);
  }

  double Ex, Ey;
  efield_at(ps, nx, ny, x, xm, xp, y, cd->inv2dx, cd->inv2dy, Ex, Ey);
  const Moments<C> m = moments(q, cd, Xo, Ex, Ey);
  C ux = m.ux, uy = m.uy;
  if (bottom) {
    ux = mux;
    uy = muy;
  }

  // charge density for the Poisson solve, summed in FP64
  {
    double sp = 0.0, sn = 0.0;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      sp += (double)q[9 + k];
      sn += (double)q[18 + k];
    }
    b.rho_e[cd->fld_off + node] = (sp - sn) + (Xo[1] - Xo[2]);
  }
  if (MACRO) {
    const long long o = cd->fld_off + node, st = b.macro_stride;
    b.macro[M_RHO * st + o] = (double)m.rho;
    b.macro[M_UX * st + o] = (double)ux;
    b.macro[M_UY * st + o] = (double)uy;
    b.macro[M_CP * st + o] = (double)(C(Xo[1]) + m.dcp);
    b.macro[M_CN * st + o] = (double)(C(Xo[2]) + m.dcn);
  }

  const C ics = C(cd->inv_cs2), icfl = C(cd->inv_cfl), ticf = ics * icfl;
  // ions: drift velocity u + K E (LBM.cu:479-488)
  {
    const C Eix = C(Ex + cd->drift_app), Eiy = C(Ey);
    const C cp = C(Xo[1]) + m.dcp, cn = C(Xo[2]) + m.dcn;
    const C kp = C(cd->kp), kn = C(cd->kn);
    const C vpx = ux + kp * Eix, vpy = uy + kp * Eiy;
    const C vnx = ux + kn * Eix, vny = uy + kn * Eiy;
    const C om_p = C(-0.5) * (vpx * vpx + vpy * vpy) * ics;
    const C om_n = C(-0.5) * (vnx * vnx + vny * vny) * ics;
    const C om_f = C(-0.5) * (ux * ux + uy * uy) * ics;
    collide_ion(q + 9, m.dcp, cp, vpx, vpy, C(cd->om_pp), C(cd->om_mp), ics, ticf, om_p);
    collide_ion(q + 18, m.dcn, cn, vnx, vny, C(cd->om_pn), C(cd->om_mn), ics, ticf,
                (cd->flags & FLAG_G0_FLUID_U) ? om_f : om_n);
  }
  if (bottom || top) {
    // full-way bounce-back of the pre-collision state, f0 unchanged
    // (LBM.cu:816-846); moving-wall terms on the top wall only
    const C ms = top ? C(cd->multis) : C(0), md = top ? C(cd->multid) : C(0);
    const C f1 = q[1], f2 = q[2], f3 = q[3], f4 = q[4];
    const C f5 = q[5], f6 = q[6], f7 = q[7], f8 = q[8];
    q[3] = f1 - ms;
    q[4] = f2;
    q[1] = f3 + ms;
    q[2] = f4;
    q[7] = f5 - md;
    q[8] = f6 + md;
    q[5] = f7 + md;
    q[6] = f8 - md;
  } else {
    collide_fluid(q, m.dr, m.rho, ux, uy, m.Fx, m.Fy, C(cd->om_pf), C(cd->om_mf), ics, icfl,
                  C(cd->src_p * cd->dt), C(cd->src_m * cd->dt));
  }
#pragma unroll
  for (int s = 0; s < 3; ++s) {
    const C j = REF ? C(rf.jo.J[s] * tc) : C(0);
#pragma unroll
    for (int d = 0; d < 9; ++d) st_pop(Q + (9 * s + d) * plane + node, (S)to_stored(q[9 * s + d], d, j));
  }
}

// ---------------------------------------------------------------------------
// k_lbm_rows: k_lbm for FP32 storage with FP64 arithmetic, for channels whose
// rows are whole work-groups (nx a multiple of LBM_WG), after the
// pre-collision (P') state.  Same operations in the same order as k_lbm, so
// the same results, with fewer registers: the row index, the reference block and
// the reference records are wave-uniform (scalar loads), and the 27 values
// read stay FP32, each species being widened to its 9 FP64 rest deviations
// only while its moments or its collision are computed (and stored right
// after its collision).  Wall rows take k_lbm's wall rules: the y = 0
// velocity mirror, the ions' anti-bounce-back of the node's own
// post-collision values and the fluid's full-way bounce-back.
// ---------------------------------------------------------------------------
// The references of a wave's 64 nodes, block k of row y (nx a multiple of
// 64): the block records k-1, k, k+1 of rows y-1, y, y+1 and those rows'
// records, all wave-uniform.
// Wave-uniform reference values are read through the constant address
// space: the step kernels never write the reference tables, and uniform
// constant loads become scalar loads into SGPRs (generic pointers would take
// vector loads, two VGPRs per value).
typedef const __attribute__((address_space(4))) double cdouble;
typedef const __attribute__((address_space(4))) float cfloat;
__device__ __forceinline__ cdouble* as_const(const double* p) { return (cdouble*)p; }
__device__ __forceinline__ cfloat* as_const(const float* p) { return (cfloat*)p; }

struct WaveRefs {
  cfloat* B[3];   // block records of rows y-1, y, y+1
  cdouble* R[3];  // row records of rows y-1, y, y+1
  int km, k, kp;  // blocks k-1, k, k+1 (periodic)
  int lane;
  __device__ double X(int r, int blk, int s) const { return (double)B[r][BLK_STRIDE * blk + s]; }
  __device__ double own(int s) const { return X(1, k, s); }
  __device__ double J(int r, int s) const { return R[r][3 + s]; }
  // X(src) - X(node) for the 9 gather sources (NodeRefs::diffs): the
  // sources of lane 0 in column x-1 and of lane 63 in column x+1 lie in the
  // neighbouring blocks, every other source in the wave's own block
  __device__ void diffs(int s, double* d) const {
    const double o = own(s);
    const bool l0 = lane == 0, l63 = lane == REF_BLOCK - 1;
    const double bm = X(0, k, s) - o, bp = X(2, k, s) - o;
    d[0] = 0.0;
    d[1] = l0 ? X(1, km, s) - o : 0.0;
    d[2] = bm;
    d[3] = l63 ? X(1, kp, s) - o : 0.0;
    d[4] = bp;
    d[5] = l0 ? X(0, km, s) - o : bm;
    d[6] = l63 ? X(0, kp, s) - o : bm;
    d[7] = l63 ? X(2, kp, s) - o : bp;
    d[8] = l0 ? X(2, km, s) - o : bp;
  }
};
__device__ __forceinline__ WaveRefs wave_refs(const StepBuffers& b, const ChanDesc* __restrict__ cd, int y, int k,
                                              int lane) {
  const int ny = cd->ny, nb = cd->nb;
  const int ym = (y == 0) ? ny - 1 : y - 1;
  const int yp = (y == ny - 1) ? 0 : y + 1;
  const float* B = b.blocks + cd->blk_off;
  const double* R = b.refs + cd->ref_off;
  const long long bs = (long long)BLK_STRIDE * nb;
  return WaveRefs{{as_const(B + ym * bs), as_const(B + y * bs), as_const(B + yp * bs)},
                  {as_const(R + (long long)REF_STRIDE * ym), as_const(R + (long long)REF_STRIDE * y),
                   as_const(R + (long long)REF_STRIDE * yp)},
                  k == 0 ? nb - 1 : k - 1, k, k + 1 == nb ? 0 : k + 1, lane};
}

// FP64 rest deviations of species sp from its gathered FP32 values (load_pre)
__device__ __forceinline__ void widen_species(const float* v, const WaveRefs& rf, int sp, double tc, double* q) {
#pragma unroll
  for (int d = 0; d < 9; ++d) q[d] = (double)v[d];
  double dX[9];
  rf.diffs(sp, dX);
  add_src_refs(q, dX, rf.J(1, sp) * tc, rf.J(0, sp) * tc, rf.J(2, sp) * tc);
}

// wall node, one ion species: anti-bounce-back of its own stored
// post-collision values p (abb_own)
__device__ __forceinline__ void abb_species(const float* p, double dcw, double j, double* q) {
  const double two_s = (2.0 * WS) * dcw, two_d = (2.0 * WD) * dcw;
  const double js = WS * j, jd = WD * j;
  q[0] = -(double)p[0] + (2.0 * W0) * dcw;
  q[3] = -((double)p[1] + js) + two_s;
  q[4] = -(double)p[2] + two_s;
  q[1] = -((double)p[3] - js) + two_s;
  q[2] = -(double)p[4] + two_s;
  q[7] = -((double)p[5] + jd) + two_d;
  q[8] = -((double)p[6] - jd) + two_d;
  q[5] = -((double)p[7] - jd) + two_d;
  q[6] = -((double)p[8] + jd) + two_d;
}

// the 9 values of species s at row y (s = 0: gathered; ions: gathered, or on
// a wall row anti-bounced from the node's own values, dcw = c_w - X_s)
__device__ __forceinline__ void species_q(const float* v, const WaveRefs& rf, int s, double tc, bool wall,
                                          double dcw, double* q) {
  if (s > 0 && wall) abb_species(v, dcw, rf.J(1, s) * tc, q);
  else widen_species(v, rf, s, tc, q);
}

// 9 FP32 values of species block B gathered for node x of row y (pull,
// periodic in x and y; rm, r0, rp are the row offsets of y-1, y, y+1, which
// fit 32 bits: a channel holds at most MAX_NX x (MAX_INTERIOR_ROWS + 2) nodes)
__device__ __forceinline__ void gather9(const float* __restrict__ B, long long plane, int rm, int r0, int rp, int x,
                                        int xm, int xp, float* v) {
  v[0] = ld_pop(B + 0 * plane + (r0 + x));
  v[1] = ld_pop(B + 1 * plane + (r0 + xm));
  v[2] = ld_pop(B + 2 * plane + (rm + x));
  v[3] = ld_pop(B + 3 * plane + (r0 + xp));
  v[4] = ld_pop(B + 4 * plane + (rp + x));
  v[5] = ld_pop(B + 5 * plane + (rm + xm));
  v[6] = ld_pop(B + 6 * plane + (rm + xp));
  v[7] = ld_pop(B + 7 * plane + (rp + xp));
  v[8] = ld_pop(B + 8 * plane + (rp + xm));
}

// one node of a row tile (WALL: y = 0 or ny-1); a separate instantiation per
// case keeps the interior code free of the wall rules' registers.  macro (the
// step writes the macroscopic planes) is a run-time, wave-uniform flag: the
// compiler allocates the interior kernel in 118 VGPRs (4 waves per SIMD) this
// way, and in 136 (3 waves) when the flag is a template parameter and false.
template <bool WALL>
__device__ __forceinline__ void lbm_row_node(const StepBuffers& b, const ChanDesc* __restrict__ cd, int y, int x,
                                             const WaveRefs& rf, const float* __restrict__ P, float* __restrict__ Q,
                                             bool macro) {
  using C = double;
  const int nx = cd->nx, ny = cd->ny;
  const long long node = (long long)y * nx + x;
  const long long plane = cd->plane;
  const double* __restrict__ ps = b.psi + cd->fld_off;
  const double tc = cd->inv_cs2 * cd->inv_cfl;
  const int xm = (x == 0) ? nx - 1 : x - 1;
  const int xp = (x == nx - 1) ? 0 : x + 1;
  const bool bottom = WALL && (y == 0), top = WALL && (y == ny - 1);
  constexpr bool wall = WALL;
  // fluid: pull gather (LBM.cu:869-905, x and y wrap); ions: the same, or
  // on a wall row the node's own values (LBM.cu:915-1023)
  float v[27];
  {
    const int r0 = y * nx;
    const int rm = (bottom ? ny - 1 : y - 1) * nx, rp = (top ? 0 : y + 1) * nx;
    gather9(P, plane, rm, r0, rp, x, xm, xp, v);
    if (wall) {
#pragma unroll
      for (int k = 9; k < 27; ++k) v[k] = P[(long long)k * plane + node];
    } else {
#pragma unroll
      for (int s = 1; s < 3; ++s) gather9(P + (long long)(9 * s) * plane, plane, rm, r0, rp, x, xm, xp, v + 9 * s);
    }
  }
  // the node's own references X(x, y)
  double Xo[3];
#pragma unroll
  for (int s = 0; s < 3; ++s) Xo[s] = rf.own(s);
  double dcwp = 0.0, dcwn = 0.0;
  if (wall) {
    const double* __restrict__ Wl = b.walls + cd->wall_off + x;
    const long long ws = b.walls_stride;
    dcwp = (bottom ? Wl[W_CP_LO * ws] : Wl[W_CP_HI * ws]) - Xo[1];
    dcwn = (bottom ? Wl[W_CN_LO * ws] : Wl[W_CN_HI * ws]) - Xo[2];
  }
  double Ex, Ey;
  efield_at(ps, nx, ny, x, xm, xp, y, cd->inv2dx, cd->inv2dy, Ex, Ey);
  // y = 0: the mirror's inputs, node (x,1)'s pre-collision values and field,
  // loaded together with the node's own
  float w[27];
  double Ex1 = 0.0, Ey1 = 0.0;
  if (bottom) {
    const int r0 = nx, rm = 0, rp = 2 * nx;
#pragma unroll
    for (int s = 0; s < 3; ++s) gather9(P + (long long)(9 * s) * plane, plane, rm, r0, rp, x, xm, xp, w + 9 * s);
    efield_at(ps, nx, ny, x, xm, xp, 1, cd->inv2dx, cd->inv2dy, Ex1, Ey1);
  }
  // every load is in flight before the first use: keep the scheduler from
  // interleaving the species' loads with the arithmetic
  __builtin_amdgcn_sched_barrier(0);

  // y = 0: velocity mirrored from node (x,1)'s pre-collision state, force
  // included (LBM.cu:402-443)
  C mux = 0, muy = 0;
  if (bottom) {
    const WaveRefs r1 = wave_refs(b, cd, 1, rf.k, rf.lane);
    C q[9], dr, jx, jy, dcp, dcn;
    widen_species(w, r1, 0, tc, q);
    dr = ((q[0] + q[1]) + (q[2] + q[3])) + ((q[4] + q[5]) + (q[6] + q[7])) + q[8];
    jx = (q[1] + q[5] + q[8]) - (q[3] + q[6] + q[7]);
    jy = (q[2] + q[5] + q[6]) - (q[4] + q[7] + q[8]);
    widen_species(w + 9, r1, 1, tc, q);
    dcp = ((q[0] + q[1]) + (q[2] + q[3])) + ((q[4] + q[5]) + (q[6] + q[7])) + q[8];
    widen_species(w + 18, r1, 2, tc, q);
    dcn = ((q[0] + q[1]) + (q[2] + q[3])) + ((q[4] + q[5]) + (q[6] + q[7])) + q[8];
    const double x0 = r1.own(0), x1 = r1.own(1), x2 = r1.own(2);
    const C rho = C(x0) + dr;
    const C rhoe = (dcp - dcn) + C(x1 - x2);
    const C Fx = C(cd->fcoef) * rhoe * C(Ex1 + cd->e_app) + C(cd->exf);
    const C Fy = C(cd->fcoef) * rhoe * C(Ey1);
    const C irho = C(1.0) / rho;
    const C icfl = C(cd->inv_cfl), hdt = C(cd->half_dt);
    mux = -((jx * icfl + Fx * hdt) * irho);
    muy = -((jy * icfl + Fy * hdt) * irho);
  }

  // moments, species by species (the expressions of moments() and of the
  // charge-density sums of k_lbm)
  Moments<C> m;
  C jx, jy, sp = 0.0, sn = 0.0;
  {
    C q[9];
    widen_species(v, rf, 0, tc, q);
    m.dr = ((q[0] + q[1]) + (q[2] + q[3])) + ((q[4] + q[5]) + (q[6] + q[7])) + q[8];
    jx = (q[1] + q[5] + q[8]) - (q[3] + q[6] + q[7]);
    jy = (q[2] + q[5] + q[6]) - (q[4] + q[7] + q[8]);
    species_q(v + 9, rf, 1, tc, wall, dcwp, q);
    m.dcp = ((q[0] + q[1]) + (q[2] + q[3])) + ((q[4] + q[5]) + (q[6] + q[7])) + q[8];
#pragma unroll
    for (int k = 0; k < 9; ++k) sp += q[k];
    species_q(v + 18, rf, 2, tc, wall, dcwn, q);
    m.dcn = ((q[0] + q[1]) + (q[2] + q[3])) + ((q[4] + q[5]) + (q[6] + q[7])) + q[8];
#pragma unroll
    for (int k = 0; k < 9; ++k) sn += q[k];
  }
  m.rho = C(Xo[0]) + m.dr;
  m.rhoe = (m.dcp - m.dcn) + C(Xo[1] - Xo[2]);
  m.Fx = C(cd->fcoef) * m.rhoe * C(Ex + cd->e_app) + C(cd->exf);
  m.Fy = C(cd->fcoef) * m.rhoe * C(Ey);
  {
    const C irho = C(1.0) / m.rho;
    const C icfl = C(cd->inv_cfl), hdt = C(cd->half_dt);
    m.ux = (jx * icfl + m.Fx * hdt) * irho;
    m.uy = (jy * icfl + m.Fy * hdt) * irho;
  }
  const C ux = bottom ? mux : m.ux, uy = bottom ? muy : m.uy;
  b.rho_e[cd->fld_off + node] = (sp - sn) + (Xo[1] - Xo[2]);
  if (macro) {
    const long long o = cd->fld_off + node, st = b.macro_stride;
    b.macro[M_RHO * st + o] = (double)m.rho;
    b.macro[M_UX * st + o] = (double)ux;
    b.macro[M_UY * st + o] = (double)uy;
    b.macro[M_CP * st + o] = (double)(C(Xo[1]) + m.dcp);
    b.macro[M_CN * st + o] = (double)(C(Xo[2]) + m.dcn);
  }

  const C ics = C(cd->inv_cs2), icfl = C(cd->inv_cfl), ticf = ics * icfl;
  // ions: drift velocity u + K E (LBM.cu:479-488), collide, store
  {
    const C Eix = C(Ex + cd->drift_app), Eiy = C(Ey);
    const C kp = C(cd->kp), kn = C(cd->kn);
    const C om_f = C(-0.5) * (ux * ux + uy * uy) * ics;
    C q[9];
    {
      const C cp = C(Xo[1]) + m.dcp;
      const C vpx = ux + kp * Eix, vpy = uy + kp * Eiy;
      const C om_p = C(-0.5) * (vpx * vpx + vpy * vpy) * ics;
      species_q(v + 9, rf, 1, tc, wall, dcwp, q);
      collide_ion(q, m.dcp, cp, vpx, vpy, C(cd->om_pp), C(cd->om_mp), ics, ticf, om_p);
      const C j = C(rf.J(1, 1) * tc);
#pragma unroll
      for (int d = 0; d < 9; ++d) st_pop(Q + (9 + d) * plane + node, (float)to_stored(q[d], d, j));
    }
    {
      const C cn = C(Xo[2]) + m.dcn;
      const C vnx = ux + kn * Eix, vny = uy + kn * Eiy;
      const C om_n = C(-0.5) * (vnx * vnx + vny * vny) * ics;
      species_q(v + 18, rf, 2, tc, wall, dcwn, q);
      collide_ion(q, m.dcn, cn, vnx, vny, C(cd->om_pn), C(cd->om_mn), ics, ticf,
                  (cd->flags & FLAG_G0_FLUID_U) ? om_f : om_n);
      const C j = C(rf.J(1, 2) * tc);
#pragma unroll
      for (int d = 0; d < 9; ++d) st_pop(Q + (18 + d) * plane + node, (float)to_stored(q[d], d, j));
    }
  }
  {
    C q[9];
    widen_species(v, rf, 0, tc, q);
    if (wall) {
      // full-way bounce-back of the pre-collision state, f0 unchanged
      // (LBM.cu:816-846); moving-wall terms on the top wall only
      const C ms = top ? C(cd->multis) : C(0), md = top ? C(cd->multid) : C(0);
      const C f1 = q[1], f2 = q[2], f3 = q[3], f4 = q[4];
      const C f5 = q[5], f6 = q[6], f7 = q[7], f8 = q[8];
      q[3] = f1 - ms;
      q[4] = f2;
      q[1] = f3 + ms;
      q[2] = f4;
      q[7] = f5 - md;
      q[8] = f6 + md;
      q[5] = f7 + md;
      q[6] = f8 - md;
    } else {
      collide_fluid(q, m.dr, m.rho, ux, uy, m.Fx, m.Fy, C(cd->om_pf), C(cd->om_mf), ics, icfl,
                    C(cd->src_p * cd->dt), C(cd->src_m * cd->dt));
    }
    const C j = C(rf.J(1, 0) * tc);
#pragma unroll
    for (int d = 0; d < 9; ++d) st_pop(Q + d * plane + node, (float)to_stored(q[d], d, j));
  }
}

// Channels with nx a multiple of 64: the 64 nodes of a wave lie in one row
// and one reference block, so the row index and its references are
// wave-uniform (scalar).  WALL: this
// launch updates the waves on wall rows, the other one the waves on interior
// rows (one kernel holding both paths would need the registers of both); a
// tile holding both kinds of row is in both launches.
// Tile shapes: wrows = 0, the work-group's waves take consecutive 64-node
// blocks from the tile's first node on (across row ends); wrows = 2 or 4,
// the tile is wrows rows x (4 / wrows) blocks whose first node is the tile's
// (waves past a row's end or past the last row return).
template <bool WALL>
__device__ __forceinline__ void lbm_rows_tile(const StepBuffers& b, int2 tile, const float* __restrict__ src,
                                              float* __restrict__ dst, bool macro, int wrows) {
  const ChanDesc* __restrict__ cd = b.chans + tile.x;
  if (!cd->active) return;  // converged channel, frozen by the residual driver
  const int nx = cd->nx, ny = cd->ny;
  const int w = (int)threadIdx.x >> 6;
  int node0 = tile.y + 64 * w;  // the wave's first node
  if (wrows > 0) {
    const int bpt = (LBM_WG / 64) / wrows, r = w / bpt, bo = w - r * bpt;
    if (tile.y % nx + 64 * bo >= nx) return;  // a block past the row's end
    node0 = tile.y + r * nx + 64 * bo;
  }
  if (node0 >= nx * ny) return;  // whole waves past the channel's last row
  const int y = __builtin_amdgcn_readfirstlane(node0 / nx);
  if ((y == 0 || y == ny - 1) != WALL) return;         // the other launch's wave
  const int x = node0 + ((int)threadIdx.x & 63) - y * nx;
  const int k = __builtin_amdgcn_readfirstlane((node0 - y * nx) >> 6);  // the wave's reference block
  const float* __restrict__ P = (cd->swap ? dst : src) + cd->pop_off;
  float* __restrict__ Q = (cd->swap ? const_cast<float*>(src) : dst) + cd->pop_off;
  const WaveRefs rf = wave_refs(b, cd, y, k, (int)threadIdx.x & 63);
  lbm_row_node<WALL>(b, cd, y, x, rf, P, Q, macro);
}

template <bool WALL>
__global__ __launch_bounds__(LBM_WG) void k_lbm_rows(StepBuffers b, const int2* __restrict__ tiles,
                                                    const float* __restrict__ src, float* __restrict__ dst,
                                                    int macro) {
  lbm_rows_tile<WALL>(b, tiles[xcd_remap(blockIdx.x, gridDim.x)], src, dst, macro != 0, WALL ? 0 : b.lbm_wrows);
}



// ---------------------------------------------------------------------------
// k_init_eq: equilibrium populations of the stored fields, unshifted
// (gpu_init_equilibrium, LBM.cu:195-305).
// ---------------------------------------------------------------------------
template <typename S>
__global__ __launch_bounds__(LBM_WG) void k_init_eq(StepBuffers b, S* __restrict__ dst) {

  const int2 tile = b.lbm_tiles[blockIdx.x];
  const ChanDesc* __restrict__ cd = b.chans + tile.x;
  const int nx = cd->nx, ny = cd->ny;
  const int node = tile.y + (int)threadIdx.x;
  if (node >= nx * ny) return;
  const int y = node / nx, x = node - (node / nx) * nx;
  const long long plane = cd->plane;
  const long long o = cd->fld_off + node, st = b.macro_stride;
  const double rho = b.macro[M_RHO * st + o], ux = b.macro[M_UX * st + o];
  const double uy = b.macro[M_UY * st + o], cp = b.macro[M_CP * st + o];
  const double cn = b.macro[M_CN * st + o];
  const int xm = (x == 0) ? nx - 1 : x - 1, xp = (x == nx - 1) ? 0 : x + 1;
  double Ex, Ey;
  efield_at(b.psi + cd->fld_off, nx, ny, x, xm, xp, y, cd->inv2dx, cd->inv2dy, Ex, Ey);
  Ex += cd->drift_app;
  const double ics = cd->inv_cs2, ticf = cd->inv_cs2 * cd->inv_cfl;
  const double vel[3][2] = {{ux, uy}, {ux + cd->kp * Ex, uy + cd->kp * Ey},
                            {ux + cd->kn * Ex, uy + cd->kn * Ey}};
  const double X[3] = {rho, cp, cn};
  constexpr bool REF = std::is_same<S, float>::value;
  NodeRefs rf{};
  if (REF) rf = node_refs(b, cd, x, y, xm, xp);
  S* __restrict__ Q = dst + cd->pop_off;
#pragma unroll
  for (int s = 0; s < 3; ++s) {
    const double j = REF ? rf.jo.J[s] * ticf : 0.0;
    const double vx = vel[s][0], vy = vel[s][1];
    const double om = -0.5 * (vx * vx + vy * vy) * ics;
    const double tx = vx * ticf, ty = vy * ticf;
    const double dX = X[s] - (REF ? rf.own(s) : 0.0);
    const double t[9] = {0.0, tx, ty, -tx, -ty, tx + ty, ty - tx, -(tx + ty), tx - ty};
#pragma unroll
    for (int d = 0; d < 9; ++d) {
      const double w = (d == 0) ? W0 : (d < 5 ? WS : WD);
      const double eq = w * (dX + X[s] * (om + t[d] * (1.0 + 0.5 * t[d])));
      Q[(9 * s + d) * plane + node] = (S)to_stored(eq, d, j);
    }
  }
}

// ---------------------------------------------------------------------------
// k_to_q: pre-collision populations of the next step, written unshifted.
// ---------------------------------------------------------------------------
template <typename S>
__global__ __launch_bounds__(LBM_WG) void k_to_q(StepBuffers b, const S* __restrict__ src,
                                                 S* __restrict__ dst) {
  using C = typename Compute<S>::T;
  const int2 tile = b.lbm_tiles[blockIdx.x];
  const ChanDesc* __restrict__ cd = b.chans + tile.x;
  const int nx = cd->nx, ny = cd->ny;
  const int node = tile.y + (int)threadIdx.x;
  if (node >= nx * ny) return;
  const int y = node / nx, x = node - (node / nx) * nx;
  const long long plane = cd->plane;
  const S* __restrict__ P = (cd->swap ? dst : src) + cd->pop_off;
  const int xm = (x == 0) ? nx - 1 : x - 1, xp = (x == nx - 1) ? 0 : x + 1;
  constexpr bool REF = std::is_same<S, float>::value;
  NodeRefs rf{};
  if (REF) rf = node_refs(b, cd, x, y, xm, xp);
  const double tc = cd->inv_cs2 * cd->inv_cfl;
  C q[27];
  load_pre<S, C, false, REF>(P, plane, nx, ny, x, y, xm, xp, rf, tc, q);
  if (y == 0 || y == ny - 1) {
    const double* __restrict__ Wl = b.walls + cd->wall_off + x;
    const long long ws = b.walls_stride;
    const double cwp = (y == 0) ? Wl[W_CP_LO * ws] : Wl[W_CP_HI * ws];
    const double cwn = (y == 0) ? Wl[W_CN_LO * ws] : Wl[W_CN_HI * ws];
    const double X1 = REF ? rf.own(1) : 0.0, X2 = REF ? rf.own(2) : 0.0;
    abb_own<S, C>(P + 9 * plane, plane, node, C(cwp - X1), REF ? C(rf.jo.J[1] * tc) : C(0), q + 9);
    abb_own<S, C>(P + 18 * plane, plane, node, C(cwn - X2), REF ? C(rf.jo.J[2] * tc) : C(0), q + 18);
  }
  S* __restrict__ Q = (cd->swap ? const_cast<S*>(src) : dst) + cd->pop_off;
#pragma unroll
  for (int s = 0; s < 3; ++s) {
    const C j = REF ? C(rf.jo.J[s] * tc) : C(0);
#pragma unroll
    for (int d = 0; d < 9; ++d) Q[(9 * s + d) * plane + node] = (S)to_stored(q[9 * s + d], d, j);
  }
}

// ---------------------------------------------------------------------------
// complex helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }

// ---------------------------------------------------------------------------
// Twiddles: LDS holds e^{-2 pi i j/nx} for j < nx/4 (one quadrant); the rest
// of [0, nx/2] follows from e^{-2 pi i (j + nx/4)/nx} = -i e^{-2 pi i j/nx}.
// ---------------------------------------------------------------------------
// Both helpers are branch-free: one table load at a clamped index, then
// selects and sign-bit flips (exact negations), so lanes of different
// quadrants do not split the wave.
__device__ __forceinline__ double neg_if(double v, bool neg) {
  return __longlong_as_double(__double_as_longlong(v) ^ ((long long)neg << 63));
}

__device__ __forceinline__ double2 tw_half(const double2* __restrict__ q, int j, int quarter) {
  // j < nx/4: q[j]; j < nx/2: -i q[j - nx/4]; j == nx/2: -1
  const bool up = j >= quarter;
  const int jj = up ? j - quarter : j;
  const double2 w = q[min(jj, quarter - 1)];
  const bool last = jj == quarter;
  const double re = up ? w.y : w.x, im = up ? neg_if(w.x, true) : w.y;
  return make_double2(last ? -1.0 : re, last ? 0.0 : im);
}

// ---------------------------------------------------------------------------
// Register DFTs of size 2, 4, 8, 16 (natural order in and out), sign SIGN in
// the exponent e^{SIGN 2 pi i jk/R}.  8 = 2 x 4 and 16 = 4 x 4 Cooley-Tukey
// with the inner twiddles W_R^{j2 u1} as constants.
// ---------------------------------------------------------------------------
// v * (SIGN i)
template <int SIGN>
__device__ __forceinline__ double2 mul_si(double2 v) {
  return SIGN < 0 ? make_double2(v.y, -v.x) : make_double2(-v.y, v.x);
}
// v * (c + SIGN i s)
template <int SIGN>
__device__ __forceinline__ double2 mul_cs(double2 v, double c, double sn) {
  const double ss = SIGN < 0 ? -sn : sn;
  return make_double2(v.x * c - v.y * ss, v.x * ss + v.y * c);
}
template <int SIGN>
__device__ __forceinline__ void dft4(double2& a, double2& b, double2& c, double2& d) {
  const double2 apc = cadd(a, c), amc = csub(a, c), bpd = cadd(b, d), jb = mul_si<SIGN>(csub(b, d));
  a = cadd(apc, bpd);
  b = cadd(amc, jb);
  c = csub(apc, bpd);
  d = csub(amc, jb);
}

constexpr double kC8 = 0.92387953251128675613;   // cos(pi/8)
constexpr double kS8 = 0.38268343236508977173;   // sin(pi/8)
constexpr double kH = 0.70710678118654752440;    // cos(pi/4)

template <int SIGN, int R>
__device__ __forceinline__ void dft_reg(double2* v) {
  if (R == 2) {
    const double2 a = v[0];
    v[0] = cadd(a, v[1]);
    v[1] = csub(a, v[1]);
  } else if (R == 4) {
    dft4<SIGN>(v[0], v[1], v[2], v[3]);
  } else if (R == 8) {
    // j = 4 j1 + j2, u = u1 + 2 u2
#pragma unroll
    for (int j2 = 0; j2 < 4; ++j2) {
      const double2 a = v[j2];
      v[j2] = cadd(a, v[4 + j2]);
      v[4 + j2] = csub(a, v[4 + j2]);
    }
    v[5] = mul_cs<SIGN>(v[5], kH, kH);
    v[6] = mul_si<SIGN>(v[6]);
    v[7] = mul_cs<SIGN>(v[7], -kH, kH);
    dft4<SIGN>(v[0], v[1], v[2], v[3]);
    dft4<SIGN>(v[4], v[5], v[6], v[7]);
    // X[u1 + 2 u2] = v[4 u1 + u2]
    const double2 t1 = v[1], t2 = v[2], t3 = v[3], t4 = v[4], t5 = v[5], t6 = v[6];
    v[1] = t4;
    v[2] = t1;
    v[3] = t5;
    v[4] = t2;
    v[5] = t6;
    v[6] = t3;
  } else {
    // R == 16: j = 4 j1 + j2, u = u1 + 4 u2
#pragma unroll
    for (int j2 = 0; j2 < 4; ++j2) dft4<SIGN>(v[j2], v[4 + j2], v[8 + j2], v[12 + j2]);
    // v[4 u1 + j2] *= W16^(u1 j2)
    v[5] = mul_cs<SIGN>(v[5], kC8, kS8);    // 1
    v[6] = mul_cs<SIGN>(v[6], kH, kH);      // 2
    v[7] = mul_cs<SIGN>(v[7], kS8, kC8);    // 3
    v[9] = mul_cs<SIGN>(v[9], kH, kH);      // 2
    v[10] = mul_si<SIGN>(v[10]);            // 4
    v[11] = mul_cs<SIGN>(v[11], -kH, kH);   // 6
    v[13] = mul_cs<SIGN>(v[13], kS8, kC8);  // 3
    v[14] = mul_cs<SIGN>(v[14], -kH, kH);   // 6
    v[15] = mul_cs<SIGN>(v[15], -kC8, -kS8);  // 9
#pragma unroll
    for (int u1 = 0; u1 < 4; ++u1) dft4<SIGN>(v[4 * u1], v[4 * u1 + 1], v[4 * u1 + 2], v[4 * u1 + 3]);
    // X[u1 + 4 u2] = v[4 u1 + u2]: a transpose of the 4 x 4 block
    double2 t[16];
#pragma unroll
    for (int i = 0; i < 16; ++i) t[i] = v[i];
#pragma unroll
    for (int u = 0; u < 16; ++u) v[u] = t[4 * (u & 3) + (u >> 2)];
  }
}

// e^{SIGN 2 pi i t/nx} for 0 <= t < nx from w = q[tw_quad_index(t, lgnx)],
// the quadrant-table value of t
__device__ __forceinline__ int tw_quad_index(int t, int lgnx) { return t & ((1 << (lgnx - 2)) - 1); }

template <int SIGN>
__device__ __forceinline__ double2 tw_quad_resolve(double2 w, int t, int lgnx) {
  // t in [0, nx): quadrant 0..3 multiplies the table value by (-i)^quad:
  // (x, y), (y, -x), (-x, -y), (-y, x); SIGN > 0 conjugates
  const int quad = t >> (lgnx - 2);
  const bool swap = quad & 1;
  const double a = swap ? w.y : w.x, b = swap ? w.x : w.y;
  const bool nre = (quad & 2) != 0;
  const bool nim = (((quad ^ (quad >> 1)) & 1) != 0) != (SIGN > 0);
  return make_double2(neg_if(a, nre), neg_if(b, nim));
}

// e^{SIGN 2 pi i t/nx} for 0 <= t < nx from the quadrant table
template <int SIGN>
__device__ __forceinline__ double2 tw_full(const double2* __restrict__ q, int t, int lgnx) {
  return tw_quad_resolve<SIGN>(q[tw_quad_index(t, lgnx)], t, lgnx);
}

// Two-level twiddle table, for rows whose quadrant table (nx / 4 values)
// would cost a resident work-group of LDS: F[f] = W^f for f < 64, then
// C[c] = W^(64 c) for c < max(1, nx / 64), W = e^{-2 pi i / nx}, both taken
// from the quadrant table; W^t = C[t >> 6] F[t & 63], one complex product
// (exact where either factor is 1).  tw2_count(nx) values.

template <int WGT>
__device__ __forceinline__ void tw2_fill(double2* t2, const double2* __restrict__ q, int nx, int lgnx) {
  const int n = tw2_count(nx);
  for (int i = threadIdx.x; i < n; i += WGT) {
    const int t = i < TW2_FINE ? i : (i - TW2_FINE) << TW2_LGF;
    t2[i] = tw_full<-1>(q, t & (nx - 1), lgnx);
  }
}

// e^{SIGN 2 pi i t/nx}, 0 <= t < nx: quadrant table or two-level table
template <int SIGN, bool TW2>
__device__ __forceinline__ double2 tw_get(const double2* __restrict__ q, int t, int lgnx) {
  if (TW2) {
    const double2 w = cmul(q[TW2_FINE + (t >> TW2_LGF)], q[t & (TW2_FINE - 1)]);
    return SIGN > 0 ? cconj(w) : w;
  }
  return tw_full<SIGN>(q, t, lgnx);
}

// e^{-2 pi i j/nx} for 0 <= j <= nx / 2 (exactly -1 at nx / 2)
template <bool TW2>
__device__ __forceinline__ double2 tw_half_get(const double2* __restrict__ q, int j, int quarter, int lgnx) {
  if (TW2) return j == 2 * quarter ? make_double2(-1.0, 0.0) : tw_get<-1, true>(q, j, lgnx);
  return tw_half(q, j, quarter);
}

// LDS position of row element i: bits 0-3 XOR bits 4-7.  The Stockham
// passes write with strides 8 (first pass) and 8 x 16 (second) elements,
// which without the swizzle put 8 (resp. 2) of every 16 lanes of a
// ds_write_b128 on the same banks; with it every 16 consecutive lanes of
// every read and write pattern here cover 16 distinct 16-byte bank groups.
// A bijection within each 256-element block; every row access uses it.
__device__ __forceinline__ int sw(int i) { return i ^ ((i >> 4) & 15); }

// ---------------------------------------------------------------------------
// Stockham autosort FFT of `rows` rows of length N = 2^lgN (4 <= N, and
// rows * N <= FFT_MAX_ELEMS) held in LDS, in place.  Stage (n, s) with radix R,
// m = n / R: butterfly (p, q) reads x[q + s(p + j m)], j < R, takes the
// R-point DFT in registers, multiplies output u by W_n^{p u} and writes
// y[q + s(R p + u)]; a barrier separates the reads and the writes.  Radices:
// one pass of 2^(lgN mod 4) when lgN is not a multiple of 4, then radix 16.
// Each lane holds 16 values per pass.  twq is the quadrant table of e^{-2 pi i j/nx}.
// ---------------------------------------------------------------------------
template <int SIGN, int R, int WGT, bool TW2>
__device__ __forceinline__ void stockham_pass(double2* x, int rows, int lgN, int lgn, int lgs,
                                              const double2* __restrict__ twq, int lgnx) {
  constexpr int LGR = R == 16 ? 4 : (R == 8 ? 3 : (R == 4 ? 2 : 1));
  // butterflies per lane: rows N / R <= FFT_MAX_ELEMS / R of them over WGT lanes
  constexpr int BPT = (FFT_MAX_ELEMS / R + WGT - 1) / WGT;
  const int lgm = lgn - LGR, s = 1 << lgs;
  const int lgper = lgN - LGR;
  const int nb = rows << lgper;
  double2 v[BPT][R];
#pragma unroll
  for (int c = 0; c < BPT; ++c) {
    const int bf = threadIdx.x + c * WGT;
    if (bf < nb) {
      const int r = bf >> lgper, jj = bf & ((1 << lgper) - 1);
      const int p = jj >> lgs, q = jj & (s - 1);
      const int base = (r << lgN) + q + s * p;
#pragma unroll
      for (int j = 0; j < R; ++j) v[c][j] = x[sw(base + ((s * j) << lgm))];
    }
  }
  __syncthreads();
#pragma unroll
  for (int c = 0; c < BPT; ++c) {
    const int bf = threadIdx.x + c * WGT;
    if (bf < nb) {
      const int r = bf >> lgper, jj = bf & ((1 << lgper) - 1);
      const int p = jj >> lgs, q = jj & (s - 1);
      dft_reg<SIGN, R>(v[c]);
      const int ob = (r << lgN) + q + s * R * p;
      x[sw(ob)] = v[c][0];
      // W^(p u): table values for u = 1, 2, 4, 8, products (at most three
      // roundings deep) for the rest
      const int t1 = p << (lgnx - lgn);
      const double2 w1 = tw_get<SIGN, TW2>(twq, t1, lgnx);
      if (R == 2) {
        x[sw(ob + s)] = cmul(v[c][1], w1);
      } else {
        const double2 w2 = tw_get<SIGN, TW2>(twq, 2 * t1, lgnx);
        x[sw(ob + s)] = cmul(v[c][1], w1);
        x[sw(ob + 2 * s)] = cmul(v[c][2], w2);
        const double2 w3 = cmul(w2, w1);
        x[sw(ob + 3 * s)] = cmul(v[c][3], w3);
        if (R >= 8) {
          const double2 w4 = tw_get<SIGN, TW2>(twq, 4 * t1, lgnx);
          x[sw(ob + 4 * s)] = cmul(v[c][4], w4);
          x[sw(ob + 5 * s)] = cmul(v[c][5], cmul(w4, w1));
          x[sw(ob + 6 * s)] = cmul(v[c][6], cmul(w4, w2));
          x[sw(ob + 7 * s)] = cmul(v[c][7], cmul(w4, w3));
          if (R == 16) {
            const double2 w8 = tw_get<SIGN, TW2>(twq, 8 * t1, lgnx);
            x[sw(ob + 8 * s)] = cmul(v[c][8], w8);
            x[sw(ob + 9 * s)] = cmul(v[c][9], cmul(w8, w1));
            x[sw(ob + 10 * s)] = cmul(v[c][10], cmul(w8, w2));
            x[sw(ob + 11 * s)] = cmul(v[c][11], cmul(w8, w3));
            const double2 w12 = cmul(w8, w4);
            x[sw(ob + 12 * s)] = cmul(v[c][12], w12);
            x[sw(ob + 13 * s)] = cmul(v[c][13], cmul(w12, w1));
            x[sw(ob + 14 * s)] = cmul(v[c][14], cmul(w12, w2));
            x[sw(ob + 15 * s)] = cmul(v[c][15], cmul(w12, w3));
          }
        }
      }
    }
  }
  __syncthreads();
}

// WGT lanes; MAXR = 16 (16 values per lane per pass) or 8 (one more pass,
// half the registers: the 512-lane one-tile kernels)
template <int SIGN, int WGT, int MAXR, bool TW2 = false>
__device__ __forceinline__ void stockham(double2* x, int rows, int N, int lgN, const double2* __restrict__ twq, int lgnx) {
  constexpr int LGM = MAXR == 16 ? 4 : 3;
  int lgn = lgN, lgs = 0;
  const int r = lgN % LGM;
  if (r == 1) {
    stockham_pass<SIGN, 2, WGT, TW2>(x, rows, lgN, lgn, lgs, twq, lgnx);
    lgn -= 1;
    lgs += 1;
  } else if (r == 2) {
    stockham_pass<SIGN, 4, WGT, TW2>(x, rows, lgN, lgn, lgs, twq, lgnx);
    lgn -= 2;
    lgs += 2;
  } else if (r == 3) {
    stockham_pass<SIGN, 8, WGT, TW2>(x, rows, lgN, lgn, lgs, twq, lgnx);
    lgn -= 3;
    lgs += 3;
  }
  while (lgn >= LGM) {
    stockham_pass<SIGN, MAXR, WGT, TW2>(x, rows, lgN, lgn, lgs, twq, lgnx);
    lgn -= LGM;
    lgs += LGM;
  }
}

__device__ __forceinline__ int ilog2(int v) { return 31 - __builtin_clz(v); }

// Spectrum layout per channel: blocks of 8 wavenumbers, [k/8][y][k%8], so
// that one wall-normal column block (8 modes x all rows) is contiguous for
// the tridiagonal sweep and a row segment of 8 modes is one 128-byte line.
__device__ __forceinline__ long long spec_idx(int ny, int y, int k) {
  return ((long long)(k >> 3) * ny + y) * 8 + (k & 7);
}

// e / rows for e < 2^20, rows <= 2^12: (e + 1/2) / rows is at least 1/(2 rows)
// away from an integer, far above the float rounding of the product
__device__ __forceinline__ int div_rows(int e, float inv_rows) { return (int)(((float)e + 0.5f) * inv_rows); }

// dst[j] = src[j] for j < n <= MAX_NX / 4, 16-byte values, by LDS
// DMA (global_load_lds_dwordx4: lane l of a wave writes M0 + 16 l, M0 = the
// wave's base in dst).  Nothing passes through registers; the data is in LDS
// after the next s_waitcnt vmcnt(0), which the compiler places before the
// next work-group barrier.  dst must be 16-byte aligned.
template <int WGT>
__device__ __forceinline__ void lds_dma_copy(double2* dst, const double2* __restrict__ src, int n) {
  const int wave_base = threadIdx.x & ~63;
#pragma unroll
  for (int c = 0; c < MAX_NX / 4 / WGT; ++c) {
    const int j = threadIdx.x + c * WGT;
    if (c * WGT < n && j < n)
      __builtin_amdgcn_global_load_lds(src + j,
                                       (__attribute__((address_space(3))) void*)(dst + c * WGT + wave_base),
                                       16, 0, 0);
  }
}

// dst[x] = src[x] for x < n by one work-group of WGT lanes, 8 loads per lane in flight
template <int WGT>
__device__ __forceinline__ void copy_row_batched(double* __restrict__ dst, const double* __restrict__ src, int n) {
  for (int x0 = 0; x0 < n; x0 += 8 * WGT) {
    double v[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const int x = x0 + threadIdx.x + c * WGT;
      v[c] = x < n ? src[x] : 0.0;
    }
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const int x = x0 + threadIdx.x + c * WGT;
      if (x < n) dst[x] = v[c];
    }
  }
}

// right-hand side of the lifted wall-normal system, scaled by dy^2
// (odd_extension, poisson.cu:229-241)
__device__ __forceinline__ double rhs_at(const double* __restrict__ re, const double* __restrict__ Wl,
                                         long long ws, const ChanDesc* __restrict__ cd, int x,
                                         int y) {
  double v = -cd->pc_dy2 * re[(long long)y * cd->nx + x];
  if (y == 1) v -= Wl[W_PSI_LO * ws + x];
  if (y == cd->ny - 2) v -= Wl[W_PSI_HI * ws + x];
  return v;
}

// ---------------------------------------------------------------------------
// k_fft_fwd: rows of RHS -> spectrum [y][k], k = 0..nx/2
// ---------------------------------------------------------------------------
// TWL: stage the quadrant twiddle table in LDS (else read it from global)
// Every global load of a tile is issued before the first one is consumed
// (twiddles, rows, wall lifts: up to 16 + 8 + 16 values per lane in flight),
// so the load phase costs one memory latency, not one per element.
template <bool POW2, bool TWL>
__global__ __launch_bounds__(FFT_WG_ROWS) void k_fft_fwd(StepBuffers b, const int4* __restrict__ tiles,
                                                   const ChanDesc* __restrict__ chans, int tile_base) {
  extern __shared__ double2 lds[];
  const int4 t = tiles[tile_base + blockIdx.x];
  const ChanDesc* __restrict__ cd = chans + t.x;
  if (!cd->active) return;
  const int nx = cd->nx, nk = cd->nk, y0 = t.y, rows = t.z;
  const double* __restrict__ re = b.rho_e + cd->fld_off;
  const double* __restrict__ Wl = b.walls + cd->wall_off;
  const long long ws = b.walls_stride;
  const double2* __restrict__ tw = b.tw + cd->tw_off;
  double2* __restrict__ spec = b.spec + cd->spec_off;
  if (POW2) {
    const int N = nx >> 1, lgN = ilog2(N), lgnx = lgN + 1, quarter = nx >> 2;
    const int total = rows << lgN;  // <= FFT_MAX_ELEMS
    const double pc = cd->pc_dy2;
    // twiddles after the (swizzled) rows: the quadrant table, or without
    // it (TWL false) the two-level table, filled once the rows' loads are
    // in flight
    double2* twl = lds + ((total + 15) & ~15);
    if (TWL) lds_dma_copy<FFT_WG_ROWS>(twl, tw, quarter);
    // the tile's rows are contiguous in rho_e: element e = (r, n) holds
    // x[2n], x[2n+1] of row y0 + r as one 16-byte value (rows start 16-byte aligned)
    const double2* __restrict__ src = reinterpret_cast<const double2*>(re + (long long)y0 * nx);
    double2 v[FFT_ROW_SLOTS];
#pragma unroll
    for (int c = 0; c < FFT_ROW_SLOTS; ++c)
      if (c * FFT_WG_ROWS < total) v[c] = src[min((int)threadIdx.x + c * FFT_WG_ROWS, total - 1)];
    // wall lifts of rows 1 and ny-2 (odd_extension, poisson.cu:229-241)
    const bool wlo = (y0 == 1), whi = (y0 + rows - 1 == cd->ny - 2);
    double2 wv[FFT_ROW_SLOTS];
    if (wlo || whi) {
#pragma unroll
      for (int c = 0; c < FFT_ROW_SLOTS; ++c) {
        wv[c] = make_double2(0.0, 0.0);
        const int e = threadIdx.x + c * FFT_WG_ROWS;
        if (c * FFT_WG_ROWS < total && e < total) {
          const int r = e >> lgN, n = e & (N - 1);
          const bool lo = wlo && r == 0, hi = whi && r == rows - 1;
          if (lo || hi) {
            const double* w = Wl + (lo ? W_PSI_LO : W_PSI_HI) * ws + 2 * n;
            wv[c] = make_double2(w[0], w[1]);
          }
        }
      }
    }
    if (!TWL) tw2_fill<FFT_WG_ROWS>(twl, tw, nx, lgnx);
    const double2* twq = twl;
#pragma unroll
    for (int c = 0; c < FFT_ROW_SLOTS; ++c) {
      const int e = threadIdx.x + c * FFT_WG_ROWS;
      if (c * FFT_WG_ROWS < total && e < total) {
        double v0 = -pc * v[c].x, v1 = -pc * v[c].y;
        if (wlo || whi) {
          v0 -= wv[c].x;
          v1 -= wv[c].y;
        }
        lds[sw(e)] = make_double2(v0, v1);
      }
    }
    __syncthreads();
    stockham<-1, FFT_WG_ROWS, 8, !TWL>(lds, rows, N, lgN, twq, lgnx);
    // split the packed transform: X[k] = E[k] + W^k O[k]
    // rows vary fastest across lanes: 8 modes x `rows` rows are contiguous
    const float inv_rows = 1.0f / (float)rows;
    for (int e = threadIdx.x; e < rows * (N + 1); e += FFT_WG_ROWS) {
      const int k = div_rows(e, inv_rows), r = e - k * rows;
      const double2 zk = lds[sw((r << lgN) + (k & (N - 1)))];
      const double2 zc = cconj(lds[sw((r << lgN) + ((N - k) & (N - 1)))]);
      const double2 E = make_double2(0.5 * (zk.x + zc.x), 0.5 * (zk.y + zc.y));
      const double2 dd = csub(zk, zc);
      const double2 O = make_double2(0.5 * dd.y, -0.5 * dd.x);  // (zk - zc)/(2i)
      spec[spec_idx(cd->ny, y0 + r, k)] = cadd(E, cmul(tw_half_get<!TWL>(twq, k, quarter, lgnx), O));
    }
  } else {
    // direct real DFT of the tile's rows: their right-hand sides in LDS
    // (row stride ldv), then one output (row, mode) per lane in the order of
    // the spectrum layout, 8 modes of a row then the next row, so that a
    // wave's stores are contiguous
    const int ldv = nx | 1;
    double* v = reinterpret_cast<double*>(lds);
    for (int e = threadIdx.x; e < rows * nx; e += FFT_WG_ROWS) {
      const int r = e / nx, n = e - r * nx;
      v[r * ldv + n] = rhs_at(re, Wl, ws, cd, n, y0 + r);
    }
    __syncthreads();
    const int nblk = (nk + 7) >> 3;
    for (int e = threadIdx.x; e < nblk * rows * 8; e += FFT_WG_ROWS) {
      const int kb = e / (rows * 8), rem = e - kb * rows * 8, r = rem >> 3, k = kb * 8 + (rem & 7);
      if (k >= nk) continue;
      const double* __restrict__ vr = v + r * ldv;
      double2 acc = make_double2(0.0, 0.0);
      int idx = 0;
      for (int n = 0; n < nx; ++n) {
        const double2 w = tw[idx];
        acc.x += vr[n] * w.x;
        acc.y += vr[n] * w.y;
        idx += k;
        if (idx >= nx) idx -= nx;
      }
      spec[spec_idx(cd->ny, y0 + r, k)] = acc;
    }
  }
}

// ---------------------------------------------------------------------------
// Persistent, software-pipelined power-of-two x-FFTs.  A work-group walks the
// tiles blockIdx.x, blockIdx.x + gridDim.x, ...; while it transforms and
// stores tile i, the loads of tile i+1 are in flight into registers.  Tile
// and channel descriptors come through restrict pointers, so they are scalar
// loads and never make the wave wait on the prefetch (barriers here wait on
// LDS only).  Tiles hold rows * N <= FFT_PERSIST_ELEMS complex values.  The
// quadrant twiddle table is (re)loaded into LDS when the channel changes.
// ---------------------------------------------------------------------------
static_assert(FFT_PF * FFT_WG == FFT_PERSIST_ELEMS, "persistent FFT tile size");
constexpr int FFT_PF_PAIRS = (FFT_PERSIST_ELEMS / 2 + FFT_MAX_ROWS) / FFT_WG;  // spectrum pairs per lane

__device__ __forceinline__ void fwd_prefetch(const StepBuffers& b, const int4* __restrict__ tiles,
                                             const ChanDesc* __restrict__ chans, int t, double2 (&pf)[FFT_PF]) {
  const int4 tt = tiles[t];
  const ChanDesc* __restrict__ cd = chans + tt.x;
  const int nx = cd->nx, total = tt.z * (nx >> 1);
  const double2* __restrict__ src = reinterpret_cast<const double2*>(b.rho_e + cd->fld_off + (long long)tt.y * nx);
#pragma unroll
  for (int c = 0; c < FFT_PF; ++c) pf[c] = src[min((int)threadIdx.x + c * FFT_WG, total - 1)];
}

__global__ __launch_bounds__(FFT_WG, 2) void k_fft_fwd_p(StepBuffers b, const int4* __restrict__ tiles,
                                                      const ChanDesc* __restrict__ chans, int n_tiles) {
  extern __shared__ double2 lds[];
  double2* twl = lds + FFT_PERSIST_ELEMS;  // quadrant twiddle table after the rows
  int tw_chan = -1;
  double2 pf[FFT_PF];
  if ((int)blockIdx.x < n_tiles) fwd_prefetch(b, tiles, chans, blockIdx.x, pf);
  for (int t = blockIdx.x; t < n_tiles; t += gridDim.x) {
    const int4 tt = tiles[t];
    const ChanDesc* __restrict__ cd = chans + tt.x;
    const int nx = cd->nx, N = nx >> 1, lgN = ilog2(N), lgnx = lgN + 1, quarter = nx >> 2;
    const int y0 = tt.y, rows = tt.z, total = rows << lgN;
    const bool active = cd->active != 0;
    if (active) {
      if (tt.x != tw_chan) {
        // through registers: an LDS-DMA copy here would make the compiler
        // wait for the prefetch before every later LDS read
        const double2* __restrict__ tw = b.tw + cd->tw_off;
        for (int j = threadIdx.x; j < quarter; j += FFT_WG) twl[j] = tw[j];
        tw_chan = tt.x;
      }
      const double pc = cd->pc_dy2;
      // wall lifts of rows 1 and ny-2 (odd_extension, poisson.cu:229-241);
      // ny >= 4, so no element is in both
      const bool wlo = (y0 == 1), whi = (y0 + rows - 1 == cd->ny - 2);
      const double* __restrict__ Wl = b.walls + cd->wall_off;
      const long long ws = b.walls_stride;
#pragma unroll
      for (int c = 0; c < FFT_PF; ++c) {
        const int e = threadIdx.x + c * FFT_WG;
        if (e < total) {
          double v0 = -pc * pf[c].x, v1 = -pc * pf[c].y;
          if (wlo || whi) {
            const int r = e >> lgN, n = e & (N - 1);
            const bool lo = wlo && r == 0, hi = whi && r == rows - 1;
            if (lo || hi) {
              const double* w = Wl + (lo ? W_PSI_LO : W_PSI_HI) * ws + 2 * n;
              v0 -= w[0];
              v1 -= w[1];
            }
          }
          lds[sw(e)] = make_double2(v0, v1);
        }
      }
    }
    // the next tile's loads overlap this tile's transform and stores
    if (t + (int)gridDim.x < n_tiles) fwd_prefetch(b, tiles, chans, t + gridDim.x, pf);
    if (active) {
      __syncthreads();
      stockham<-1, FFT_WG, 16>(lds, rows, N, lgN, twl, lgnx);
      double2* __restrict__ spec = b.spec + cd->spec_off;
      const float inv_rows = 1.0f / (float)rows;
      for (int e = threadIdx.x; e < rows * (N + 1); e += FFT_WG) {
        const int k = div_rows(e, inv_rows), r = e - k * rows;
        const double2 zk = lds[sw((r << lgN) + (k & (N - 1)))];
        const double2 zc = cconj(lds[sw((r << lgN) + ((N - k) & (N - 1)))]);
        const double2 E = make_double2(0.5 * (zk.x + zc.x), 0.5 * (zk.y + zc.y));
        const double2 dd = csub(zk, zc);
        const double2 O = make_double2(0.5 * dd.y, -0.5 * dd.x);  // (zk - zc)/(2i)
        spec[spec_idx(cd->ny, y0 + r, k)] = cadd(E, cmul(tw_half(twl, k, quarter), O));
      }
      __syncthreads();  // LDS rows and twiddles are rewritten next
    }
  }
}

// spectrum pairs (X[p], X[N-p]) of tile t, p = e / rows, r = e - p rows
__device__ __forceinline__ void inv_prefetch(const StepBuffers& b, const int4* __restrict__ tiles,
                                             const ChanDesc* __restrict__ chans, int t,
                                             double2 (&xa)[FFT_PF_PAIRS], double2 (&xb)[FFT_PF_PAIRS]) {
  const int4 tt = tiles[t];
  const ChanDesc* __restrict__ cd = chans + tt.x;
  const int N = cd->nx >> 1, rows = tt.z, ny = cd->ny;
  const int npair = rows * ((N >> 1) + 1);
  const float inv_rows = 1.0f / (float)rows;
  const double2* __restrict__ spec = b.spec + cd->spec_off;
#pragma unroll
  for (int c = 0; c < FFT_PF_PAIRS; ++c) {
    const int e = min((int)threadIdx.x + c * FFT_WG, npair - 1);
    const int p = div_rows(e, inv_rows), r = e - p * rows;
    xa[c] = spec[spec_idx(ny, tt.y + r, p)];
    xb[c] = spec[spec_idx(ny, tt.y + r, N - p)];
  }
}

__global__ __launch_bounds__(FFT_WG, 2) void k_fft_inv_p(StepBuffers b, const int4* __restrict__ tiles,
                                                      const ChanDesc* __restrict__ chans, int n_tiles) {
  extern __shared__ double2 lds[];
  double2* twl = lds + FFT_PERSIST_ELEMS;  // quadrant twiddle table after the rows
  int tw_chan = -1;
  double2 xa[FFT_PF_PAIRS], xb[FFT_PF_PAIRS];
  if ((int)blockIdx.x < n_tiles) inv_prefetch(b, tiles, chans, blockIdx.x, xa, xb);
  for (int t = blockIdx.x; t < n_tiles; t += gridDim.x) {
    const int4 tt = tiles[t];
    const ChanDesc* __restrict__ cd = chans + tt.x;
    const int nx = cd->nx, N = nx >> 1, lgN = ilog2(N), lgnx = lgN + 1, quarter = nx >> 2, half = N >> 1;
    const int y0 = tt.y, rows = tt.z, ny = cd->ny;
    const int npair = rows * (half + 1);
    const bool active = cd->active != 0;
    if (active) {
      if (tt.x != tw_chan) {
        // through registers (see k_fft_fwd_p)
        const double2* __restrict__ tw = b.tw + cd->tw_off;
        for (int j = threadIdx.x; j < quarter; j += FFT_WG) twl[j] = tw[j];
        tw_chan = tt.x;
        __syncthreads();
      }
      // z[k] = E + i O, E = (X[k] + conj X[N-k]) / 2,
      // O = (X[k] - conj X[N-k]) / 2 * conj W^k, for k = p and k = N - p
      const float inv_rows = 1.0f / (float)rows;
#pragma unroll
      for (int c = 0; c < FFT_PF_PAIRS; ++c) {
        const int e = threadIdx.x + c * FFT_WG;
        if (e < npair) {
          const int p = div_rows(e, inv_rows), r = e - p * rows;
          const int rb = r << lgN;
          {
            const double2 xk = xa[c], xc = cconj(xb[c]);
            const double2 E = make_double2(0.5 * (xk.x + xc.x), 0.5 * (xk.y + xc.y));
            const double2 O = cmul(make_double2(0.5 * (xk.x - xc.x), 0.5 * (xk.y - xc.y)),
                                   cconj(tw_half(twl, p, quarter)));
            lds[sw(rb + p)] = make_double2(E.x - O.y, E.y + O.x);
          }
          if (p > 0 && p < half) {
            const double2 xk = xb[c], xc = cconj(xa[c]);
            const double2 E = make_double2(0.5 * (xk.x + xc.x), 0.5 * (xk.y + xc.y));
            const double2 O = cmul(make_double2(0.5 * (xk.x - xc.x), 0.5 * (xk.y - xc.y)),
                                   cconj(tw_half(twl, N - p, quarter)));
            lds[sw(rb + N - p)] = make_double2(E.x - O.y, E.y + O.x);
          }
        }
      }
    }
    if (t + (int)gridDim.x < n_tiles) inv_prefetch(b, tiles, chans, t + gridDim.x, xa, xb);
    if (active) {
      __syncthreads();
      stockham<+1, FFT_WG, 16>(lds, rows, N, lgN, twl, lgnx);
      double* __restrict__ ps = b.psi + cd->fld_off;
      const double invN = 1.0 / (double)N;
      double2* __restrict__ dst = reinterpret_cast<double2*>(ps + (long long)y0 * nx);
      for (int e = threadIdx.x; e < rows * N; e += FFT_WG) {
        const double2 z = lds[sw(e)];
        dst[e] = make_double2(z.x * invN, z.y * invN);
      }
      // Dirichlet rows (odd_extract, poisson.cu:297-303)
      const double* __restrict__ Wl = b.walls + cd->wall_off;
      const long long ws = b.walls_stride;
      if (y0 == 1) copy_row_batched<FFT_WG>(ps, Wl + W_PSI_LO * ws, nx);
      if (y0 + rows - 1 == ny - 2) copy_row_batched<FFT_WG>(ps + (long long)(ny - 1) * nx, Wl + W_PSI_HI * ws, nx);
      __syncthreads();  // LDS rows and twiddles are rewritten next
    }
  }
}

// ---------------------------------------------------------------------------
// k_fft_inv: spectrum rows -> psi rows (normalised by 1/nx)
// ---------------------------------------------------------------------------
// Loads as in k_fft_fwd: all of a tile's spectrum values (as pairs X[p],
// X[N-p]) and twiddles are in flight before the first is consumed.
template <bool POW2, bool TWL>
__global__ __launch_bounds__(FFT_WG_ROWS) void k_fft_inv(StepBuffers b, const int4* __restrict__ tiles,
                                                   const ChanDesc* __restrict__ chans, int tile_base) {
  extern __shared__ double2 lds[];
  const int4 t = tiles[tile_base + blockIdx.x];
  const ChanDesc* __restrict__ cd = chans + t.x;
  if (!cd->active) return;
  const int nx = cd->nx, nk = cd->nk, ny = cd->ny, y0 = t.y, rows = t.z;
  const double2* __restrict__ tw = b.tw + cd->tw_off;
  const double2* __restrict__ spec = b.spec + cd->spec_off;
  double* __restrict__ ps = b.psi + cd->fld_off;
  if (POW2) {
    const int N = nx >> 1, lgN = ilog2(N), lgnx = lgN + 1, quarter = nx >> 2, half = N >> 1;
    const int npair = rows * (half + 1);  // pairs (p, N - p), p = 0..N/2, of each row
    double2* twl = lds + (((rows << lgN) + 15) & ~15);
    if (TWL) lds_dma_copy<FFT_WG_ROWS>(twl, tw, quarter);
    // pair e: p = e / rows, r = e - p rows (rows vary fastest: the 8 modes of
    // a spectrum block are one 128-byte line per row)
    const float inv_rows = 1.0f / (float)rows;
    double2 xa[FFT_PAIR_SLOTS], xb[FFT_PAIR_SLOTS];
#pragma unroll
    for (int c = 0; c < FFT_PAIR_SLOTS; ++c)
      if (c * FFT_WG_ROWS < npair) {
        const int e = min((int)threadIdx.x + c * FFT_WG_ROWS, npair - 1);
        const int p = div_rows(e, inv_rows), r = e - p * rows;
        xa[c] = spec[spec_idx(ny, y0 + r, p)];
        xb[c] = spec[spec_idx(ny, y0 + r, N - p)];
      }
    if (!TWL) tw2_fill<FFT_WG_ROWS>(twl, tw, nx, lgnx);  // the spectrum's loads are in flight
    __syncthreads();  // twiddles in LDS
    const double2* twq = twl;
    // z[k] = E + i O, E = (X[k] + conj X[N-k]) / 2,
    // O = (X[k] - conj X[N-k]) / 2 * conj W^k, for k = p and k = N - p
#pragma unroll
    for (int c = 0; c < FFT_PAIR_SLOTS; ++c) {
      const int e = threadIdx.x + c * FFT_WG_ROWS;
      if (c * FFT_WG_ROWS < npair && e < npair) {
        const int p = div_rows(e, inv_rows), r = e - p * rows;
        const int rb = r << lgN;
        {
          const double2 xk = xa[c], xc = cconj(xb[c]);
          const double2 E = make_double2(0.5 * (xk.x + xc.x), 0.5 * (xk.y + xc.y));
          const double2 O = cmul(make_double2(0.5 * (xk.x - xc.x), 0.5 * (xk.y - xc.y)),
                                 cconj(tw_half_get<!TWL>(twq, p, quarter, lgnx)));
          lds[sw(rb + p)] = make_double2(E.x - O.y, E.y + O.x);
        }
        if (p > 0 && p < half) {
          const double2 xk = xb[c], xc = cconj(xa[c]);
          const double2 E = make_double2(0.5 * (xk.x + xc.x), 0.5 * (xk.y + xc.y));
          const double2 O = cmul(make_double2(0.5 * (xk.x - xc.x), 0.5 * (xk.y - xc.y)),
                                 cconj(tw_half_get<!TWL>(twq, N - p, quarter, lgnx)));
          lds[sw(rb + N - p)] = make_double2(E.x - O.y, E.y + O.x);
        }
      }
    }
    __syncthreads();
    stockham<+1, FFT_WG_ROWS, 8, !TWL>(lds, rows, N, lgN, twq, lgnx);
    const double invN = 1.0 / (double)N;
    double2* __restrict__ dst = reinterpret_cast<double2*>(ps + (long long)y0 * nx);
    for (int e = threadIdx.x; e < rows * N; e += FFT_WG_ROWS) {
      const double2 z = lds[sw(e)];
      dst[e] = make_double2(z.x * invN, z.y * invN);
    }
    // Dirichlet rows psi(0) = voltage, psi(ny-1) = voltage2 (odd_extract,
    // poisson.cu:297-303), written by the tiles holding rows 1 and ny-2
    const double* __restrict__ Wl = b.walls + cd->wall_off;
    const long long ws = b.walls_stride;
    if (y0 == 1) copy_row_batched<FFT_WG_ROWS>(ps, Wl + W_PSI_LO * ws, nx);
    if (y0 + rows - 1 == ny - 2) copy_row_batched<FFT_WG_ROWS>(ps + (long long)(ny - 1) * nx, Wl + W_PSI_HI * ws, nx);
  } else {
    const double* __restrict__ Wl = b.walls + cd->wall_off;
    const long long ws = b.walls_stride;
    if (y0 == 1) copy_row_batched<FFT_WG_ROWS>(ps, Wl + W_PSI_LO * ws, nx);
    if (y0 + rows - 1 == ny - 2) copy_row_batched<FFT_WG_ROWS>(ps + (long long)(ny - 1) * nx, Wl + W_PSI_HI * ws, nx);
    // the tile's spectra in LDS, [row][mode] (loads in the order of the
    // spectrum layout), then one output node per lane, x fastest
    const int nblk = (nk + 7) >> 3;
    for (int e = threadIdx.x; e < nblk * rows * 8; e += FFT_WG_ROWS) {
      const int kb = e / (rows * 8), rem = e - kb * rows * 8, r = rem >> 3, k = kb * 8 + (rem & 7);
      if (k < nk) lds[r * nk + k] = spec[spec_idx(ny, y0 + r, k)];
    }
    __syncthreads();
    const int kmax = (nx - 1) / 2;  // modes counted twice
    const bool nyq = (nx & 1) == 0;
    for (int e = threadIdx.x; e < rows * nx; e += FFT_WG_ROWS) {
      const int r = e / nx, n = e - r * nx;
      const double2* __restrict__ z = lds + r * nk;
      double acc = z[0].x;
      int idx = 0;
      for (int k = 1; k <= kmax; ++k) {
        idx += n;
        if (idx >= nx) idx -= nx;
        const double2 w = tw[idx];
        acc += 2.0 * (z[k].x * w.x + z[k].y * w.y);
      }
      if (nyq) acc += (n & 1) ? -z[nx / 2].x : z[nx / 2].x;
      ps[(long long)(y0 + r) * nx + n] = acc / (double)nx;
    }
  }
}

// ---------------------------------------------------------------------------
// Row-pair x-FFTs.  A work-group takes the rows A = y0 and B = y0 + 1 of a
// tile (a channel's last tile may hold A only) and transforms them one after
// the other.  Every load of both rows is issued before the first is used,
// row B's after row A's, so row B's data arrives while row A is transformed,
// and row A's stores drain while row B is transformed.  Loads and stores are
// straight-line and unconditional (indices past the end are clamped: a
// duplicate store writes the same value to the same place), so the wait the
// compiler puts before each use counts exactly the memory operations issued
// after the one needed; in particular row B's waits do not include row A's
// stores.  S = N / FFT_WG_ROWS values per lane and row; twiddles pass
// through registers into LDS with row A.
// ---------------------------------------------------------------------------
constexpr int clog2(int v) { return v <= 1 ? 0 : 1 + clog2(v >> 1); }

template <int S>
__global__ __launch_bounds__(FFT_WG_ROWS) void k_fft_fwd2(StepBuffers b, const int4* __restrict__ tiles,
                                                          const ChanDesc* __restrict__ chans) {
  extern __shared__ double2 lds[];
  constexpr int WG = FFT_WG_ROWS, N = S * WG, lgN = clog2(N), lgnx = lgN + 1, quarter = N / 2;
  constexpr int TQ = quarter / WG;  // twiddle values per lane
  static_assert(S >= 2 && quarter % WG == 0, "row-pair FFT sizes");
  const int4 t = tiles[blockIdx.x];
  const ChanDesc* __restrict__ cd = chans + t.x;
  if (!cd->active) return;
  const int nx = 2 * N, ny = cd->ny, y0 = t.y;
  const bool two = t.z == 2;
  const double* __restrict__ re = b.rho_e + cd->fld_off;
  const double* __restrict__ Wl = b.walls + cd->wall_off;
  const long long ws = b.walls_stride;
  const double2* __restrict__ tw = b.tw + cd->tw_off;
  double2* __restrict__ spec = b.spec + cd->spec_off;
  const double pc = cd->pc_dy2;
  const int tid = threadIdx.x;
  double2* rowA = lds;
  double2* rowB = lds + N;
  double2* twl = lds + 2 * N;
  // wall lifts (odd_extension, poisson.cu:229-241): row 1 takes psi_lo, row
  // ny-2 psi_hi; loaded first, only by the tiles that hold those rows, and
  // read only under the same conditions (left unset otherwise, so that no
  // merge of the loaded values makes the other tiles wait for them)
  const bool wlo = y0 == 1, whi = y0 + t.z - 1 == ny - 2;
  double2 wl[S], wh[S];
  if (wlo) {
    const double2* w = reinterpret_cast<const double2*>(Wl + W_PSI_LO * ws);
#pragma unroll
    for (int c = 0; c < S; ++c) wl[c] = w[tid + c * WG];
  }
  if (whi) {
    const double2* w = reinterpret_cast<const double2*>(Wl + W_PSI_HI * ws);
#pragma unroll
    for (int c = 0; c < S; ++c) wh[c] = w[tid + c * WG];
  }
  double2 tq[TQ], va[S], vb[S];
#pragma unroll
  for (int c = 0; c < TQ; ++c) tq[c] = tw[tid + c * WG];
  const double2* __restrict__ srcA = reinterpret_cast<const double2*>(re + (long long)y0 * nx);
  const double2* __restrict__ srcB = reinterpret_cast<const double2*>(re + (long long)(y0 + (two ? 1 : 0)) * nx);
#pragma unroll
  for (int c = 0; c < S; ++c) va[c] = srcA[tid + c * WG];
#pragma unroll
  for (int c = 0; c < S; ++c) vb[c] = srcB[tid + c * WG];

#pragma unroll
  for (int c = 0; c < TQ; ++c) twl[tid + c * WG] = tq[c];
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    if (r == 1 && !two) break;
    double2* x = r == 0 ? rowA : rowB;
    const bool lift = r == 0 ? wlo : (whi && two);
    const bool lift_hi = r == 0 && whi && !two;
#pragma unroll
    for (int c = 0; c < S; ++c) {
      const double2 v = r == 0 ? va[c] : vb[c];
      double v0 = -pc * v.x, v1 = -pc * v.y;
      if (r == 0 && lift) {
        v0 -= wl[c].x;
        v1 -= wl[c].y;
      }
      if ((r == 1 && lift) || lift_hi) {
        v0 -= wh[c].x;
        v1 -= wh[c].y;
      }
      x[sw(tid + c * WG)] = make_double2(v0, v1);
    }
    __syncthreads();
    stockham<-1, WG, 8>(x, 1, N, lgN, twl, lgnx);
    // split X[k] = E[k] + W^k O[k], k = 0..N (lanes past N repeat k = N)
    const int y = y0 + r;
#pragma unroll
    for (int c = 0; c <= S; ++c) {
      const int k = min(tid + c * WG, N);
      const double2 zk = x[sw(k & (N - 1))];
      const double2 zc = cconj(x[sw((N - k) & (N - 1))]);
      const double2 E = make_double2(0.5 * (zk.x + zc.x), 0.5 * (zk.y + zc.y));
      const double2 dd = csub(zk, zc);
      const double2 O = make_double2(0.5 * dd.y, -0.5 * dd.x);
      spec[spec_idx(ny, y, k)] = cadd(E, cmul(tw_half(twl, k, quarter), O));
    }
  }
}

template <int S>
__global__ __launch_bounds__(FFT_WG_ROWS) void k_fft_inv2(StepBuffers b, const int4* __restrict__ tiles,
                                                          const ChanDesc* __restrict__ chans) {
  extern __shared__ double2 lds[];
  constexpr int WG = FFT_WG_ROWS, N = S * WG, lgN = clog2(N), lgnx = lgN + 1, quarter = N / 2, half = N / 2;
  constexpr int TQ = quarter / WG, P = half / WG + 1;  // twiddles, pair slots per lane
  static_assert(S >= 2 && quarter % WG == 0, "row-pair FFT sizes");
  const int4 t = tiles[blockIdx.x];
  const ChanDesc* __restrict__ cd = chans + t.x;
  if (!cd->active) return;
  const int nx = 2 * N, ny = cd->ny, y0 = t.y;
  const bool two = t.z == 2;
  const double2* __restrict__ tw = b.tw + cd->tw_off;
  const double2* __restrict__ spec = b.spec + cd->spec_off;
  double* __restrict__ ps = b.psi + cd->fld_off;
  const int tid = threadIdx.x;
  double2* rowA = lds;
  double2* rowB = lds + N;
  double2* twl = lds + 2 * N;
  double2 tq[TQ], xa[2][P], xb[2][P];
#pragma unroll
  for (int c = 0; c < TQ; ++c) tq[c] = tw[tid + c * WG];
  // pairs X[p], X[N - p], p = 0..N/2 (lanes past N/2 repeat p = N/2)
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const int y = y0 + (r == 1 && two ? 1 : 0);
#pragma unroll
    for (int c = 0; c < P; ++c) {
      const int p = min(tid + c * WG, half);
      xa[r][c] = spec[spec_idx(ny, y, p)];
      xb[r][c] = spec[spec_idx(ny, y, N - p)];
    }
  }
#pragma unroll
  for (int c = 0; c < TQ; ++c) twl[tid + c * WG] = tq[c];
  __syncthreads();  // twiddles in LDS before the merge reads them
  const double invN = 1.0 / (double)N;
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    if (r == 1 && !two) break;
    double2* x = r == 0 ? rowA : rowB;
    // z[k] = E + i O, E = (X[k] + conj X[N-k]) / 2,
    // O = (X[k] - conj X[N-k]) / 2 * conj W^k, for k = p and k = N - p
#pragma unroll
    for (int c = 0; c < P; ++c) {
      const int p = min(tid + c * WG, half);
      {
        const double2 xk = xa[r][c], xc = cconj(xb[r][c]);
        const double2 E = make_double2(0.5 * (xk.x + xc.x), 0.5 * (xk.y + xc.y));
        const double2 O = cmul(make_double2(0.5 * (xk.x - xc.x), 0.5 * (xk.y - xc.y)),
                               cconj(tw_half(twl, p, quarter)));
        x[sw(p)] = make_double2(E.x - O.y, E.y + O.x);
      }
      if (p > 0 && p < half) {
        const double2 xk = xb[r][c], xc = cconj(xa[r][c]);
        const double2 E = make_double2(0.5 * (xk.x + xc.x), 0.5 * (xk.y + xc.y));
        const double2 O = cmul(make_double2(0.5 * (xk.x - xc.x), 0.5 * (xk.y - xc.y)),
                               cconj(tw_half(twl, N - p, quarter)));
        x[sw(N - p)] = make_double2(E.x - O.y, E.y + O.x);
      }
    }
    __syncthreads();
    stockham<+1, WG, 8>(x, 1, N, lgN, twl, lgnx);
    double2* __restrict__ dst = reinterpret_cast<double2*>(ps + (long long)(y0 + r) * nx);
#pragma unroll
    for (int c = 0; c < S; ++c) {
      const double2 z = x[sw(tid + c * WG)];
      dst[tid + c * WG] = make_double2(z.x * invN, z.y * invN);
    }
  }
  // Dirichlet rows psi(0) = voltage, psi(ny-1) = voltage2 (odd_extract,
  // poisson.cu:297-303), written by the tiles holding rows 1 and ny-2
  const double* __restrict__ Wl = b.walls + cd->wall_off;
  const long long ws = b.walls_stride;
  if (y0 == 1) copy_row_batched<WG>(ps, Wl + W_PSI_LO * ws, nx);
  if (y0 + t.z - 1 == ny - 2) copy_row_batched<WG>(ps + (long long)(ny - 1) * nx, Wl + W_PSI_HI * ws, nx);
}

// ---------------------------------------------------------------------------
// Row-pair x-FFTs for rows too wide for two row buffers per work-group
// (nx = 8192: two 64 KB rows would leave one work-group per CU).  One row
// buffer and the two-level twiddle table in LDS (67 KB, two work-groups per
// CU); row B waits in registers while row A is transformed and moves into the
// buffer once every lane has read row A's result out of it.  Load order: the
// twiddle table's quadrant-table values, row A, row B, so that filling the
// table waits for its own loads only and row A's first use for row A's.  The
// wall lifts of rows 1 and ny-2 are read where they are applied, by the two
// tiles holding those rows.  The operations of the one-tile kernels with the
// two-level table (results equal to FP64 rounding: the compiler contracts
// multiply-adds per kernel).  Four waves per SIMD (two work-groups per CU,
// the LDS limit) cap the registers at 128.
// ---------------------------------------------------------------------------
template <int S>
__global__ __launch_bounds__(FFT_WG_ROWS) __attribute__((amdgpu_waves_per_eu(4))) void k_fft_fwd2w(
    StepBuffers b, const int4* __restrict__ tiles, const ChanDesc* __restrict__ chans) {
  extern __shared__ double2 lds[];
  constexpr int WG = FFT_WG_ROWS, N = S * WG, lgN = clog2(N), lgnx = lgN + 1, nx = 2 * N, quarter = N / 2;
  constexpr int NT2 = tw2_count(nx);
  static_assert(NT2 <= WG && N % 16 == 0, "wide row-pair FFT sizes");
  const int4 t = tiles[blockIdx.x];
  const ChanDesc* __restrict__ cd = chans + t.x;
  if (!cd->active) return;
  const int ny = cd->ny, y0 = t.y;
  const bool two = t.z == 2;
  const double* __restrict__ re = b.rho_e + cd->fld_off;
  const double2* __restrict__ tw = b.tw + cd->tw_off;
  double2* __restrict__ spec = b.spec + cd->spec_off;
  const double pc = cd->pc_dy2;
  const int tid = threadIdx.x;
  double2* x = lds;
  double2* t2 = lds + N;
  const int i2 = min(tid, NT2 - 1);
  const int tt = (i2 < TW2_FINE ? i2 : (i2 - TW2_FINE) << TW2_LGF) & (nx - 1);
  const double2 wq = tw[tw_quad_index(tt, lgnx)];
  const double2* __restrict__ srcA = reinterpret_cast<const double2*>(re + (long long)y0 * nx);
  const double2* __restrict__ srcB = reinterpret_cast<const double2*>(re + (long long)(y0 + (two ? 1 : 0)) * nx);
  double2 va[S], vb[S];
  // the three groups of loads stay in this order (no instruction is moved
  // across a scheduling barrier), so each wait below counts only the loads
  // issued after the ones it needs
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int c = 0; c < S; ++c) va[c] = srcA[tid + c * WG];
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int c = 0; c < S; ++c) vb[c] = srcB[tid + c * WG];
  __builtin_amdgcn_sched_barrier(0);
  // every lane writes (lanes past the table rewrite its last value), so the
  // table's load is not moved into a branch behind the rows' loads
  t2[i2] = tw_quad_resolve<-1>(wq, tt, lgnx);
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    if (r == 1 && !two) break;
    if (r == 1) __syncthreads();  // row A's split has read the row buffer
    const int y = y0 + r;
#pragma unroll
    for (int c = 0; c < S; ++c) {
      const double2 v = r == 0 ? va[c] : vb[c];
      x[sw(tid + c * WG)] = make_double2(-pc * v.x, -pc * v.y);
    }
    // wall lifts (odd_extension, poisson.cu:229-241): row 1 takes psi_lo,
    // row ny-2 psi_hi (distinct rows: ny >= 4); each lane updates its own
    // elements, so the barrier below orders them for the transform
    if (y == 1 || y == ny - 2) {
      const double2* __restrict__ w = reinterpret_cast<const double2*>(
          b.walls + cd->wall_off + (y == 1 ? W_PSI_LO : W_PSI_HI) * b.walls_stride);
#pragma unroll
      for (int c = 0; c < S; ++c) {
        const int e = tid + c * WG;
        const double2 l = w[e], v = x[sw(e)];
        x[sw(e)] = make_double2(v.x - l.x, v.y - l.y);
      }
    }
    __syncthreads();
    stockham<-1, WG, 8, true>(x, 1, N, lgN, t2, lgnx);
    // split X[k] = E[k] + W^k O[k], k = 0..N (lanes past N repeat k = N)
#pragma unroll
    for (int c = 0; c <= S; ++c) {
      const int k = min(tid + c * WG, N);
      const double2 zk = x[sw(k & (N - 1))];
      const double2 zc = cconj(x[sw((N - k) & (N - 1))]);
      const double2 E = make_double2(0.5 * (zk.x + zc.x), 0.5 * (zk.y + zc.y));
      const double2 dd = csub(zk, zc);
      const double2 O = make_double2(0.5 * dd.y, -0.5 * dd.x);
      spec[spec_idx(ny, y, k)] = cadd(E, cmul(tw_half_get<true>(t2, k, quarter, lgnx), O));
    }
  }
}

template <int S>
__global__ __launch_bounds__(FFT_WG_ROWS) __attribute__((amdgpu_waves_per_eu(4))) void k_fft_inv2w(
    StepBuffers b, const int4* __restrict__ tiles, const ChanDesc* __restrict__ chans) {
  extern __shared__ double2 lds[];
  constexpr int WG = FFT_WG_ROWS, N = S * WG, lgN = clog2(N), lgnx = lgN + 1, nx = 2 * N, quarter = N / 2;
  constexpr int half = N / 2, P = half / WG + 1, NT2 = tw2_count(nx);  // pair slots per lane
  static_assert(NT2 <= WG && N % 16 == 0, "wide row-pair FFT sizes");
  const int4 t = tiles[blockIdx.x];
  const ChanDesc* __restrict__ cd = chans + t.x;
  if (!cd->active) return;
  const int ny = cd->ny, y0 = t.y;
  const bool two = t.z == 2;
  const double2* __restrict__ tw = b.tw + cd->tw_off;
  const double2* __restrict__ spec = b.spec + cd->spec_off;
  double* __restrict__ ps = b.psi + cd->fld_off;
  const int tid = threadIdx.x;
  double2* x = lds;
  double2* t2 = lds + N;
  const int i2 = min(tid, NT2 - 1);
  const int tt = (i2 < TW2_FINE ? i2 : (i2 - TW2_FINE) << TW2_LGF) & (nx - 1);
  const double2 wq = tw[tw_quad_index(tt, lgnx)];
  // pairs X[p], X[N - p], p = 0..N/2 (lanes past N/2 repeat p = N/2); the
  // table's load, row A's and row B's stay in this order (scheduling barriers)
  double2 xa[2][P], xb[2][P];
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    __builtin_amdgcn_sched_barrier(0);
    const int y = y0 + (r == 1 && two ? 1 : 0);
#pragma unroll
    for (int c = 0; c < P; ++c) {
      const int p = min(tid + c * WG, half);
      xa[r][c] = spec[spec_idx(ny, y, p)];
      xb[r][c] = spec[spec_idx(ny, y, N - p)];
    }
  }
  __builtin_amdgcn_sched_barrier(0);
  t2[i2] = tw_quad_resolve<-1>(wq, tt, lgnx);  // every lane: no branch around the load
  __syncthreads();  // twiddles in LDS before the merge reads them
  const double invN = 1.0 / (double)N;
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    if (r == 1 && !two) break;
    if (r == 1) __syncthreads();  // row A's output has been read from the row buffer
    // z[k] = E + i O, E = (X[k] + conj X[N-k]) / 2,
    // O = (X[k] - conj X[N-k]) / 2 * conj W^k, for k = p and k = N - p
#pragma unroll
    for (int c = 0; c < P; ++c) {
      const int p = min(tid + c * WG, half);
      {
        const double2 xk = xa[r][c], xc = cconj(xb[r][c]);
        const double2 E = make_double2(0.5 * (xk.x + xc.x), 0.5 * (xk.y + xc.y));
        const double2 O = cmul(make_double2(0.5 * (xk.x - xc.x), 0.5 * (xk.y - xc.y)),
                               cconj(tw_half_get<true>(t2, p, quarter, lgnx)));
        x[sw(p)] = make_double2(E.x - O.y, E.y + O.x);
      }
      if (p > 0 && p < half) {
        const double2 xk = xb[r][c], xc = cconj(xa[r][c]);
        const double2 E = make_double2(0.5 * (xk.x + xc.x), 0.5 * (xk.y + xc.y));
        const double2 O = cmul(make_double2(0.5 * (xk.x - xc.x), 0.5 * (xk.y - xc.y)),
                               cconj(tw_half_get<true>(t2, N - p, quarter, lgnx)));
        x[sw(N - p)] = make_double2(E.x - O.y, E.y + O.x);
      }
    }
    __syncthreads();
    stockham<+1, WG, 8, true>(x, 1, N, lgN, t2, lgnx);
    double2* __restrict__ dst = reinterpret_cast<double2*>(ps + (long long)(y0 + r) * nx);
#pragma unroll
    for (int c = 0; c < S; ++c) {
      const double2 z = x[sw(tid + c * WG)];
      dst[tid + c * WG] = make_double2(z.x * invN, z.y * invN);
    }
  }
  // Dirichlet rows psi(0) = voltage, psi(ny-1) = voltage2 (odd_extract,
  // poisson.cu:297-303), written by the tiles holding rows 1 and ny-2
  const double* __restrict__ Wl = b.walls + cd->wall_off;
  const long long ws = b.walls_stride;
  if (y0 == 1) copy_row_batched<WG>(ps, Wl + W_PSI_LO * ws, nx);
  if (y0 + t.z - 1 == ny - 2) copy_row_batched<WG>(ps + (long long)(ny - 1) * nx, Wl + W_PSI_HI * ws, nx);
}

// ---------------------------------------------------------------------------
// k_tridiag: for each wavenumber k the system
//   x_{i-1} + a x_i + x_{i+1} = r_i,  i = 1..M,  x_0 = x_{M+1} = 0,
//   a = -(2 + (kx dy)^2) = -2 cosh(theta)
// (the FD-y / spectral-x operator of poisson.cu:274-283, times dy^2).
// Thomas: c_i = 1/(a - c_{i-1}), y_i = c_i (r_i - y_{i-1}),
//         x_i = y_i - c_i x_{i+1}.
// Both sweeps are first-order affine recurrences, so the column is cut into
// C chunks of L rows (one lane each); every lane sweeps its chunk locally,
// an LDS scan of the chunk maps (A, B) -> (A A', A B' + B) supplies the
// missing carry, and a second pass adds carry * running product.  The
// Thomas coefficient entering a chunk has the closed form
//   c_i = -sinh(i theta)/sinh((i+1) theta).
// Lanes: mode m = tid % G, chunk j = tid / G.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double cprime_closed(int i, double kd) {
  if (kd == 0.0) return -(double)i / (double)(i + 1);
  const double th = 2.0 * asinh(0.5 * kd);
  return -exp(-th) * expm1(-2.0 * (double)i * th) / expm1(-2.0 * (double)(i + 1) * th);
}

// Exclusive scan of affine maps v -> A v + B along the chunks of one mode.
// FWD: chunk j receives the composition of chunks 0..j-1 applied to 0 (the
// forward carry); otherwise chunks j+1..C-1 (the backward carry).  Lanes of
// one mode sit G apart; a wave holds cpw chunks; sA/sB hold one total per
// (wave, mode).
template <bool FWD>
__device__ __forceinline__ double2 chunk_scan(double A, double2 B, int G, int wv, int jw, int cpw, int m,
                                              double* sA, double2* sB) {
  // inclusive scan inside the wave
  for (int d = 1; d < cpw; d <<= 1) {
    const int off = d * G;
    const double An = FWD ? __shfl_up(A, off, 64) : __shfl_down(A, off, 64);
    const double Bx = FWD ? __shfl_up(B.x, off, 64) : __shfl_down(B.x, off, 64);
    const double By = FWD ? __shfl_up(B.y, off, 64) : __shfl_down(B.y, off, 64);
    const bool has = FWD ? (jw >= d) : (jw + d < cpw);
    if (has) {
      B = make_double2(A * Bx + B.x, A * By + B.y);
      A = A * An;
    }
  }
  // wave totals: the last chunk of the wave in scan order
  const int nw = TRI_WG / 64;
  if (jw == (FWD ? cpw - 1 : 0)) {
    sA[wv * 64 + m] = A;
    sB[wv * 64 + m] = B;
  }
  // this lane's exclusive value inside the wave
  const int off = G;
  const double Ae = FWD ? __shfl_up(A, off, 64) : __shfl_down(A, off, 64);
  const double Bex = FWD ? __shfl_up(B.x, off, 64) : __shfl_down(B.x, off, 64);
  const double Bey = FWD ? __shfl_up(B.y, off, 64) : __shfl_down(B.y, off, 64);
  const bool first = FWD ? (jw == 0) : (jw == cpw - 1);
  __syncthreads();
  // carry entering this wave: the totals of the waves before it, in order
  double2 Y = make_double2(0.0, 0.0);
  if (FWD) {
    for (int w = 0; w < wv; ++w) {
      const double a = sA[w * 64 + m];
      const double2 bb = sB[w * 64 + m];
      Y = make_double2(a * Y.x + bb.x, a * Y.y + bb.y);
    }
  } else {
    for (int w = nw - 1; w > wv; --w) {
      const double a = sA[w * 64 + m];
      const double2 bb = sB[w * 64 + m];
      Y = make_double2(a * Y.x + bb.x, a * Y.y + bb.y);
    }
  }
  if (first) return Y;
  return make_double2(Ae * Y.x + Bex, Ae * Y.y + Bey);
}

template <int LMAX>
__device__ __forceinline__ void tridiag_tile(const StepBuffers& b, int2 t, double* sA, double2* sB) {
  const ChanDesc* __restrict__ cd = b.chans + t.x;
  if (!cd->active) return;
  const int G = cd->tri_G, L = cd->tri_L;
  const int nk = cd->nk, M = cd->ny - 2;
  const int ks = 8;  // row stride inside a column block (spec_idx)
  const int tid = threadIdx.x;
  const int m = tid & (G - 1), j = tid / G;
  const int k = t.y + m;
  const bool valid = (k < nk);
  const int i0 = 1 + j * L;
  int len = M - j * L;
  len = len < 0 ? 0 : (len > L ? L : len);
  double2* __restrict__ col = b.spec + cd->spec_off + (valid ? spec_idx(cd->ny, 0, k) : 0);
  const double kd = (double)k * cd->kdy_scale;
  const double a = -(2.0 + kd * kd);

  // Straight-line sweeps over LMAX rows; rows at or past `len` are masked
  // with selects (no branches, so the arrays stay in registers).
  double cc[LMAX], yr[LMAX], yi[LMAX];
#pragma unroll
  for (int ii = 0; ii < LMAX; ++ii) {
    const int row = min(i0 + ii, M);  // clamped: always a valid row
    const double2 v = col[(long long)row * ks];
    yr[ii] = v.x;
    yi[ii] = v.y;
  }
  // local forward sweep with y_{i0-1} = 0; the coefficient entering the
  // chunk, cprime_closed(i0 - 1, kd), comes from the plan's table
  double cprev = (valid && i0 > 1 && len > 0) ? b.cprime[cd->cp_off + (long long)j * nk + k] : 0.0;
  double A = 1.0, pr = 0.0, pim = 0.0;
#pragma unroll
  for (int ii = 0; ii < LMAX; ++ii) {
    const bool on = ii < len;
    const double c = 1.0 / (a - cprev);
    cc[ii] = on ? c : 0.0;
    cprev = on ? c : cprev;
    pr = on ? c * (yr[ii] - pr) : pr;
    pim = on ? c * (yi[ii] - pim) : pim;
    yr[ii] = pr;
    yi[ii] = pim;
    A = on ? -c * A : A;
  }
  // Exclusive scan of the chunk maps y -> A y + B over j: shuffles inside a
  // wave (the wave holds 64/G consecutive chunks of its G modes), one LDS
  // exchange of wave totals across the work-group.
  const int wv = tid >> 6, jw = (tid & 63) / G, cpw = 64 / G;
  const double2 Yin = chunk_scan<true>(A, make_double2(pr, pim), G, wv, jw, cpw, m, sA, sB);
  // carry into the forward sweep, then local backward sweep with x_{e+1} = 0
  // (masked rows have cc = 0: the carry product becomes 0 there, and the
  // backward recurrence passes x through unchanged only for rows < len)
  double P = 1.0;
#pragma unroll
  for (int ii = 0; ii < LMAX; ++ii) {
    P *= -cc[ii];
    yr[ii] += P * Yin.x;
    yi[ii] += P * Yin.y;
  }
  double xr = 0.0, xim = 0.0;
#pragma unroll
  for (int ii = LMAX - 1; ii >= 0; --ii) {
    const bool on = ii < len;
    const double nxr = yr[ii] - cc[ii] * xr;
    const double nxi = yi[ii] - cc[ii] * xim;
    xr = on ? nxr : xr;
    xim = on ? nxi : xim;
    yr[ii] = xr;
    yi[ii] = xim;
  }
  __syncthreads();  // the forward exchange's LDS is reused
  const double2 Xin = chunk_scan<false>(A, make_double2(xr, xim), G, wv, jw, cpw, m, sA, sB);
  P = 1.0;
#pragma unroll
  for (int ii = LMAX - 1; ii >= 0; --ii) {
    const bool on = ii < len;
    P = on ? -cc[ii] * P : P;
    yr[ii] += P * Xin.x;
    yi[ii] += P * Xin.y;
  }
  if (valid) {
#pragma unroll
    for (int ii = 0; ii < LMAX; ++ii)
      if (ii < len) col[(long long)(i0 + ii) * ks] = make_double2(yr[ii], yi[ii]);
  }
}

template <int LMAX>
__global__ __launch_bounds__(TRI_WG) void k_tridiag(StepBuffers b) {
  __shared__ double sA[TRI_WG];
  __shared__ double2 sB[TRI_WG];
  // consecutive tiles (the 8-mode lines of one spectrum block) share an XCD's L2
  tridiag_tile<LMAX>(b, b.tri_tiles[xcd_remap(blockIdx.x, gridDim.x)], sA, sB);
}

// The tridiagonal launch of a coupled step with the wall rows of the step's
// collide-stream added: the first nw work-groups (the wall tiles, padded to a
// multiple of 8 so that the tridiagonal tiles keep their XCD placement) update
// the row-kernel wall tiles (as k_lbm_rows<true>), the rest solve the
// tridiagonal tiles.  Wall work-groups go first so that they start with the
// first round of tridiagonal tiles instead of waiting for a late slot.  Valid because the
// wall-row update reads only the step's source populations, psi and the
// storage references, none of which the forward x-FFT or the tridiagonal
// writes, and writes rows 0 and ny-1, which neither of them reads; the wall
// path needs fewer registers than the tridiagonal, so the launch keeps its
// occupancy, and its work-groups fill slots the tridiagonal leaves free.
static_assert(TRI_WG == LBM_WG, "wall tiles are LBM work-group sized");
template <int LMAX, bool MACRO>
__global__ __launch_bounds__(TRI_WG) void k_tridiag_walls(StepBuffers b, const float* __restrict__ src,
                                                          float* __restrict__ dst) {
  __shared__ double sA[TRI_WG];
  __shared__ double2 sB[TRI_WG];
  const int nt = b.n_tri_tiles, nw = (b.n_lbm_wall_tiles + 7) & ~7;
  const int i = blockIdx.x;
  if (i >= nw)
    tridiag_tile<LMAX>(b, b.tri_tiles[xcd_remap(i - nw, nt)], sA, sB);
  else if (i < b.n_lbm_wall_tiles)
    lbm_rows_tile<true>(b, b.lbm_wall_tiles[i], src, dst, MACRO, 0);
}

// table[j][k] = cprime_closed(j L, k kdy_scale) for chunks j = 1..C-1 (row
// j L is the last row of chunk j-1); row 0 of the table is unused
// every channel's table in one launch: channel chan0 + blockIdx.y, its
// C x nk entries grid-strided over x
__global__ __launch_bounds__(256) void k_cprime_table(StepBuffers b, int chan0) {
  const ChanDesc* __restrict__ cd = b.chans + chan0 + (int)blockIdx.y;
  const int nk = cd->nk, C = cd->tri_C, L = cd->tri_L;
  for (int t = blockIdx.x * 256 + (int)threadIdx.x; t < C * nk; t += (int)gridDim.x * 256) {
    const int j = t / nk, k = t - j * nk;
    b.cprime[cd->cp_off + t] = j == 0 ? 0.0 : cprime_closed(j * L, (double)k * cd->kdy_scale);
  }
}

// ---------------------------------------------------------------------------
// FP64 reductions per channel: partial sums per LBM tile (wave shuffles +
// LDS), then one work-group per channel adds its tiles in a fixed order.
// ---------------------------------------------------------------------------
// FP64 sum over the 64 lanes of a wave, result in lane 0
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

__global__ __launch_bounds__(LBM_WG) void k_reduce_partial(StepBuffers b, int snapshot) {
  __shared__ double red[LBM_WG / 64][8];
  const int2 tile = b.lbm_tiles[blockIdx.x];
  const ChanDesc* __restrict__ cd = b.chans + tile.x;
  const int nx = cd->nx, ny = cd->ny;
  const int node = tile.y + (int)threadIdx.x;
  double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  if (node < nx * ny) {
    const int y = node / nx;
    const long long o = cd->fld_off + node, st = b.macro_stride;
    const double rho = b.macro[M_RHO * st + o], ux = b.macro[M_UX * st + o];
    const double uy = b.macro[M_UY * st + o];
    const double cp = b.macro[M_CP * st + o], cn = b.macro[M_CN * st + o];
    const double ps = b.psi[o];
    const double uxp = b.ux_prev[o];
    v[0] = rho;
    v[1] = cp;
    v[2] = cn;
    if (y > 0 && y < ny - 1) {
      v[3] = cp - cn;
      v[4] = (ux - uxp) * (ux - uxp);
      v[5] = ux * ux;
    }
    v[6] = (isfinite(rho) && isfinite(ux) && isfinite(uy) && isfinite(cp) && isfinite(cn) && isfinite(ps)) ? 0.0
                                                                                                          : 1.0;
    if (snapshot) b.ux_prev[o] = ux;
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const double s = wave_sum(v[k]);
    if (lane == 0) red[wv][k] = s;
  }
  __syncthreads();
  if (threadIdx.x < 8) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < LBM_WG / 64; ++w) s += red[w][threadIdx.x];
    b.partials[(long long)blockIdx.x * 8 + threadIdx.x] = s;
  }
}

__global__ __launch_bounds__(256) void k_reduce_final(StepBuffers b) {
  __shared__ double red[4][8];
  const int c = blockIdx.x;
  const int t0 = b.red_first_tile[c], t1 = b.red_first_tile[c + 1];
  double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int t = t0 + (int)threadIdx.x; t < t1; t += 256) {
#pragma unroll
    for (int k = 0; k < 8; ++k) v[k] += b.partials[(long long)t * 8 + k];
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const double s = wave_sum(v[k]);
    if (lane == 0) red[wv][k] = s;
  }
  __syncthreads();
  if (threadIdx.x < 8)
    b.red_out[c * 8 + threadIdx.x] = (red[0][threadIdx.x] + red[1][threadIdx.x]) +
                                     (red[2][threadIdx.x] + red[3][threadIdx.x]);
}

// ---------------------------------------------------------------------------

// Storage references: row means of rho, c+, c- and their x-fluxes (one
// work-group per channel row) and the row's block records, then the exact
// change of reference of every stored population (delta += w_i (ref_old -
// ref_new), in FP64, rounded once).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_row_means(StepBuffers b, double* __restrict__ out,
                                                   float* __restrict__ blocks_out) {
  __shared__ double red[4][REF_ROWS];
  const int2 t = b.row_tiles[blockIdx.x];
  const ChanDesc* __restrict__ cd = b.chans + t.x;
  const int nx = cd->nx, y = t.y, nb = cd->nb;
  const long long rec = cd->ref_off + (long long)REF_STRIDE * y;
  const long long kn0 = cd->blk_off + (long long)BLK_STRIDE * nb * y;
  if (!cd->active) {
    // a frozen channel keeps the references its stored state is expressed in
    for (int i = threadIdx.x; i < REF_STRIDE; i += 256) out[rec + i] = b.refs[rec + i];
    for (int i = threadIdx.x; i < BLK_STRIDE * nb; i += 256) blocks_out[kn0 + i] = b.blocks[kn0 + i];
    return;
  }
  const long long o = cd->fld_off + (long long)y * nx, st = b.macro_stride;
  double v[REF_ROWS] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int x = threadIdx.x; x < nx; x += 256) {
    const double r = b.macro[M_RHO * st + o + x], u = b.macro[M_UX * st + o + x];
    const double cp = b.macro[M_CP * st + o + x], cn = b.macro[M_CN * st + o + x];
    v[0] += r;
    v[1] += cp;
    v[2] += cn;
    v[3] += r * u;
    v[4] += cp * u;
    v[5] += cn * u;
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < REF_ROWS; ++k) {
    const double s = wave_sum(v[k]);
    if (lane == 0) red[wv][k] = s;
  }
  // blocks: the mean over columns 64 k .. min(64 k + 64, nx) - 1, one wave
  // per block
  for (int k = wv; k < nb; k += 4) {
    const int x = REF_BLOCK * k + lane, cnt = min(REF_BLOCK, nx - REF_BLOCK * k);
    double br = 0.0, bp = 0.0, bc = 0.0;
    if (lane < cnt) {
      br = b.macro[M_RHO * st + o + x];
      bp = b.macro[M_CP * st + o + x];
      bc = b.macro[M_CN * st + o + x];
    }
    br = wave_sum(br);
    bp = wave_sum(bp);
    bc = wave_sum(bc);
    if (lane == 0) {
      float* bout = blocks_out + kn0 + (long long)BLK_STRIDE * k;
      bout[0] = (float)(br / (double)cnt);
      bout[1] = (float)(bp / (double)cnt);
      bout[2] = (float)(bc / (double)cnt);
      bout[3] = 0.0f;
    }
  }
  __syncthreads();
  if (threadIdx.x < REF_STRIDE) {
    const int i = threadIdx.x;
    const double s = i < REF_ROWS ? (red[0][i] + red[1][i]) + (red[2][i] + red[3][i]) : 0.0;
    // wall rows store reversed (fluid) or wall-rule (ions) populations whose
    // x-flux is not the row's macroscopic one: no flux reference there
    const bool wall = (y == 0 || y == cd->ny - 1);
    out[rec + i] = (i >= REF_ROWS || (wall && i >= 3)) ? 0.0 : s / (double)nx;
  }
}

// Frozen channels (not active) are left alone: their state may sit in either
// lattice, and k_row_means kept their references.
template <typename S>
__global__ __launch_bounds__(LBM_WG) void k_rebase(StepBuffers b, S* __restrict__ cur, S* __restrict__ other,
                                                   const double* __restrict__ new_refs,
                                                   const float* __restrict__ new_blocks) {
  const int2 tile = b.lbm_tiles[blockIdx.x];
  const ChanDesc* __restrict__ cd = b.chans + tile.x;
  if (!cd->active) return;
  const int nx = cd->nx, ny = cd->ny, nb = cd->nb;
  const int node = tile.y + (int)threadIdx.x;
  if (node >= nx * ny) return;
  const int y = node / nx, x = node - y * nx;
  S* __restrict__ P = (cd->swap ? other : cur) + cd->pop_off;
  const double* __restrict__ R0 = b.refs + cd->ref_off + (long long)REF_STRIDE * y;
  const double* __restrict__ R1 = new_refs + cd->ref_off + (long long)REF_STRIDE * y;
  const long long krow = cd->blk_off + (long long)BLK_STRIDE * nb * y;
  const float* __restrict__ K0 = b.blocks + krow;
  const float* __restrict__ K1 = new_blocks + krow;
  const double tc = cd->inv_cs2 * cd->inv_cfl;
#pragma unroll
  for (int s = 0; s < 3; ++s) {
    const double d = block_x(K0, x, s) - block_x(K1, x, s);
    const double dj = (R0[3 + s] - R1[3 + s]) * tc;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const double w = (k == 0) ? W0 : (k < 5 ? WS : WD);
      S* p = P + (long long)(9 * s + k) * cd->plane + node;
      *p = (S)((double)*p + w * (d + CX[k] * dj));
    }
  }
}

// ---------------------------------------------------------------------------
// PB-Picard start (LBM.cu:106-145): gpu_PBE and gpu_PBE_phi, all nodes
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(LBM_WG) void k_pb_fields(StepBuffers b) {
  const int2 tile = b.lbm_tiles[blockIdx.x];
  const ChanDesc* __restrict__ cd = b.chans + tile.x;
  const int node = tile.y + (int)threadIdx.x;
  if (node >= cd->nx * cd->ny) return;
  const long long o = cd->fld_off + node, st = b.macro_stride;
  const double psi = b.psi[o];
  const double cp = cd->c_bulk * exp(-cd->inv_vt * psi);  // LBM.cu:143
  const double cn = cd->c_bulk * exp(cd->inv_vt * psi);   // LBM.cu:144
  b.macro[M_CP * st + o] = cp;
  b.macro[M_CN * st + o] = cn;
  b.rho_e[o] = cp - cn;
}

__global__ __launch_bounds__(LBM_WG) void k_pb_relax(StepBuffers b, double* __restrict__ psi_old) {
  const int2 tile = b.lbm_tiles[blockIdx.x];
  const ChanDesc* __restrict__ cd = b.chans + tile.x;
  const int node = tile.y + (int)threadIdx.x;
  if (node >= cd->nx * cd->ny) return;
  const long long o = cd->fld_off + node;
  const double w = cd->pb_omega;
  const double v = w * b.psi[o] + (1.0 - w) * psi_old[o];  // LBM.cu:136
  b.psi[o] = v;
  psi_old[o] = v;
}

// ---------------------------------------------------------------------------
// Iterative Poisson solve (the reference's poisson_phi / gpu_poisson,
// poisson.cu:35-130): the 5-point finite-difference system
//   (psi_L + psi_R - 2 psi)/dx^2 + (psi_U + psi_D - 2 psi)/dy^2 = -pc rho_e
// on rows 1..ny-2, periodic in x, psi(0) and psi(ny-1) the wall potentials.
// gpu_poisson relaxes all nodes in place at once (a racy mix of Jacobi and
// Gauss-Seidel); one sweep here is a red-black SOR pass instead, the nodes
// with (x + y) even (color 0) and then those with (x + y) odd, each
//   psi_gs = (r (psi_L + psi_R) + psi_U + psi_D + pc dy^2 rho_e) * inv,
//   r = (dy/dx)^2, inv = 1 / (2 (1 + r)),
//   psi <- psi_gs (omega = 1) or psi + omega (psi_gs - psi),
// which for omega = 1 and dx = dy is gpu_poisson's update (poisson.cu:122-124).
// nx is even, so the colouring survives the periodic wrap.  Each work-group
// writes the largest |change| of its nodes to partials[blockIdx.x].
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
  return v;
}

__global__ __launch_bounds__(LBM_WG) void k_sor_sweep(StepBuffers b, int chan, int color, double omega, double r,
                                                     double inv, double* __restrict__ partials) {
  __shared__ double wmax[LBM_WG / 64];
  const ChanDesc* __restrict__ cd = b.chans + chan;
  const int nx = cd->nx, ny = cd->ny, hx = nx >> 1;
  const int i = blockIdx.x * LBM_WG + (int)threadIdx.x;
  double change = 0.0;
  if (i < hx * (ny - 2)) {
    const int y = 1 + i / hx;
    const int x = 2 * (i - (y - 1) * hx) + ((y + color) & 1);
    const int xm = x == 0 ? nx - 1 : x - 1, xp = x == nx - 1 ? 0 : x + 1;
    double* __restrict__ ps = b.psi + cd->fld_off;
    const double* __restrict__ row = ps + (long long)y * nx;
    const double* __restrict__ Wl = b.walls + cd->wall_off;
    const long long ws = b.walls_stride;
    const double up = (y == ny - 2) ? Wl[W_PSI_HI * ws + x] : row[nx + x];
    const double dn = (y == 1) ? Wl[W_PSI_LO * ws + x] : row[x - nx];
    const double old = row[x];
    const double src = cd->pc_dy2 * b.rho_e[cd->fld_off + (long long)y * nx + x];
    const double gs = (r * (row[xm] + row[xp]) + up + dn + src) * inv;
    const double nv = (omega == 1.0) ? gs : old + omega * (gs - old);
    ps[(long long)y * nx + x] = nv;
    change = fabs(nv - old);
  }
  change = wave_max(change);
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = change;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = wmax[0];
#pragma unroll
    for (int w = 1; w < LBM_WG / 64; ++w) m = fmax(m, wmax[w]);
    partials[blockIdx.x] = m;
  }
}

// out[0] = max of v[0..n) (one work-group)
__global__ __launch_bounds__(256) void k_max_final(const double* __restrict__ v, int n, double* __restrict__ out) {
  __shared__ double wmax[4];
  double m = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) m = fmax(m, v[i]);
  m = wave_max(m);
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) out[0] = fmax(fmax(wmax[0], wmax[1]), fmax(wmax[2], wmax[3]));
}

// psi(0) = W_PSI_LO, psi(ny-1) = W_PSI_HI of one channel (poisson.cu:96-103)
__global__ __launch_bounds__(256) void k_psi_wall_rows(StepBuffers b, int chan) {
  const ChanDesc* __restrict__ cd = b.chans + chan;
  const int nx = cd->nx, ny = cd->ny;
  const int x = blockIdx.x * 256 + (int)threadIdx.x;
  if (x >= nx) return;
  double* __restrict__ ps = b.psi + cd->fld_off;
  const double* __restrict__ Wl = b.walls + cd->wall_off;
  const long long ws = b.walls_stride;
  ps[x] = Wl[W_PSI_LO * ws + x];
  ps[(long long)(ny - 1) * nx + x] = Wl[W_PSI_HI * ws + x];
}

// ---------------------------------------------------------------------------
// Streaming-copy ceiling: dst = src over n float4 values with non-temporal
// loads and stores, grid-stride (the fastest plain copy measured on this
// part: the reference rate for the LBM kernel's HBM traffic and the kernel
// against which the FETCH_SIZE / WRITE_SIZE counters are calibrated)
// ---------------------------------------------------------------------------
typedef float v4f __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void k_copy_nt(const v4f* __restrict__ src, v4f* __restrict__ dst, long long n) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const v4f v = __builtin_nontemporal_load(src + i);
    __builtin_nontemporal_store(v, dst + i);
  }
}

hipError_t launch_copy_nt(const void* src, void* dst, long long n_float4, int grid, hipStream_t s) {
  hipLaunchKernelGGL(k_copy_nt, dim3(grid), dim3(256), 0, s, static_cast<const v4f*>(src), static_cast<v4f*>(dst),
                     n_float4);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <typename S, typename C>
static hipError_t lbm_dispatch(bool from_q, bool macro, bool walls, const StepBuffers& b, const void* src,
                               void* dst, hipStream_t s) {
  const S* sp = static_cast<const S*>(src);
  S* dp = static_cast<S*>(dst);
  const dim3 block(LBM_WG);
  if (from_q) {
    const dim3 grid(b.n_lbm_tiles);
    if (macro) hipLaunchKernelGGL((k_lbm<S, C, true, true>), grid, block, 0, s, b, b.lbm_tiles, sp, dp);
    else hipLaunchKernelGGL((k_lbm<S, C, true, false>), grid, block, 0, s, b, b.lbm_tiles, sp, dp);
    return hipGetLastError();
  }
  // the row kernel exists for FP32 storage with FP64 arithmetic
  const bool rows = std::is_same<S, float>::value && std::is_same<C, double>::value &&
                    (b.n_lbm_row_tiles > 0 || b.n_lbm_wall_tiles > 0);
  const int2* gen = rows ? b.lbm_gen_tiles : b.lbm_tiles;
  const int n_gen = rows ? b.n_lbm_gen_tiles : b.n_lbm_tiles;
  if (n_gen > 0) {
    const dim3 grid(n_gen);
    if (macro) hipLaunchKernelGGL((k_lbm<S, C, false, true>), grid, block, 0, s, b, gen, sp, dp);
    else hipLaunchKernelGGL((k_lbm<S, C, false, false>), grid, block, 0, s, b, gen, sp, dp);
  }
  if (rows) {
    const float* fs = reinterpret_cast<const float*>(src);
    float* fd = reinterpret_cast<float*>(dst);
    if (walls && b.n_lbm_wall_tiles > 0)
      hipLaunchKernelGGL((k_lbm_rows<true>), dim3(b.n_lbm_wall_tiles), block, 0, s, b, b.lbm_wall_tiles, fs, fd,
                         macro ? 1 : 0);
    if (b.n_lbm_row_tiles > 0)
      hipLaunchKernelGGL((k_lbm_rows<false>), dim3(b.n_lbm_row_tiles), block, 0, s, b, b.lbm_row_tiles, fs, fd,
                         macro ? 1 : 0);
  }
  return hipGetLastError();
}

// the row kernel's interior launch alone (no macroscopic planes): the LBM
// tile-shape autotuner's trial
hipError_t launch_lbm_rows_interior(const StepBuffers& b, const void* src, void* dst, hipStream_t s) {
  if (b.n_lbm_row_tiles > 0)
    hipLaunchKernelGGL((k_lbm_rows<false>), dim3(b.n_lbm_row_tiles), dim3(LBM_WG), 0, s, b, b.lbm_row_tiles,
                       static_cast<const float*>(src), static_cast<float*>(dst), 0);
  return hipGetLastError();
}

hipError_t launch_lbm(int precision, bool compute64, bool from_q, bool macro, const StepBuffers& b,
                      const void* src, void* dst, hipStream_t s, bool walls) {
  if (precision == 0) {
    if (compute64) return lbm_dispatch<float, double>(from_q, macro, walls, b, src, dst, s);
    return lbm_dispatch<float, float>(from_q, macro, walls, b, src, dst, s);
  }
  return lbm_dispatch<double, double>(from_q, macro, walls, b, src, dst, s);
}

// The FFT tile table holds the power-of-two tiles first (n_fft_pow2 of them),
// then the direct-DFT tiles.
hipError_t launch_poisson_fwd(const StepBuffers& b, hipStream_t s) {
  if (b.n_fft_pow2 > 0 && b.fft_pair_s > 0) {
    const dim3 grid(b.n_fft_pow2), block(FFT_WG_ROWS);
    if (b.fft_pair_s == 2) hipLaunchKernelGGL(k_fft_fwd2<2>, grid, block, b.fft_lds_pow2, s, b, b.fft_tiles, b.chans);
    else if (b.fft_pair_s == 4) hipLaunchKernelGGL(k_fft_fwd2<4>, grid, block, b.fft_lds_pow2, s, b, b.fft_tiles, b.chans);
    else if (b.fft_pair_s == FFT_PAIR_WIDE_S)
      hipLaunchKernelGGL(k_fft_fwd2w<FFT_PAIR_WIDE_S>, grid, block, b.fft_lds_pow2, s, b, b.fft_tiles, b.chans);
    else return hipErrorInvalidValue;
  } else if (b.n_fft_pow2 > 0 && b.fft_persist) {
    hipLaunchKernelGGL(k_fft_fwd_p, dim3(b.fft_grid), dim3(FFT_WG), FFT_LDS_PERSIST, s, b, b.fft_tiles, b.chans,
                       b.n_fft_pow2);
  } else if (b.n_fft_pow2 > 0) {
    if (b.fft_tw_lds)
      hipLaunchKernelGGL((k_fft_fwd<true, true>), dim3(b.n_fft_pow2), dim3(FFT_WG_ROWS), b.fft_lds_pow2, s, b, b.fft_tiles, b.chans, 0);
    else
      hipLaunchKernelGGL((k_fft_fwd<true, false>), dim3(b.n_fft_pow2), dim3(FFT_WG_ROWS), b.fft_lds_pow2, s, b, b.fft_tiles, b.chans, 0);
  }
  if (b.n_fft_tiles > b.n_fft_pow2)
    hipLaunchKernelGGL((k_fft_fwd<false, false>), dim3(b.n_fft_tiles - b.n_fft_pow2), dim3(FFT_WG_ROWS),
                       b.fft_lds_dft, s, b, b.fft_tiles, b.chans, b.n_fft_pow2);
  return hipGetLastError();
}

// resident work-groups per CU of the coupled step's tridiagonal launch
int tridiag_walls_blocks_per_cu() {
  int n = 0;
  const hipError_t e =
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_tridiag_walls<TRI_LMAX, false>, TRI_WG, 0);
  return e == hipSuccess ? n : 0;
}

int fft_one_tile_blocks_per_cu(bool twl, unsigned lds) {
  int n = 0;
  hipError_t e = twl ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_fft_fwd<true, true>, FFT_WG_ROWS, lds)
                     : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_fft_fwd<true, false>, FFT_WG_ROWS, lds);
  return e == hipSuccess ? n : 0;
}

hipError_t launch_poisson_inv(const StepBuffers& b, hipStream_t s) {
  if (b.n_fft_pow2 > 0 && b.fft_pair_s > 0) {
    const dim3 grid(b.n_fft_pow2), block(FFT_WG_ROWS);
    if (b.fft_pair_s == 2) hipLaunchKernelGGL(k_fft_inv2<2>, grid, block, b.fft_lds_pow2, s, b, b.fft_tiles, b.chans);
    else if (b.fft_pair_s == 4) hipLaunchKernelGGL(k_fft_inv2<4>, grid, block, b.fft_lds_pow2, s, b, b.fft_tiles, b.chans);
    else if (b.fft_pair_s == FFT_PAIR_WIDE_S)
      hipLaunchKernelGGL(k_fft_inv2w<FFT_PAIR_WIDE_S>, grid, block, b.fft_lds_pow2, s, b, b.fft_tiles, b.chans);
    else return hipErrorInvalidValue;
  } else if (b.n_fft_pow2 > 0 && b.fft_persist) {
    hipLaunchKernelGGL(k_fft_inv_p, dim3(b.fft_grid), dim3(FFT_WG), FFT_LDS_PERSIST, s, b, b.fft_tiles, b.chans,
                       b.n_fft_pow2);
  } else if (b.n_fft_pow2 > 0) {
    if (b.fft_tw_lds)
      hipLaunchKernelGGL((k_fft_inv<true, true>), dim3(b.n_fft_pow2), dim3(FFT_WG_ROWS), b.fft_lds_pow2, s, b, b.fft_tiles, b.chans, 0);
    else
      hipLaunchKernelGGL((k_fft_inv<true, false>), dim3(b.n_fft_pow2), dim3(FFT_WG_ROWS), b.fft_lds_pow2, s, b, b.fft_tiles, b.chans, 0);
  }
  if (b.n_fft_tiles > b.n_fft_pow2)
    hipLaunchKernelGGL((k_fft_inv<false, false>), dim3(b.n_fft_tiles - b.n_fft_pow2), dim3(FFT_WG_ROWS),
                       b.fft_lds_dft, s, b, b.fft_tiles, b.chans, b.n_fft_pow2);
  return hipGetLastError();
}

hipError_t launch_cprime_tables(const StepBuffers& b, int n_channels, long long max_entries, hipStream_t s) {
  // one work-group per 256 entries of the largest table, about 4096 in all
  // (the kernel strides over the rest)
  const long long cap = std::max<long long>(8, 4096 / std::max(1, n_channels));
  const int gx = (int)std::min<long long>(cap, std::max<long long>(1, (max_entries + 255) / 256));
  for (int c0 = 0; c0 < n_channels; c0 += 65535) {
    const int nc = std::min(65535, n_channels - c0);
    hipLaunchKernelGGL(k_cprime_table, dim3(gx, nc), dim3(256), 0, s, b, c0);
  }
  return hipGetLastError();
}

hipError_t launch_poisson_tri(const StepBuffers& b, hipStream_t s) {
  // the plan keeps every chunk at <= TRI_LMAX rows
  hipLaunchKernelGGL((k_tridiag<TRI_LMAX>), dim3(b.n_tri_tiles), dim3(TRI_WG), 0, s, b);
  return hipGetLastError();
}

hipError_t launch_poisson_tri_walls(bool macro, const StepBuffers& b, const void* src, void* dst, hipStream_t s) {
  const float* fs = static_cast<const float*>(src);
  float* fd = static_cast<float*>(dst);
  const dim3 grid(b.n_tri_tiles + ((b.n_lbm_wall_tiles + 7) & ~7));
  if (macro) hipLaunchKernelGGL((k_tridiag_walls<TRI_LMAX, true>), grid, dim3(TRI_WG), 0, s, b, fs, fd);
  else hipLaunchKernelGGL((k_tridiag_walls<TRI_LMAX, false>), grid, dim3(TRI_WG), 0, s, b, fs, fd);
  return hipGetLastError();
}

hipError_t launch_init_equilibrium(int precision, const StepBuffers& b, void* dst, hipStream_t s) {
  const dim3 grid(b.n_lbm_tiles), block(LBM_WG);
  if (precision == 0) hipLaunchKernelGGL((k_init_eq<float>), grid, block, 0, s, b, static_cast<float*>(dst));
  else hipLaunchKernelGGL((k_init_eq<double>), grid, block, 0, s, b, static_cast<double*>(dst));
  return hipGetLastError();
}

hipError_t launch_to_q(int precision, const StepBuffers& b, const void* src, void* dst, hipStream_t s) {
  const dim3 grid(b.n_lbm_tiles), block(LBM_WG);
  if (precision == 0)
    hipLaunchKernelGGL((k_to_q<float>), grid, block, 0, s, b, static_cast<const float*>(src),
                       static_cast<float*>(dst));
  else
    hipLaunchKernelGGL((k_to_q<double>), grid, block, 0, s, b, static_cast<const double*>(src),
                       static_cast<double*>(dst));
  return hipGetLastError();
}

hipError_t launch_row_means(const StepBuffers& b, double* out, float* blocks_out, hipStream_t s) {
  hipLaunchKernelGGL(k_row_means, dim3(b.n_row_tiles), dim3(256), 0, s, b, out, blocks_out);
  return hipGetLastError();
}

hipError_t launch_rebase(int precision, const StepBuffers& b, void* cur, void* other, const double* new_refs,
                         const float* new_blocks, hipStream_t s) {
  const dim3 grid(b.n_lbm_tiles), block(LBM_WG);
  if (precision == 0)
    hipLaunchKernelGGL((k_rebase<float>), grid, block, 0, s, b, static_cast<float*>(cur),
                       static_cast<float*>(other), new_refs, new_blocks);
  else
    hipLaunchKernelGGL((k_rebase<double>), grid, block, 0, s, b, static_cast<double*>(cur),
                       static_cast<double*>(other), new_refs, new_blocks);
  return hipGetLastError();
}

hipError_t launch_pb_fields(const StepBuffers& b, hipStream_t s) {
  hipLaunchKernelGGL(k_pb_fields, dim3(b.n_lbm_tiles), dim3(LBM_WG), 0, s, b);
  return hipGetLastError();
}

hipError_t launch_pb_relax(const StepBuffers& b, double* psi_old, hipStream_t s) {
  hipLaunchKernelGGL(k_pb_relax, dim3(b.n_lbm_tiles), dim3(LBM_WG), 0, s, b, psi_old);
  return hipGetLastError();
}

hipError_t launch_sor_sweep(const StepBuffers& b, int chan, int nx, int ny, int color, double omega, double r,
                            double inv, double* partials, hipStream_t s) {
  const int n = (nx / 2) * (ny - 2);
  hipLaunchKernelGGL(k_sor_sweep, dim3((n + LBM_WG - 1) / LBM_WG), dim3(LBM_WG), 0, s, b, chan, color, omega, r, inv,
                     partials);
  return hipGetLastError();
}

hipError_t launch_max_final(const double* v, int n, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_max_final, dim3(1), dim3(256), 0, s, v, n, out);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Run-time diagnostics of one channel from its macroscopic planes and psi:
//  * umax: the largest u_y over the channel, starting from 0, with the top
//    row taken as 2 u_y(ny-2) - u_y(ny-3) (record_umax, LBM.cu:1419-1451);
//  * the current's sum over the top row of c+ E_y, c+ extrapolated the same
//    way and E_y(ny-1) = E_y(ny-2) = (psi(ny-3) - psi(ny-1)) / (2 dy)
//    (current, LBM.cu:1394-1417; gpu_efield / gpu_bc, poisson.cu:145-174).
// k_diag_partial: partials[2 wg] = the largest u_y of the work-group's nodes
// (grid-stride), partials[2 wg + 1] = its share of the top-row sum;
// k_diag_final combines them in a fixed order into out[0] = umax,
// out[1] = the sum.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_diag_partial(const double* __restrict__ uy, const double* __restrict__ cp,
                                                      const double* __restrict__ psi, int nx, int ny,
                                                      double half_inv_dy, double* __restrict__ partials) {
  __shared__ double wm[4], ws[4];
  const long long nn = (long long)nx * ny, top = (long long)(ny - 1) * nx, nx2 = 2LL * nx;
  double m = 0.0, sum = 0.0;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < nn; i += (long long)gridDim.x * 256) {
    if (i < top) {
      m = fmax(m, uy[i]);
    } else {
      m = fmax(m, 2.0 * uy[i - nx] - uy[i - nx2]);
      const double c = 2.0 * cp[i - nx] - cp[i - nx2];
      const double ey = (psi[i - nx2] - psi[i]) * half_inv_dy;
      sum += c * ey;
    }
  }
  m = wave_max(m);
  sum = wave_sum(sum);
  if ((threadIdx.x & 63) == 0) {
    wm[threadIdx.x >> 6] = m;
    ws[threadIdx.x >> 6] = sum;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = fmax(fmax(wm[0], wm[1]), fmax(wm[2], wm[3]));
    partials[2 * blockIdx.x + 1] = (ws[0] + ws[1]) + (ws[2] + ws[3]);
  }
}

__global__ __launch_bounds__(256) void k_diag_final(const double* __restrict__ partials, int n,
                                                    double* __restrict__ out) {
  __shared__ double wm[4], ws[4];
  double m = 0.0, sum = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) {
    m = fmax(m, partials[2 * i]);
    sum += partials[2 * i + 1];
  }
  m = wave_max(m);
  sum = wave_sum(sum);
  if ((threadIdx.x & 63) == 0) {
    wm[threadIdx.x >> 6] = m;
    ws[threadIdx.x >> 6] = sum;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    out[0] = fmax(fmax(wm[0], wm[1]), fmax(wm[2], wm[3]));
    out[1] = (ws[0] + ws[1]) + (ws[2] + ws[3]);
  }
}

hipError_t launch_diagnostics(const double* uy, const double* cp, const double* psi, int nx, int ny,
                              double half_inv_dy, double* partials, int grid, double* out, hipStream_t s) {
  if (ny < 3 || grid < 1 || grid > DIAG_MAX_WG) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_diag_partial, dim3(grid), dim3(256), 0, s, uy, cp, psi, nx, ny, half_inv_dy, partials);
  hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(256), 0, s, partials, grid, out);
  return hipGetLastError();
}

hipError_t launch_psi_wall_rows(const StepBuffers& b, int chan, int nx, hipStream_t s) {
  hipLaunchKernelGGL(k_psi_wall_rows, dim3((nx + 255) / 256), dim3(256), 0, s, b, chan);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Scratch of the allocation-free reductions: a file-scope device array, one
// copy per GPU.  Its host-side name is not a device address: every kernel
// argument, copy or fill uses the address hipGetSymbolAddress returns for
// the current device, looked up once per device (noalloc_scratch_init, at
// context creation) and kept here.
// ---------------------------------------------------------------------------
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH_DOUBLES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

hipError_t noalloc_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return hipSuccess;
  int cur = -1;
  hipError_t e = hipGetDevice(&cur);
  if (e != hipSuccess) return e;
  if (cur != device) return hipErrorInvalidDevice;  // the caller makes the device current
  void* ptr = nullptr;
  e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  g_scratch_ptr[device] = static_cast<double*>(ptr);
  return hipSuccess;
}

double* noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return g_scratch_ptr[device];
}

hipError_t launch_reductions(const StepBuffers& b, int snapshot, hipStream_t s) {
  hipLaunchKernelGGL(k_reduce_partial, dim3(b.n_lbm_tiles), dim3(LBM_WG), 0, s, b, snapshot);
  hipLaunchKernelGGL(k_reduce_final, dim3(b.n_channels), dim3(256), 0, s, b);
  return hipGetLastError();
}

}  // namespace ekpnp
