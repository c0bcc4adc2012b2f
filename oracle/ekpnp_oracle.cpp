/*
 * ekpnp_oracle.cpp -- FP64 CPU restatement of the gyf135/EK-PNP coupled time
 * step.  TEST INFRASTRUCTURE ONLY: this library is the checker used by tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py.  Nothing in
 * the product (ek-pnp_amd/) links, loads or calls it.
 *
 * It keeps the reference's own storage and stage order so that every stage
 * can be compared line by line with the CUDA source:
 *   populations  f0[ny][nx] + f1/f2[8][ny][nx], same for h and temp (g)
 *                (LBM.h:124-126, index NX*(NY*(d-1)+y)+x, LBM.cu:36-39)
 *   stage 1  collide_save   LBM.cu:331-804
 *   stage 2  boundary       LBM.cu:806-860
 *   stage 3  stream         LBM.cu:862-906
 *   stage 4  bc_charge      LBM.cu:908-1024
 *   Poisson  method 0: odd extension + 2-D DFT + division by -mu
 *                     (poisson.cu:180-306, k tables main.cu:119-137)
 *            method 1: x-DFT + Thomas solve along y per wavenumber
 *                     (the same operator, see the note at poisson_xdft_thomas)
 *   E field  poisson.cu:145-174
 *   init     LBM.cu:147-160 (gpu_initialization), LBM.cu:195-305
 *            (gpu_init_equilibrium)
 *
 * Deliberate differences from the shipped source (each a parameter or flag,
 * so the literal behaviour stays reachable):
 *   - body force: force_coef (the commented formula of LBM.cu:388-389);
 *     force_coef = 0 is the shipped zeroed force.  The y = 0 mirror evaluates
 *     the force with node (x,1)'s own charge and field (the values that
 *     LBM.cu:434-437 computes for it).
 *   - the racy y = 0 read of f0/h0/temp0 at (x,1) (LBM.cu:403-405) always sees
 *     the pre-collision value: row 0 is collided before the other rows.
 *   - EKPNP_FLAG_G0_FLUID_U selects LBM.cu:493 (fluid |u|^2 in g_eq0);
 *     without it the drift velocity is used, as LBM.cu:242 does.
 *   - EKPNP_FLAG_ION_DRIFT_APPLIED adds Ext to the ion drift (LBM.cu:479-488
 *     uses the solved field only).
 *   - wall potentials and wall concentrations may vary along x.
 */
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/ekpnp.h"
#include "../include/ekpnp_seed.h"

typedef std::complex<double> cplx;

namespace {

const int EX[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
const int EY[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};

struct Oracle {
  ekpnp_channel_params p;
  int nx, ny;
  size_t n;  // nx*ny
  // populations, reference layout
  std::vector<double> f0, f1, f2, h0, h1, h2, g0, g1, g2;
  // macroscopic fields
  std::vector<double> rho, ux, uy, cp, cn, psi, ex, ey;
  std::vector<double> f0bc;  // [2][nx], LBM.cu:347-349
  // per-x wall values
  std::vector<double> psi_lo, psi_hi, cp_lo, cp_hi, cn_lo, cn_hi;
  long long steps;
};

inline size_t sidx(const Oracle& o, int x, int y) { return (size_t)o.nx * y + x; }
inline size_t nidx(const Oracle& o, int x, int y, int d) {
  return (size_t)o.nx * ((size_t)o.ny * (d - 1) + y) + x;
}

// ---------------------------------------------------------------------------
// Stage 1: gpu_collide_save, LBM.cu:331-804, one node.
// ---------------------------------------------------------------------------
struct NodeMoments {
  double rho, charge, temp, ux, uy;
};

// moments and force-corrected velocity of node (x,y) from the "0"/"1" arrays
// (LBM.cu:382-392); also used for the y = 0 mirror source (LBM.cu:403-442)
NodeMoments node_moments(const Oracle& o, int x, int y) {
  const ekpnp_channel_params& p = o.p;
  double ft[9], ht[9], tt[9];
  ft[0] = o.f0[sidx(o, x, y)];
  ht[0] = o.h0[sidx(o, x, y)];
  tt[0] = o.g0[sidx(o, x, y)];
  for (int d = 1; d < 9; ++d) {
    ft[d] = o.f1[nidx(o, x, y, d)];
    ht[d] = o.h1[nidx(o, x, y, d)];
    tt[d] = o.g1[nidx(o, x, y, d)];
  }
  NodeMoments m;
  m.rho = ft[0] + ft[1] + ft[2] + ft[3] + ft[4] + ft[5] + ft[6] + ft[7] + ft[8];
  double rhoinv = 1.0 / m.rho;
  m.charge = ht[0] + ht[1] + ht[2] + ht[3] + ht[4] + ht[5] + ht[6] + ht[7] + ht[8];
  m.temp = tt[0] + tt[1] + tt[2] + tt[3] + tt[4] + tt[5] + tt[6] + tt[7] + tt[8];
  double Ex = o.ex[sidx(o, x, y)];
  double Ey = o.ey[sidx(o, x, y)];
  double forcex = p.force_coef * (m.charge - m.temp) * (Ex + p.e_applied) + p.exf;
  double forcey = p.force_coef * (m.charge - m.temp) * Ey;
  m.ux = rhoinv * ((ft[1] + ft[5] + ft[8] - (ft[3] + ft[6] + ft[7])) / p.cfl + forcex * p.dt * 0.5);
  m.uy = rhoinv * ((ft[2] + ft[5] + ft[6] - (ft[4] + ft[7] + ft[8])) / p.cfl + forcey * p.dt * 0.5);
  return m;
}

void collide_node(Oracle& o, int x, int y) {
  const ekpnp_channel_params& p = o.p;
  const double cs2 = p.cs2, CFL = p.cfl, dt = p.dt;
  const double w0 = 4.0 / 9.0, ws = 1.0 / 9.0, wd = 1.0 / 36.0;  // LBM.h:105-107
  // LBM.cu:336-341
  double omega_plus = 1.0 / (p.nu / cs2 / dt + 1.0 / 2.0) / dt;
  double omega_minus = 1.0 / (p.magic_f / (p.nu / cs2 / dt) + 1.0 / 2.0) / dt;
  double omega_c_minus = 1.0 / (p.diff_pos / cs2 / dt + 1.0 / 2.0) / dt;
  double omega_c_plus = 1.0 / (p.magic_pos / (p.diff_pos / cs2 / dt) + 1.0 / 2.0) / dt;
  double omega_T_minus = 1.0 / (p.diff_neg / cs2 / dt + 1.0 / 2.0) / dt;
  double omega_T_plus = 1.0 / (p.magic_neg / (p.diff_neg / cs2 / dt) + 1.0 / 2.0) / dt;

  const size_t s = sidx(o, x, y);
  // LBM.cu:347-349
  if (y == 0) o.f0bc[x] = o.f0[s];
  if (y == o.ny - 1) o.f0bc[o.nx + x] = o.f0[s];

  double ft[9], ht[9], tt[9];
  ft[0] = o.f0[s];
  ht[0] = o.h0[s];
  tt[0] = o.g0[s];
  for (int d = 1; d < 9; ++d) {
    ft[d] = o.f1[nidx(o, x, y, d)];
    ht[d] = o.h1[nidx(o, x, y, d)];
    tt[d] = o.g1[nidx(o, x, y, d)];
  }
  NodeMoments m = node_moments(o, x, y);
  double rho = m.rho, charge = m.charge, temp = m.temp;
  double Ex = o.ex[s], Ey = o.ey[s];
  double forcex = p.force_coef * (charge - temp) * (Ex + p.e_applied) + p.exf;
  double forcey = p.force_coef * (charge - temp) * Ey;
  double ux = m.ux, uy = m.uy;
  if (y == 0) {
    // LBM.cu:402-443: velocity mirrored from node (x,1), pre-collision state
    NodeMoments mm = node_moments(o, x, 1);
    ux = -mm.ux;
    uy = -mm.uy;
  }
  // LBM.cu:447-451
  o.rho[s] = rho;
  o.ux[s] = ux;
  o.uy[s] = uy;
  o.cp[s] = charge;
  o.cn[s] = temp;

  // LBM.cu:468-552 equilibria
  double Exi = Ex + ((p.flags & EKPNP_FLAG_ION_DRIFT_APPLIED) ? p.e_applied : 0.0);
  double K = p.mob_pos, Kn = p.mob_neg;
  double omusq = 1.0 - 0.5 * (ux * ux + uy * uy) / cs2;
  double omusq_c = 1.0 - 0.5 * ((ux + K * Exi) * (ux + K * Exi) + (uy + K * Ey) * (uy + K * Ey)) / cs2;
  double omusq_cn = 1.0 - 0.5 * ((ux + Kn * Exi) * (ux + Kn * Exi) + (uy + Kn * Ey) * (uy + Kn * Ey)) / cs2;
  double tux = ux / cs2 / CFL, tuy = uy / cs2 / CFL;
  double tux_c = (ux + K * Exi) / cs2 / CFL, tuy_c = (uy + K * Ey) / cs2 / CFL;
  double tux_cn = (ux + Kn * Exi) / cs2 / CFL, tuy_cn = (uy + Kn * Ey) / cs2 / CFL;
  double fe[9], he[9], te[9];
  fe[0] = w0 * rho * omusq;
  he[0] = w0 * charge * omusq_c;
  te[0] = w0 * temp * ((p.flags & EKPNP_FLAG_G0_FLUID_U) ? omusq : omusq_cn);
  for (int d = 1; d < 9; ++d) {
    double w = (d < 5) ? ws : wd;
    double c = EX[d] * tux + EY[d] * tuy;
    fe[d] = w * rho * (omusq + c * (1.0 + 0.5 * c));
    c = EX[d] * tux_c + EY[d] * tuy_c;
    he[d] = w * charge * (omusq_c + c * (1.0 + 0.5 * c));
    c = EX[d] * tux_cn + EY[d] * tuy_cn;
    te[d] = w * temp * (omusq_cn + c * (1.0 + 0.5 * c));
  }
  // LBM.cu:556-572 Guo force populations
  double coe0 = w0 / cs2, coes = ws / cs2, coed = wd / cs2;
  double cflinv = 1.0 / CFL;
  double fpop[9];
  fpop[0] = coe0 * (-ux * forcex - uy * forcey);
  fpop[1] = coes * (((cflinv - ux) + (cflinv * ux) * cflinv / cs2) * forcex - uy * forcey);
  fpop[2] = coes * (-ux * forcex + ((cflinv - uy) + (cflinv * uy) * cflinv / cs2) * forcey);
  fpop[3] = coes * (((-cflinv - ux) + (cflinv * ux) * cflinv / cs2) * forcex - uy * forcey);
  fpop[4] = coes * (-ux * forcex + ((-cflinv - uy) + (cflinv * uy) * cflinv / cs2) * forcey);
  double cflinv2 = cflinv * cflinv / cs2;
  fpop[5] = coed * (((cflinv - ux) + (ux + uy) * cflinv2) * forcex + ((cflinv - uy) + (ux + uy) * cflinv2) * forcey);
  fpop[6] = coed * (((-cflinv - ux) + (ux - uy) * cflinv2) * forcex + ((cflinv - uy) + (-ux + uy) * cflinv2) * forcey);
  fpop[7] = coed * (((-cflinv - ux) + (ux + uy) * cflinv2) * forcex + ((-cflinv - uy) + (ux + uy) * cflinv2) * forcey);
  fpop[8] = coed * (((cflinv - ux) + (ux - uy) * cflinv2) * forcex + ((-cflinv - uy) + (-ux + uy) * cflinv2) * forcey);

  // LBM.cu:574-743: symmetric / antisymmetric parts; pairs (1,3) (2,4) (5,7) (6,8)
  const int OPP[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
  double sp = 1.0 - 0.5 * dt * omega_plus;   // LBM.cu:745-746
  double sm = 1.0 - 0.5 * dt * omega_minus;
  double tw0rp = omega_plus * dt, tw0rm = omega_minus * dt;
  double tw0cp = omega_c_plus * dt, tw0cm = omega_c_minus * dt;
  double tw0Tp = omega_T_plus * dt, tw0Tm = omega_T_minus * dt;
  double fout[9], hout[9], tout[9];
  for (int d = 0; d < 9; ++d) {
    int b = OPP[d];
    double fp, fm, fep, fem, hp, hm, hep, hem, tp, tm, tep, tem, forcep, forcem;
    if (d == 0) {
      fp = ft[0]; fm = 0.0; fep = fe[0]; fem = 0.0;
      hp = ht[0]; hm = 0.0; hep = he[0]; hem = 0.0;
      tp = tt[0]; tm = 0.0; tep = te[0]; tem = 0.0;
      forcep = fpop[0]; forcem = 0.0;
    } else {
      // for d in {1,2,5,6}: p = 0.5(d + opp), m = 0.5(d - opp); for the
      // opposite members the same p and the negated m (LBM.cu:577-597)
      int lo = (d == 1 || d == 2 || d == 5 || d == 6) ? d : b;
      int hi = OPP[lo];
      double sgn = (lo == d) ? 1.0 : -1.0;
      fp = 0.5 * (ft[lo] + ft[hi]); fm = sgn * (0.5 * (ft[lo] - ft[hi]));
      fep = 0.5 * (fe[lo] + fe[hi]); fem = sgn * (0.5 * (fe[lo] - fe[hi]));
      hp = 0.5 * (ht[lo] + ht[hi]); hm = sgn * (0.5 * (ht[lo] - ht[hi]));
      hep = 0.5 * (he[lo] + he[hi]); hem = sgn * (0.5 * (he[lo] - he[hi]));
      tp = 0.5 * (tt[lo] + tt[hi]); tm = sgn * (0.5 * (tt[lo] - tt[hi]));
      tep = 0.5 * (te[lo] + te[hi]); tem = sgn * (0.5 * (te[lo] - te[hi]));
      forcep = 0.5 * (fpop[lo] + fpop[hi]); forcem = sgn * (0.5 * (fpop[lo] - fpop[hi]));
    }
    double source = (d == 0) ? sp * fpop[0] : sp * forcep + sm * forcem;  // LBM.cu:748-756
    // LBM.cu:775-803
    fout[d] = ft[d] - (tw0rp * (fp - fep) + tw0rm * (fm - fem)) + dt * source;
    hout[d] = ht[d] - (tw0cp * (hp - hep) + tw0cm * (hm - hem));
    tout[d] = tt[d] - (tw0Tp * (tp - tep) + tw0Tm * (tm - tem));
  }
  o.f0[s] = fout[0];
  o.h0[s] = hout[0];
  o.g0[s] = tout[0];
  for (int d = 1; d < 9; ++d) {
    o.f2[nidx(o, x, y, d)] = fout[d];
    o.h2[nidx(o, x, y, d)] = hout[d];
    o.g2[nidx(o, x, y, d)] = tout[d];
  }
}

void stage_collide(Oracle& o) {
  // row 0 first: its mirror reads row 1 before row 1 is collided (F4)
  for (int x = 0; x < o.nx; ++x) collide_node(o, x, 0);
#pragma omp parallel for schedule(static)
  for (int y = 1; y < o.ny; ++y)
    for (int x = 0; x < o.nx; ++x) collide_node(o, x, y);
}

// ---------------------------------------------------------------------------
// Stage 2: gpu_boundary, LBM.cu:806-860
// ---------------------------------------------------------------------------
void stage_boundary(Oracle& o) {
  const ekpnp_channel_params& p = o.p;
  const double ws = 1.0 / 9.0, wd = 1.0 / 36.0;
  double multis = 2.0 * p.rho0 * p.uw / p.cs2 * ws / p.cfl;
  double multid = 2.0 * p.rho0 * p.uw / p.cs2 * wd / p.cfl;
  const int ny = o.ny;
  for (int x = 0; x < o.nx; ++x) {
    o.f0[sidx(o, x, 0)] = o.f0bc[x];
    o.f2[nidx(o, x, 0, 3)] = o.f1[nidx(o, x, 0, 1)];
    o.f2[nidx(o, x, 0, 4)] = o.f1[nidx(o, x, 0, 2)];
    o.f2[nidx(o, x, 0, 1)] = o.f1[nidx(o, x, 0, 3)];
    o.f2[nidx(o, x, 0, 2)] = o.f1[nidx(o, x, 0, 4)];
    o.f2[nidx(o, x, 0, 7)] = o.f1[nidx(o, x, 0, 5)];
    o.f2[nidx(o, x, 0, 8)] = o.f1[nidx(o, x, 0, 6)];
    o.f2[nidx(o, x, 0, 5)] = o.f1[nidx(o, x, 0, 7)];
    o.f2[nidx(o, x, 0, 6)] = o.f1[nidx(o, x, 0, 8)];

    int y = ny - 1;
    o.f0[sidx(o, x, y)] = o.f0bc[o.nx + x];
    o.f2[nidx(o, x, y, 3)] = o.f1[nidx(o, x, y, 1)] - multis;
    o.f2[nidx(o, x, y, 4)] = o.f1[nidx(o, x, y, 2)];
    o.f2[nidx(o, x, y, 1)] = o.f1[nidx(o, x, y, 3)] + multis;
    o.f2[nidx(o, x, y, 2)] = o.f1[nidx(o, x, y, 4)];
    o.f2[nidx(o, x, y, 7)] = o.f1[nidx(o, x, y, 5)] - multid;
    o.f2[nidx(o, x, y, 8)] = o.f1[nidx(o, x, y, 6)] + multid;
    o.f2[nidx(o, x, y, 5)] = o.f1[nidx(o, x, y, 7)] + multid;
    o.f2[nidx(o, x, y, 6)] = o.f1[nidx(o, x, y, 8)] - multid;
  }
}

// ---------------------------------------------------------------------------
// Stage 3: gpu_stream, LBM.cu:862-906 (pull, periodic in x and y)
// ---------------------------------------------------------------------------
void stage_stream(Oracle& o) {
  const int nx = o.nx, ny = o.ny;
#pragma omp parallel for schedule(static)
  for (int y = 0; y < ny; ++y) {
    for (int x = 0; x < nx; ++x) {
      int xp1 = (x + 1) % nx, yp1 = (y + 1) % ny;
      int xm1 = (nx + x - 1) % nx, ym1 = (ny + y - 1) % ny;
      std::vector<double>* A1[3] = {&o.f1, &o.h1, &o.g1};
      std::vector<double>* A2[3] = {&o.f2, &o.h2, &o.g2};
      for (int s = 0; s < 3; ++s) {
        std::vector<double>& d1 = *A1[s];
        const std::vector<double>& d2 = *A2[s];
        d1[nidx(o, x, y, 1)] = d2[nidx(o, xm1, y, 1)];
        d1[nidx(o, x, y, 2)] = d2[nidx(o, x, ym1, 2)];
        d1[nidx(o, x, y, 3)] = d2[nidx(o, xp1, y, 3)];
        d1[nidx(o, x, y, 4)] = d2[nidx(o, x, yp1, 4)];
        d1[nidx(o, x, y, 5)] = d2[nidx(o, xm1, ym1, 5)];
        d1[nidx(o, x, y, 6)] = d2[nidx(o, xp1, ym1, 6)];
        d1[nidx(o, x, y, 7)] = d2[nidx(o, xp1, yp1, 7)];
        d1[nidx(o, x, y, 8)] = d2[nidx(o, xm1, yp1, 8)];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Stage 4: gpu_bc_charge, LBM.cu:908-1024 (anti-bounce-back of the wall
// node's own post-collision populations)
// ---------------------------------------------------------------------------
void stage_bc_charge(Oracle& o) {
  const double w0 = 4.0 / 9.0, ws = 1.0 / 9.0, wd = 1.0 / 36.0;
  for (int wall = 0; wall < 2; ++wall) {
    int y = wall == 0 ? 0 : o.ny - 1;
    for (int x = 0; x < o.nx; ++x) {
      double cw[2] = {wall == 0 ? o.cp_lo[x] : o.cp_hi[x], wall == 0 ? o.cn_lo[x] : o.cn_hi[x]};
      std::vector<double>* A0[2] = {&o.h0, &o.g0};
      std::vector<double>* A1[2] = {&o.h1, &o.g1};
      std::vector<double>* A2[2] = {&o.h2, &o.g2};
      for (int s = 0; s < 2; ++s) {
        double multi0 = 2.0 * cw[s] * w0, multis = 2.0 * cw[s] * ws, multid = 2.0 * cw[s] * wd;
        std::vector<double>& a0 = *A0[s];
        std::vector<double>& a1 = *A1[s];
        const std::vector<double>& a2 = *A2[s];
        double t[9];
        for (int d = 1; d < 9; ++d) t[d] = a2[nidx(o, x, y, d)];
        a0[sidx(o, x, y)] = -a0[sidx(o, x, y)] + multi0;
        a1[nidx(o, x, y, 3)] = -t[1] + multis;
        a1[nidx(o, x, y, 4)] = -t[2] + multis;
        a1[nidx(o, x, y, 1)] = -t[3] + multis;
        a1[nidx(o, x, y, 2)] = -t[4] + multis;
        a1[nidx(o, x, y, 7)] = -t[5] + multid;
        a1[nidx(o, x, y, 8)] = -t[6] + multid;
        a1[nidx(o, x, y, 5)] = -t[7] + multid;
        a1[nidx(o, x, y, 6)] = -t[8] + multid;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// E field: gpu_efield + gpu_bc, poisson.cu:145-174
// ---------------------------------------------------------------------------
void efield(Oracle& o) {
  const int nx = o.nx, ny = o.ny;
  const double dx = o.p.dx, dy = o.p.dy;
  for (int y = 0; y < ny; ++y) {
    for (int x = 0; x < nx; ++x) {
      int xp1 = (x + 1) % nx, yp1 = (y + 1) % ny;
      int xm1 = (nx + x - 1) % nx, ym1 = (ny + y - 1) % ny;
      double phiL = o.psi[sidx(o, xm1, y)], phiR = o.psi[sidx(o, xp1, y)];
      double phiU = o.psi[sidx(o, x, yp1)], phiD = o.psi[sidx(o, x, ym1)];
      o.ex[sidx(o, x, y)] = 0.5 * (phiL - phiR) / dx;
      o.ey[sidx(o, x, y)] = 0.5 * (phiD - phiU) / dy;
    }
  }
  for (int x = 0; x < nx; ++x) {
    o.ey[sidx(o, x, 0)] = o.ey[sidx(o, x, 1)];
    o.ey[sidx(o, x, ny - 1)] = o.ey[sidx(o, x, ny - 2)];
  }
}

// ---------------------------------------------------------------------------
// discrete Fourier transforms (unnormalised; sign -1 forward, +1 inverse)
// ---------------------------------------------------------------------------
bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

// naive DFT of length n with stride; twiddle angle reduced mod n first
void dft(const cplx* in, cplx* out, int n, int sign) {
  std::vector<cplx> w(n);
  for (int j = 0; j < n; ++j) {
    double a = 2.0 * M_PI * (double)j / (double)n;
    w[j] = cplx(std::cos(a), sign * std::sin(a));
  }
  for (int k = 0; k < n; ++k) {
    cplx acc(0.0, 0.0);
    for (int t = 0; t < n; ++t) acc += in[t] * w[(long)k * t % n];
    out[k] = acc;
  }
}

// iterative radix-2 Cooley-Tukey, in place, n a power of two
void fft_pow2(cplx* a, int n, int sign) {
  for (int i = 1, j = 0; i < n; ++i) {
    int bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) std::swap(a[i], a[j]);
  }
  for (int len = 2; len <= n; len <<= 1) {
    for (int i = 0; i < n; i += len) {
      for (int k = 0; k < len / 2; ++k) {
        double ang = 2.0 * M_PI * (double)k / (double)len;
        cplx w(std::cos(ang), sign * std::sin(ang));
        cplx u = a[i + k], v = a[i + k + len / 2] * w;
        a[i + k] = u + v;
        a[i + k + len / 2] = u - v;
      }
    }
  }
}

void transform(std::vector<cplx>& v, int sign) {
  int n = (int)v.size();
  if (is_pow2(n)) {
    fft_pow2(v.data(), n, sign);
  } else {
    std::vector<cplx> out(n);
    dft(v.data(), out.data(), n, sign);
    v.swap(out);
  }
}

// kx table, main.cu:119-128
double kx_of(const Oracle& o, int i) {
  int NX = o.nx;
  if (i <= NX / 2) return (double)i * 2.0 * M_PI / o.p.lx;
  return ((double)i - NX) * 2.0 * M_PI / o.p.lx;
}

// ---------------------------------------------------------------------------
// Poisson method 0: the reference's odd extension + 2-D DFT + (-1/mu)
// (poisson.cu:180-306; ky table main.cu:129-137)
// ---------------------------------------------------------------------------
void poisson_odd_extension(Oracle& o) {
  const int NX = o.nx, NY = o.ny, NE = 2 * (NY - 1);
  const double dy = o.p.dy, pc = o.p.poisson_coef;
  std::vector<cplx> ext((size_t)NE * NX, cplx(0, 0));
  auto q = [&](int x, int y) { return pc * (o.cp[sidx(o, x, y)] - o.cn[sidx(o, x, y)]); };
  // odd_extension, poisson.cu:220-263
  for (int y = 0; y < NE; ++y) {
    for (int x = 0; x < NX; ++x) {
      double v = 0.0;
      if (y == 0) v = 0.0;
      else if (y == 1) v = -q(x, y) - o.psi_lo[x] / dy / dy;
      else if (y > 1 && y < NY - 2) v = -q(x, y);
      else if (y == NY - 2) v = -q(x, y) - o.psi_hi[x] / dy / dy;
      else if (y == NY - 1) v = 0.0;
      else if (y == NY) v = q(x, NE - y) + o.psi_hi[x] / dy / dy;
      else if (y > NY && y < NE - 1) v = q(x, NE - y);
      else if (y == NE - 1) v = q(x, 1) + o.psi_lo[x] / dy / dy;
      ext[(size_t)y * NX + x] = cplx(v, 0.0);
    }
  }
  // forward 2-D transform: rows (x) then columns (y)
  auto transform2d = [&](int sign) {
    std::vector<cplx> row(NX), col(NE);
    for (int y = 0; y < NE; ++y) {
      for (int x = 0; x < NX; ++x) row[x] = ext[(size_t)y * NX + x];
      transform(row, sign);
      for (int x = 0; x < NX; ++x) ext[(size_t)y * NX + x] = row[x];
    }
    for (int x = 0; x < NX; ++x) {
      for (int y = 0; y < NE; ++y) col[y] = ext[(size_t)y * NX + x];
      transform(col, sign);
      for (int y = 0; y < NE; ++y) ext[(size_t)y * NX + x] = col[y];
    }
  };
  transform2d(-1);
  // gpu_derivative, poisson.cu:274-283
  for (int j = 0; j < NE; ++j) {
    double J = (j <= NE / 2) ? (double)j * 2.0 * M_PI / (NE * dy)
                             : ((double)j - NE) * 2.0 * M_PI / (NE * dy);
    for (int i = 0; i < NX; ++i) {
      double I = kx_of(o, i);
      double mu = (4.0 / dy / dy) * (std::sin(J * dy * 0.5) * std::sin(J * dy * 0.5)) + I * I;
      if (j == 0 && i == 0) mu = 1.0;
      ext[(size_t)j * NX + i] = -ext[(size_t)j * NX + i] / mu;
    }
  }
  transform2d(+1);
  // odd_extract, poisson.cu:294-306
  const double SIZE = (double)NX * NE;
  for (int y = 0; y < NY; ++y) {
    for (int x = 0; x < NX; ++x) {
      if (y == 0) o.psi[sidx(o, x, y)] = o.psi_lo[x];
      else if (y == NY - 1) o.psi[sidx(o, x, y)] = o.psi_hi[x];
      else o.psi[sidx(o, x, y)] = ext[(size_t)y * NX + x].real() / SIZE;
    }
  }
}

// ---------------------------------------------------------------------------
// Poisson method 1: DFT along x, then per wavenumber the tridiagonal system
//   (psi_{y-1} - 2 psi_y + psi_{y+1})/dy^2 - kx^2 psi_y = rhs_y,
//   y = 1..NY-2, psi_0 = psi_{NY-1} = 0 after the wall values are lifted into
//   rows 1 and NY-2.  The odd extension of method 0 diagonalises exactly
//   this operator in y (its symbol 4/dy^2 sin^2(ky dy/2) is the spectrum of
//   the second difference with Dirichlet ends), so both methods solve the
//   same linear system.
// ---------------------------------------------------------------------------
void poisson_xdft_thomas(Oracle& o) {
  const int NX = o.nx, NY = o.ny, M = NY - 2;
  const double dy = o.p.dy, pc = o.p.poisson_coef;
  std::vector<cplx> spec((size_t)M * NX);
#pragma omp parallel for schedule(static)
  for (int y = 1; y <= NY - 2; ++y) {
    std::vector<cplx> row(NX);
    for (int x = 0; x < NX; ++x) {
      double v = -pc * (o.cp[sidx(o, x, y)] - o.cn[sidx(o, x, y)]);
      if (y == 1) v -= o.psi_lo[x] / dy / dy;
      if (y == NY - 2) v -= o.psi_hi[x] / dy / dy;
      row[x] = cplx(v, 0.0);
    }
    transform(row, -1);
    for (int k = 0; k < NX; ++k) spec[(size_t)(y - 1) * NX + k] = row[k];
  }
#pragma omp parallel for schedule(static)
  for (int k = 0; k < NX; ++k) {
    double kx = kx_of(o, k);
    double b = 1.0 / (dy * dy);
    double d = -2.0 / (dy * dy) - kx * kx;
    std::vector<double> cprime(M);
    std::vector<cplx> dprime(M);
    cprime[0] = b / d;
    dprime[0] = spec[k] / d;
    for (int i = 1; i < M; ++i) {
      double m = d - b * cprime[i - 1];
      cprime[i] = b / m;
      dprime[i] = (spec[(size_t)i * NX + k] - b * dprime[i - 1]) / m;
    }
    spec[(size_t)(M - 1) * NX + k] = dprime[M - 1];
    for (int i = M - 2; i >= 0; --i)
      spec[(size_t)i * NX + k] = dprime[i] - cprime[i] * spec[(size_t)(i + 1) * NX + k];
  }
#pragma omp parallel for schedule(static)
  for (int y = 1; y <= NY - 2; ++y) {
    std::vector<cplx> row(NX);
    for (int k = 0; k < NX; ++k) row[k] = spec[(size_t)(y - 1) * NX + k];
    transform(row, +1);
    for (int x = 0; x < NX; ++x) o.psi[sidx(o, x, y)] = row[x].real() / (double)NX;
  }
  for (int x = 0; x < NX; ++x) {
    o.psi[sidx(o, x, 0)] = o.psi_lo[x];
    o.psi[sidx(o, x, NY - 1)] = o.psi_hi[x];
  }
}

void fast_poisson(Oracle& o, int method) {
  if (method == 0) poisson_odd_extension(o);
  else poisson_xdft_thomas(o);
  efield(o);
}

// ---------------------------------------------------------------------------
// gpu_init_equilibrium, LBM.cu:195-305 (writes the "0" and "1" arrays)
// ---------------------------------------------------------------------------
void init_equilibrium(Oracle& o) {
  const ekpnp_channel_params& p = o.p;
  const double cs2 = p.cs2, CFL = p.cfl;
  const double w0 = 4.0 / 9.0, ws = 1.0 / 9.0, wd = 1.0 / 36.0;
  const double K = p.mob_pos, Kn = p.mob_neg;
#pragma omp parallel for schedule(static)
  for (int y = 0; y < o.ny; ++y) {
    for (int x = 0; x < o.nx; ++x) {
      size_t s = sidx(o, x, y);
      double rho = o.rho[s], ux = o.ux[s], uy = o.uy[s], charge = o.cp[s];
      double Ex = o.ex[s] + ((p.flags & EKPNP_FLAG_ION_DRIFT_APPLIED) ? p.e_applied : 0.0);
      double Ey = o.ey[s], Temp = o.cn[s];
      double omusq = 1.0 - 0.5 * (ux * ux + uy * uy) / cs2;
      double omusq_c = 1.0 - 0.5 * ((ux + K * Ex) * (ux + K * Ex) + (uy + K * Ey) * (uy + K * Ey)) / cs2;
      double omusq_cn = 1.0 - 0.5 * ((ux + Kn * Ex) * (ux + Kn * Ex) + (uy + Kn * Ey) * (uy + Kn * Ey)) / cs2;
      double tux = ux / cs2 / CFL, tuy = uy / cs2 / CFL;
      double tux_c = (ux + K * Ex) / cs2 / CFL, tuy_c = (uy + K * Ey) / cs2 / CFL;
      double tux_cn = (ux + Kn * Ex) / cs2 / CFL, tuy_cn = (uy + Kn * Ey) / cs2 / CFL;
      o.f0[s] = w0 * rho * omusq;
      o.h0[s] = w0 * charge * omusq_c;
      o.g0[s] = w0 * Temp * omusq_cn;
      for (int d = 1; d < 9; ++d) {
        double w = (d < 5) ? ws : wd;
        double c = EX[d] * tux + EY[d] * tuy;
        o.f1[nidx(o, x, y, d)] = w * rho * (omusq + c * (1.0 + 0.5 * c));
        c = EX[d] * tux_c + EY[d] * tuy_c;
        o.h1[nidx(o, x, y, d)] = w * charge * (omusq_c + c * (1.0 + 0.5 * c));
        c = EX[d] * tux_cn + EY[d] * tuy_cn;
        o.g1[nidx(o, x, y, d)] = w * Temp * (omusq_cn + c * (1.0 + 0.5 * c));
      }
    }
  }
}

}  // namespace

// ===========================================================================
// C interface (ctypes)
// ===========================================================================
extern "C" {

typedef Oracle ekpnp_oracle;

void ekpnp_oracle_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

int ekpnp_oracle_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

ekpnp_oracle* ekpnp_oracle_create(const ekpnp_channel_params* p) {
  if (!p || p->nx < 2 || p->ny < 4) return nullptr;
  Oracle* o = new Oracle();
  o->p = *p;
  o->nx = p->nx;
  o->ny = p->ny;
  o->n = (size_t)p->nx * p->ny;
  size_t n = o->n, n8 = 8 * o->n;
  o->f0.assign(n, 0); o->h0.assign(n, 0); o->g0.assign(n, 0);
  o->f1.assign(n8, 0); o->h1.assign(n8, 0); o->g1.assign(n8, 0);
  o->f2.assign(n8, 0); o->h2.assign(n8, 0); o->g2.assign(n8, 0);
  o->rho.assign(n, 0); o->ux.assign(n, 0); o->uy.assign(n, 0);
  o->cp.assign(n, 0); o->cn.assign(n, 0); o->psi.assign(n, 0);
  o->ex.assign(n, 0); o->ey.assign(n, 0);
  o->f0bc.assign(2 * (size_t)p->nx, 0);
  o->psi_lo.assign(p->nx, p->psi_lo); o->psi_hi.assign(p->nx, p->psi_hi);
  o->cp_lo.assign(p->nx, p->cpos_lo); o->cp_hi.assign(p->nx, p->cpos_hi);
  o->cn_lo.assign(p->nx, p->cneg_lo); o->cn_hi.assign(p->nx, p->cneg_hi);
  o->steps = 0;
  return o;
}

void ekpnp_oracle_destroy(ekpnp_oracle* o) { delete o; }

void ekpnp_oracle_set_wall_profile(ekpnp_oracle* o, const double* psi_lo, const double* psi_hi,
                                   const double* cp_lo, const double* cp_hi,
                                   const double* cn_lo, const double* cn_hi) {
  int nx = o->nx;
  if (psi_lo) o->psi_lo.assign(psi_lo, psi_lo + nx);
  if (psi_hi) o->psi_hi.assign(psi_hi, psi_hi + nx);
  if (cp_lo) o->cp_lo.assign(cp_lo, cp_lo + nx);
  if (cp_hi) o->cp_hi.assign(cp_hi, cp_hi + nx);
  if (cn_lo) o->cn_lo.assign(cn_lo, cn_lo + nx);
  if (cn_hi) o->cn_hi.assign(cn_hi, cn_hi + nx);
}

// mode EKPNP_INIT_COLD_NOISE or EKPNP_INIT_REFERENCE; channel = RNG stream id
int ekpnp_oracle_initialization(ekpnp_oracle* o, int mode, uint64_t seed, int channel,
                                int poisson_method) {
  const ekpnp_channel_params& p = o->p;
  std::fill(o->ux.begin(), o->ux.end(), 0.0);
  std::fill(o->uy.begin(), o->uy.end(), 0.0);
  std::fill(o->rho.begin(), o->rho.end(), p.rho0);
  if (mode == EKPNP_INIT_REFERENCE || mode == EKPNP_INIT_PB_PICARD) {
    // LBM.cu:147-160
    std::fill(o->cp.begin(), o->cp.end(), 0.0);
    std::fill(o->cn.begin(), o->cn.end(), 0.0);
    std::fill(o->psi.begin(), o->psi.end(), p.psi_lo);
    std::fill(o->ex.begin(), o->ex.end(), 0.0);
    std::fill(o->ey.begin(), o->ey.end(), 0.0);
    if (mode == EKPNP_INIT_REFERENCE) return 0;
    // LBM.cu:107-127: 501 Picard iterations with under-relaxation PB_omega
    std::vector<double> psi_old = o->psi;
    const int iters = (int)p.pb_iterations;
    for (int it = 0; it < iters; ++it) {
      for (size_t i = 0; i < o->n; ++i) {  // gpu_PBE, LBM.cu:139-145
        o->cp[i] = p.c_bulk * std::exp(-p.inv_thermal_voltage * o->psi[i]);
        o->cn[i] = p.c_bulk * std::exp(p.inv_thermal_voltage * o->psi[i]);
      }
      fast_poisson(*o, poisson_method);
      for (size_t i = 0; i < o->n; ++i) {  // gpu_PBE_phi, LBM.cu:132-137
        o->psi[i] = p.pb_omega * o->psi[i] + (1.0 - p.pb_omega) * psi_old[i];
        psi_old[i] = o->psi[i];
      }
    }
    return 0;
  }
  if (mode != EKPNP_INIT_COLD_NOISE) return -1;
  ekpnp_cold_field(o->cp.data(), o->nx, o->ny, p.c_bulk, p.init_noise, o->cp_lo.data(),
                   o->cp_hi.data(), seed, channel, 0);
  ekpnp_cold_field(o->cn.data(), o->nx, o->ny, p.c_bulk, p.init_noise, o->cn_lo.data(),
                   o->cn_hi.data(), seed, channel, 1);
  fast_poisson(*o, poisson_method);
  return 0;
}

// psi may be NULL: solved from c+ - c- with the given method
void ekpnp_oracle_set_fields(ekpnp_oracle* o, const double* rho, const double* ux,
                             const double* uy, const double* cp, const double* cn,
                             const double* psi, int poisson_method) {
  size_t n = o->n;
  o->rho.assign(rho, rho + n);
  o->ux.assign(ux, ux + n);
  o->uy.assign(uy, uy + n);
  o->cp.assign(cp, cp + n);
  o->cn.assign(cn, cn + n);
  if (psi) {
    o->psi.assign(psi, psi + n);
    efield(*o);
  } else {
    fast_poisson(*o, poisson_method);
  }
}

void ekpnp_oracle_init_equilibrium(ekpnp_oracle* o) { init_equilibrium(*o); }

// individual stages (unit tests)
void ekpnp_oracle_stage_collide(ekpnp_oracle* o) { stage_collide(*o); }
void ekpnp_oracle_stage_boundary(ekpnp_oracle* o) { stage_boundary(*o); }
void ekpnp_oracle_stage_stream(ekpnp_oracle* o) { stage_stream(*o); }
void ekpnp_oracle_stage_bc_charge(ekpnp_oracle* o) { stage_bc_charge(*o); }
void ekpnp_oracle_efield(ekpnp_oracle* o) { efield(*o); }

// stream_collide_save, LBM.cu:307-329
void ekpnp_oracle_stream_collide_save(ekpnp_oracle* o) {
  stage_collide(*o);
  stage_boundary(*o);
  stage_stream(*o);
  stage_bc_charge(*o);
}

void ekpnp_oracle_fast_poisson(ekpnp_oracle* o, int method) { fast_poisson(*o, method); }

// main.cu:194-199, n times
void ekpnp_oracle_step(ekpnp_oracle* o, int n, int method) {
  for (int i = 0; i < n; ++i) {
    ekpnp_oracle_stream_collide_save(o);
    fast_poisson(*o, method);
    o->steps++;
  }
}

// wall-clock seconds of n steps (init and I/O excluded)
double ekpnp_oracle_time_steps(ekpnp_oracle* o, int n, int method) {
  auto t0 = std::chrono::steady_clock::now();
  ekpnp_oracle_step(o, n, method);
  auto t1 = std::chrono::steady_clock::now();
  return std::chrono::duration<double>(t1 - t0).count();
}

void ekpnp_oracle_get_field(const ekpnp_oracle* o, int field, double* out) {
  const std::vector<double>* src = nullptr;
  switch (field) {
    case EKPNP_FIELD_RHO: src = &o->rho; break;
    case EKPNP_FIELD_UX: src = &o->ux; break;
    case EKPNP_FIELD_UY: src = &o->uy; break;
    case EKPNP_FIELD_CPOS: src = &o->cp; break;
    case EKPNP_FIELD_CNEG: src = &o->cn; break;
    case EKPNP_FIELD_PSI: src = &o->psi; break;
    case EKPNP_FIELD_EX: src = &o->ex; break;
    case EKPNP_FIELD_EY: src = &o->ey; break;
    default: return;
  }
  std::memcpy(out, src->data(), o->n * sizeof(double));
}

// [9][ny][nx]: 0 = rest array, 1..8 = the "1" arrays
void ekpnp_oracle_get_populations(const ekpnp_oracle* o, int species, double* out) {
  const std::vector<double>* a0 = species == 0 ? &o->f0 : species == 1 ? &o->h0 : &o->g0;
  const std::vector<double>* a1 = species == 0 ? &o->f1 : species == 1 ? &o->h1 : &o->g1;
  std::memcpy(out, a0->data(), o->n * sizeof(double));
  std::memcpy(out + o->n, a1->data(), 8 * o->n * sizeof(double));
}

void ekpnp_oracle_set_populations(ekpnp_oracle* o, int species, const double* in) {
  std::vector<double>* a0 = species == 0 ? &o->f0 : species == 1 ? &o->h0 : &o->g0;
  std::vector<double>* a1 = species == 0 ? &o->f1 : species == 1 ? &o->h1 : &o->g1;
  std::memcpy(a0->data(), in, o->n * sizeof(double));
  std::memcpy(a1->data(), in + o->n, 8 * o->n * sizeof(double));
}

// the post-collision "2" arrays (stage tests)
void ekpnp_oracle_get_post(const ekpnp_oracle* o, int species, double* out) {
  const std::vector<double>* a2 = species == 0 ? &o->f2 : species == 1 ? &o->h2 : &o->g2;
  std::memcpy(out, a2->data(), 8 * o->n * sizeof(double));
}

void ekpnp_oracle_set_post(ekpnp_oracle* o, int species, const double* in) {
  std::vector<double>* a2 = species == 0 ? &o->f2 : species == 1 ? &o->h2 : &o->g2;
  std::memcpy(a2->data(), in, 8 * o->n * sizeof(double));
}

void ekpnp_oracle_set_field(ekpnp_oracle* o, int field, const double* in) {
  std::vector<double>* dst = nullptr;
  switch (field) {
    case EKPNP_FIELD_RHO: dst = &o->rho; break;
    case EKPNP_FIELD_UX: dst = &o->ux; break;
    case EKPNP_FIELD_UY: dst = &o->uy; break;
    case EKPNP_FIELD_CPOS: dst = &o->cp; break;
    case EKPNP_FIELD_CNEG: dst = &o->cn; break;
    case EKPNP_FIELD_PSI: dst = &o->psi; break;
    case EKPNP_FIELD_EX: dst = &o->ex; break;
    case EKPNP_FIELD_EY: dst = &o->ey; break;
    default: return;
  }
  std::memcpy(dst->data(), in, o->n * sizeof(double));
}

long long ekpnp_oracle_steps(const ekpnp_oracle* o) { return o->steps; }

}  // extern "C"
