// sanitize_main.cpp -- drives the FP64 CPU oracle (test infrastructure)
// through every stage on small grids for the AddressSanitizer /
// UndefinedBehaviorSanitizer build (make -C oracle sanitize): a 10-wide
// channel (direct DFT Poisson) with both Poisson methods, the
// gpu_initialization and PB-Picard starts, a 16-wide channel with x-varying
// walls, moving top wall and both flags, stage by stage, and the
// field/population accessors.  A sanitizer finding aborts.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../include/ekpnp.h"

extern "C" {
struct Oracle;
Oracle* ekpnp_oracle_create(const ekpnp_channel_params* p);
void ekpnp_oracle_destroy(Oracle* o);
void ekpnp_oracle_set_threads(int n);
void ekpnp_oracle_set_wall_profile(Oracle* o, const double*, const double*, const double*, const double*,
                                   const double*, const double*);
int ekpnp_oracle_initialization(Oracle* o, int mode, uint64_t seed, int channel, int method);
void ekpnp_oracle_init_equilibrium(Oracle* o);
void ekpnp_oracle_stage_collide(Oracle* o);
void ekpnp_oracle_stage_boundary(Oracle* o);
void ekpnp_oracle_stage_stream(Oracle* o);
void ekpnp_oracle_stage_bc_charge(Oracle* o);
void ekpnp_oracle_efield(Oracle* o);
void ekpnp_oracle_fast_poisson(Oracle* o, int method);
void ekpnp_oracle_step(Oracle* o, int n, int method);
void ekpnp_oracle_get_field(const Oracle* o, int field, double* out);
void ekpnp_oracle_set_field(Oracle* o, int field, const double* in);
void ekpnp_oracle_get_populations(const Oracle* o, int species, double* out);
void ekpnp_oracle_set_populations(Oracle* o, int species, const double* in);
void ekpnp_oracle_get_post(const Oracle* o, int species, double* out);
void ekpnp_oracle_set_post(Oracle* o, int species, const double* in);
}

// lattice units: dx = dt = 1, cs^2 = 1/3, TRT magic 1/4, D = K = 0.05
static ekpnp_channel_params lattice(int nx, int ny, double tau, double zeta, double kappa_h, double ex) {
  ekpnp_channel_params p;
  std::memset(&p, 0, sizeof p);
  p.nx = nx;
  p.ny = ny;
  p.dx = p.dy = p.dt = p.cfl = 1.0;
  p.cs2 = 1.0 / 3.0;
  p.rho0 = 1.0;
  p.nu = (tau - 0.5) / 3.0;
  p.diff_pos = p.diff_neg = 0.05;
  p.mob_pos = 0.05;
  p.mob_neg = -0.05;
  p.magic_f = p.magic_pos = p.magic_neg = 0.25;
  p.psi_lo = p.psi_hi = zeta;
  p.cpos_lo = p.cpos_hi = std::exp(-zeta);
  p.cneg_lo = p.cneg_hi = std::exp(zeta);
  p.c_bulk = 1.0;
  const double kappa = kappa_h / (0.5 * (ny - 1));
  p.poisson_coef = p.force_coef = 0.5 * kappa * kappa;
  p.e_applied = ex;
  p.lx = nx;
  p.init_noise = 1e-3;
  p.inv_thermal_voltage = 1.0;
  p.pb_omega = 0.05;
  p.pb_iterations = 20;
  return p;
}

static bool finite(const std::vector<double>& v) {
  for (double x : v)
    if (!std::isfinite(x)) return false;
  return true;
}

int main() {
  ekpnp_oracle_set_threads(2);
  int bad = 0;
  // non-power-of-two x rows (direct DFT), both Poisson methods,
  // gpu_initialization and the PB-Picard start
  ekpnp_channel_params r = lattice(10, 21, 0.9, -1.0, 3.0, 1e-4);
  for (int method = 0; method < 2; ++method) {
    for (int mode = 1; mode <= 2; ++mode) {
      Oracle* o = ekpnp_oracle_create(&r);
      ekpnp_oracle_initialization(o, mode, 0, 0, method);
      ekpnp_oracle_init_equilibrium(o);
      ekpnp_oracle_step(o, 5, method);
      std::vector<double> f((size_t)r.nx * r.ny);
      for (int k = 0; k < EKPNP_FIELD_COUNT; ++k) {
        ekpnp_oracle_get_field(o, k, f.data());
        bad += !finite(f);
      }
      ekpnp_oracle_destroy(o);
    }
  }
  // a lattice channel with x-varying walls, stage by stage, accessors
  ekpnp_channel_params p = lattice(16, 12, 0.8, -1.5, 3.0, 1e-4);
  p.flags = EKPNP_FLAG_ION_DRIFT_APPLIED | EKPNP_FLAG_G0_FLUID_U;
  p.uw = 1e-3;
  std::vector<double> z(p.nx), cpw(p.nx), cnw(p.nx);
  for (int x = 0; x < p.nx; ++x) {
    z[x] = -1.5 + 0.3 * std::sin(6.283185307179586 * x / p.nx);
    cpw[x] = std::exp(-z[x]);
    cnw[x] = std::exp(z[x]);
  }
  Oracle* o = ekpnp_oracle_create(&p);
  ekpnp_oracle_set_wall_profile(o, z.data(), z.data(), cpw.data(), cpw.data(), cnw.data(), cnw.data());
  ekpnp_oracle_initialization(o, 0, 3, 1, 1);
  ekpnp_oracle_init_equilibrium(o);
  for (int s = 0; s < 3; ++s) {
    ekpnp_oracle_stage_collide(o);
    ekpnp_oracle_stage_boundary(o);
    ekpnp_oracle_stage_stream(o);
    ekpnp_oracle_stage_bc_charge(o);
    ekpnp_oracle_fast_poisson(o, s & 1);
    ekpnp_oracle_efield(o);
  }
  const size_t n = (size_t)p.nx * p.ny;
  std::vector<double> pops(9 * n), post(8 * n), fld(n);
  for (int sp = 0; sp < 3; ++sp) {
    ekpnp_oracle_get_populations(o, sp, pops.data());
    ekpnp_oracle_set_populations(o, sp, pops.data());
    ekpnp_oracle_get_post(o, sp, post.data());
    ekpnp_oracle_set_post(o, sp, post.data());
    bad += !finite(pops) + !finite(post);
  }
  ekpnp_oracle_get_field(o, EKPNP_FIELD_PSI, fld.data());
  ekpnp_oracle_set_field(o, EKPNP_FIELD_PSI, fld.data());
  ekpnp_oracle_step(o, 10, 0);
  for (int k = 0; k < EKPNP_FIELD_COUNT; ++k) {
    ekpnp_oracle_get_field(o, k, fld.data());
    bad += !finite(fld);
  }
  ekpnp_oracle_destroy(o);
  if (ekpnp_oracle_create(nullptr) != nullptr) ++bad;
  if (bad) {
    std::fprintf(stderr, "oracle sanitize run: %d non-finite results\n", bad);
    return 1;
  }
  std::printf("oracle sanitize run ok\n");
  return 0;
}
