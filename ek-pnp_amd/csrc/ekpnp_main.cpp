// ekpnp_main.cpp -- the reference's main.cu time loop (main.cu:28-307) on the
// C ABI: create -> initialization -> init_equilibrium -> N x (stream_collide_save,
// fast_Poisson) with periodic diagnostics -> performance print -> data_end.dat.
//
// usage: ekpnp_run [--config ref|c0|c1] [--steps N] [--fp32|--fp64]
//                  [--print-every K] [--out data_end.dat]
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/ekpnp.h"

namespace {

#define CHECK(call)                                                                  \
  do {                                                                               \
    int rc_ = (call);                                                                \
    if (rc_ != EKPNP_OK) {                                                           \
      fprintf(stderr, "%s failed (%d): %s\n", #call, rc_, ekpnp_last_error(ctx));    \
      return 1;                                                                      \
    }                                                                                \
  } while (0)

// wall current I = K dy sum_x c(x, NY-1) Ey(x, NY-1), with c at the two wall
// rows linearly extrapolated from the interior first (LBM.cu:1394-1417)
double wall_current(const ekpnp_channel_params& p, std::vector<double> c, const std::vector<double>& ey) {
  const int nx = p.nx, ny = p.ny;
  for (int x = 0; x < nx; ++x) {
    c[x] = 2.0 * c[(size_t)nx + x] - c[2 * (size_t)nx + x];
    c[(size_t)(ny - 1) * nx + x] = 2.0 * c[(size_t)(ny - 2) * nx + x] - c[(size_t)(ny - 3) * nx + x];
  }
  double I = 0.0;
  for (int x = 0; x < nx; ++x) I += c[(size_t)(ny - 1) * nx + x] * ey[(size_t)(ny - 1) * nx + x];
  return I * p.mob_pos * p.dy;
}

// max u_y after extrapolating the top row (LBM.cu:1419-1451)
double max_uy(const ekpnp_channel_params& p, std::vector<double> uy) {
  const int nx = p.nx, ny = p.ny;
  for (int x = 0; x < nx; ++x)
    uy[(size_t)(ny - 1) * nx + x] = 2.0 * uy[(size_t)(ny - 2) * nx + x] - uy[(size_t)(ny - 3) * nx + x];
  double m = 0.0;
  for (double v : uy) m = v > m ? v : m;
  return m;
}

}  // namespace

int main(int argc, char** argv) {
  std::string config = "ref", out = "data_end.dat";
  long nsteps = -1, print_every = 2000;
  int precision = -1;
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    if (a == "--config" && i + 1 < argc) config = argv[++i];
    else if (a == "--steps" && i + 1 < argc) nsteps = atol(argv[++i]);
    else if (a == "--fp32") precision = EKPNP_POP_FP32;
    else if (a == "--fp64") precision = EKPNP_POP_FP64;
    else if (a == "--print-every" && i + 1 < argc) print_every = atol(argv[++i]);
    else if (a == "--out" && i + 1 < argc) out = argv[++i];
    else {
      fprintf(stderr, "usage: %s [--config ref|c0|c1] [--steps N] [--fp32|--fp64] [--print-every K] [--out FILE]\n", argv[0]);
      return 2;
    }
  }
  ekpnp_channel_params p;
  int mode = EKPNP_INIT_COLD_NOISE;
  uint64_t seed = 0;
  if (config == "ref") {
    ekpnp_reference_params(&p);
    mode = EKPNP_INIT_REFERENCE;
    if (nsteps < 0) nsteps = 100000;  // LBM.h:115
    if (precision < 0) precision = EKPNP_POP_FP64;
  } else if (config == "c0") {
    ekpnp_lattice_params(&p, 64, 128, 0.8, -1.0, 8.0, 1e-4, 0.05);
    if (nsteps < 0) nsteps = 20000;
    if (precision < 0) precision = EKPNP_POP_FP64;
  } else if (config == "c1") {
    ekpnp_lattice_params(&p, 4096, 1024, 0.6, -2.0, 32.0, 5e-5, 0.05);
    seed = 1;
    if (nsteps < 0) nsteps = 50000;
    if (precision < 0) precision = EKPNP_POP_FP32;
  } else {
    fprintf(stderr, "unknown config %s\n", config.c_str());
    return 2;
  }
  printf("Simulating EK-PNP in 2D\n      domain size: %dx%d\n        timesteps: %ld\n\n", p.nx, p.ny, nsteps);

  ekpnp_options o;
  ekpnp_default_options(&o);
  o.precision = precision;
  ekpnp_ctx* ctx = nullptr;
  int rc = ekpnp_create(&p, 1, &o, &ctx);
  if (rc != EKPNP_OK) {
    fprintf(stderr, "ekpnp_create failed (%d): %s\n", rc, ekpnp_create_error());
    return 1;
  }
  CHECK(ekpnp_initialization(ctx, mode, seed));
  CHECK(ekpnp_init_equilibrium(ctx));

  const size_t nn = (size_t)p.nx * p.ny;
  std::vector<double> c(nn), ey(nn), uy(nn);
  auto t0 = std::chrono::steady_clock::now();
  // the reference runs NSTEPS+1 iterations and counts NSTEPS (main.cu:190, 248)
  long done = 0;
  while (done <= nsteps) {
    long chunk = print_every - (done % print_every);
    if (done + chunk > nsteps + 1) chunk = nsteps + 1 - done;
    CHECK(ekpnp_step(ctx, (int)chunk));
    done += chunk;
    if (done % print_every == 1 || done == nsteps + 1) {
      CHECK(ekpnp_get_field(ctx, 0, EKPNP_FIELD_CPOS, c.data()));
      CHECK(ekpnp_get_field(ctx, 0, EKPNP_FIELD_EY, ey.data()));
      CHECK(ekpnp_get_field(ctx, 0, EKPNP_FIELD_UY, uy.data()));
      printf("Iteration: %ld, Current = %g, max uy = %g\n", done - 1, wall_current(p, c, ey), max_uy(p, uy));
    }
  }
  CHECK(ekpnp_synchronize(ctx));
  auto t1 = std::chrono::steady_clock::now();
  const double runtime = std::chrono::duration<double>(t1 - t0).count();
  const double mlups = (double)nsteps * (double)nn / (1e6 * runtime);  // main.cu:248-250
  const double bytes = (precision == EKPNP_POP_FP32 ? 248.0 : 464.0) * (double)nsteps * (double)nn;
  printf(" ----- performance information -----\n");
  printf("  memory allocated (GPU): %.1f (MiB)\n", ekpnp_device_bytes(ctx) / 1048576.0);
  printf("               timesteps: %ld\n", nsteps);
  printf("           clock runtime: %.3f (s)\n", runtime);
  printf("                   speed: %.2f (Mlups)\n", mlups);
  printf("  algorithmic bandwidth: %.1f (GB/s)\n", bytes / runtime / 1e9);

  // data_end.dat: t ux uy rho c phi Ex Ey T per node, x fastest (LBM.cu:1294-1302),
  // written with 17 significant digits so that it restarts exactly
  std::vector<std::vector<double>> f(8, std::vector<double>(nn));
  const int order[8] = {EKPNP_FIELD_UX, EKPNP_FIELD_UY, EKPNP_FIELD_RHO, EKPNP_FIELD_CPOS,
                        EKPNP_FIELD_PSI, EKPNP_FIELD_EX, EKPNP_FIELD_EY, EKPNP_FIELD_CNEG};
  for (int k = 0; k < 8; ++k) CHECK(ekpnp_get_field(ctx, 0, order[k], f[k].data()));
  FILE* fend = fopen(out.c_str(), "w");
  if (!fend) {
    fprintf(stderr, "cannot open %s\n", out.c_str());
    ekpnp_destroy(ctx);
    return 1;
  }
  const double t = p.dt * (double)(nsteps + 1);
  for (size_t i = 0; i < nn; ++i) {
    fprintf(fend, "%.17g", t);
    for (int k = 0; k < 8; ++k) fprintf(fend, " %.17g", f[k][i]);
    fputc('\n', fend);
  }
  fclose(fend);
  ekpnp_destroy(ctx);
  return 0;
}
