// ekpnp_main.cpp -- the reference's main.cu time loop (main.cu:28-307) on the
// C ABI: create -> initialization (or read_data restart) -> init_equilibrium
// -> N x (stream_collide_save, fast_Poisson) with the reference's periodic
// outputs -> performance print -> data_end.dat.
//
// usage: ekpnp_run [--config ref|c0|c1] [--steps N] [--fp32|--fp64]
//                  [--nsave K] [--ndmd K] [--print-current K]
//                  [--restart data_end.dat] [--checkpoint FILE] [--resume FILE]
//                  [--dir OUTDIR]
// Outputs (in OUTDIR, default "."): data.dat (Tecplot, every nsave steps),
// charge_data (every ndmd), umax.dat and the printed wall current (every
// print-current), data_end.dat at the end (LBM.h:115-119 intervals by
// default).
#include <algorithm>
#include <chrono>
#include <climits>
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

// first iteration index j >= i with j % k == 1 (the reference's output test)
long next_event(long i, long k) {
  if (k < 2) return LONG_MAX;  // never fires
  long r = i % k;
  return r <= 1 ? i + (1 - r) : i + (k - r) + 1;
}

}  // namespace

int main(int argc, char** argv) {
  std::string config = "ref", dir = ".", restart, ckpt_out, resume;
  long nsteps = -1, nsave = -1, ndmd = -1, print_current = -1;
  int precision = -1;
  for (int i = 1; i < argc; ++i) {
    std::string a = argv[i];
    auto next = [&]() { return std::string(i + 1 < argc ? argv[++i] : ""); };
    if (a == "--config") config = next();
    else if (a == "--steps") nsteps = atol(next().c_str());
    else if (a == "--fp32") precision = EKPNP_POP_FP32;
    else if (a == "--fp64") precision = EKPNP_POP_FP64;
    else if (a == "--nsave") nsave = atol(next().c_str());
    else if (a == "--ndmd") ndmd = atol(next().c_str());
    else if (a == "--print-current") print_current = atol(next().c_str());
    else if (a == "--restart") restart = next();
    else if (a == "--checkpoint") ckpt_out = next();
    else if (a == "--resume") resume = next();
    else if (a == "--dir") dir = next();
    else {
      fprintf(stderr,
              "usage: %s [--config ref|c0|c1] [--steps N] [--fp32|--fp64] [--nsave K] [--ndmd K] "
              "[--print-current K] [--restart data_end.dat] [--checkpoint FILE] [--resume FILE] [--dir D]\n",
              argv[0]);
      return 2;
    }
  }
  ekpnp_channel_params p;
  int mode = EKPNP_INIT_COLD_NOISE;
  uint64_t seed = 0;
  if (config == "ref") {
    ekpnp_reference_params(&p);
    mode = EKPNP_INIT_PB_PICARD;  // initialization(), LBM.cu:95-145
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
  // LBM.h:116-119; NDMD's literal 5000000000000 wraps to 658067456 in the
  // reference's unsigned int, so charge_data is written once, after step 1
  if (nsave < 0) nsave = nsteps / 5;
  if (ndmd < 0) ndmd = 658067456L;
  if (print_current < 0) print_current = 2000;
  const std::string f_tec = dir + "/data.dat", f_umax = dir + "/umax.dat", f_dmd = dir + "/charge_data",
                    f_end = dir + "/data_end.dat";
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
  double t = 0.0;
  if (!resume.empty()) {
    CHECK(ekpnp_load_checkpoint(ctx, resume.c_str()));
    t = p.dt * (double)ekpnp_steps_done(ctx);
  } else {
    CHECK(ekpnp_initialization(ctx, mode, seed));
    if (!restart.empty()) CHECK(ekpnp_read_data(ctx, 0, restart.c_str(), &t));  // main.cu:165-167
    CHECK(ekpnp_init_equilibrium(ctx));                                          // main.cu:174
  }
  CHECK(ekpnp_save_tecplot(ctx, 0, t, f_tec.c_str(), 1));  // main.cu:178
  std::remove(f_umax.c_str());

  const size_t nn = (size_t)p.nx * p.ny;
  auto t0 = std::chrono::steady_clock::now();
  // iterations i = 0..NSTEPS (main.cu:190); after iteration i the reference
  // writes when i % K == 1, then advances t
  long i = 0;
  while (i <= nsteps) {
    long stop = nsteps;
    for (long k : {nsave, ndmd, print_current}) stop = std::min(stop, next_event(i, k));
    CHECK(ekpnp_step(ctx, (int)(stop - i + 1)));
    t += p.dt * (double)(stop - i);  // t of iteration `stop`
    i = stop;
    if (nsave > 0 && i % nsave == 1) CHECK(ekpnp_save_tecplot(ctx, 0, t, f_tec.c_str(), 0));
    if (ndmd > 0 && i % ndmd == 1) {
      printf("Iteration: %ld, physical time: %g.\n", i, t);
      CHECK(ekpnp_save_charge_data(ctx, 0, f_dmd.c_str()));
    }
    if (print_current > 0 && i % print_current == 1) {
      double I = 0.0, umax = 0.0;
      CHECK(ekpnp_current(ctx, 0, &I));
      printf("Iteration: %ld, physical time: %g, Current = %g\n", i, t, I);
      CHECK(ekpnp_record_umax(ctx, 0, t, f_umax.c_str(), &umax));
    }
    t += p.dt;
    ++i;
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
  CHECK(ekpnp_save_data_end(ctx, 0, t, f_end.c_str()));
  if (!ckpt_out.empty()) CHECK(ekpnp_save_checkpoint(ctx, ckpt_out.c_str()));
  ekpnp_destroy(ctx);
  return 0;
}
