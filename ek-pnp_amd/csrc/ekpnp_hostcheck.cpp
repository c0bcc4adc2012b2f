// ekpnp_hostcheck.cpp -- exercises every device-free path of the C ABI
// (parameter helpers, the resolution check, option defaults, argument and
// parameter validation, null contexts) for the AddressSanitizer /
// UndefinedBehaviorSanitizer build of the host code (make -C ek-pnp_amd
// sanitize).  Needs no GPU.  Exits 0 when every result is as documented in
// include/ekpnp.h; a sanitizer finding aborts it.
#include <cstdio>
#include <cstring>

#include "../../include/ekpnp.h"

static int failures = 0;
#define EXPECT(cond)                                                  \
  do {                                                                \
    if (!(cond)) {                                                    \
      std::fprintf(stderr, "hostcheck: %s:%d: %s\n", __FILE__, __LINE__, #cond); \
      ++failures;                                                     \
    }                                                                 \
  } while (0)

int main() {
  ekpnp_channel_params p[4];
  ekpnp_reference_params(&p[0]);
  EXPECT(p[0].nx == 10 && p[0].ny == 101);
  EXPECT(ekpnp_lattice_params(&p[1], 64, 128, 0.8, -1.0, 8.0, 1e-4, 0.05) == EKPNP_OK);
  EXPECT(ekpnp_lattice_params(&p[2], 64, 40, 0.8, -0.5, 9.0, 1e-4, 0.05) == EKPNP_OK);   // kappa dy 0.46
  EXPECT(ekpnp_lattice_params(&p[3], 64, 200, 0.8, -4.0, 30.0, 1e-4, 0.05) == EKPNP_OK); // fast wall drift
  EXPECT(ekpnp_lattice_params(&p[3], 64, 3, 0.8, -1.0, 8.0, 1e-4, 0.05) == EKPNP_ERR_INVALID);
  EXPECT(ekpnp_lattice_params(&p[3], 64, 64, 0.5, -1.0, 8.0, 1e-4, 0.05) == EKPNP_ERR_INVALID);
  EXPECT(ekpnp_lattice_params(nullptr, 64, 64, 0.8, -1.0, 8.0, 1e-4, 0.05) == EKPNP_ERR_INVALID);
  EXPECT(ekpnp_lattice_params(&p[3], 64, 200, 0.8, -4.0, 30.0, 1e-4, 0.05) == EKPNP_OK);

  uint32_t flags[4] = {9, 9, 9, 9};
  EXPECT(ekpnp_check_params(p, 4, flags) == 2);
  EXPECT(flags[0] == 0 && flags[1] == 0);
  EXPECT(flags[2] == EKPNP_CHECK_DEBYE_UNRESOLVED && flags[3] == EKPNP_CHECK_WALL_DRIFT);
  EXPECT(ekpnp_check_params(p, 0, flags) == EKPNP_ERR_INVALID);
  EXPECT(ekpnp_check_params(nullptr, 1, flags) == EKPNP_ERR_INVALID);
  EXPECT(ekpnp_check_params(p, 4, nullptr) == 2);
  // tau 0.51 under a strong field: electro-osmotic flow |zeta| E / nu = 6
  ekpnp_channel_params q;
  EXPECT(ekpnp_lattice_params(&q, 64, 128, 0.51, -2.0, 8.0, 1e-2, 0.05) == EKPNP_OK);
  uint32_t fq = 0;
  EXPECT(ekpnp_check_params(&q, 1, &fq) == 1 && fq == EKPNP_CHECK_FLOW_FAST);

  ekpnp_options o;
  std::memset(&o, 0x5a, sizeof o);
  ekpnp_default_options(&o);
  EXPECT(o.precision == EKPNP_POP_FP32 && o.use_graph == 1 && o.compute_fp64 == 1 && o.reject_unresolved == 0);
  ekpnp_default_options(nullptr);

  // creation is refused before any device call
  ekpnp_ctx* ctx = reinterpret_cast<ekpnp_ctx*>(0x1);
  EXPECT(ekpnp_create(p, 0, &o, &ctx) == EKPNP_ERR_INVALID && ctx == nullptr);
  EXPECT(ekpnp_create(p, 1, &o, nullptr) == EKPNP_ERR_INVALID);
  ekpnp_channel_params bad = p[1];
  bad.nu = -1.0;
  EXPECT(ekpnp_create(&bad, 1, &o, &ctx) == EKPNP_ERR_INVALID);
  EXPECT(std::strstr(ekpnp_create_error(), "viscosity") != nullptr);
  bad = p[1];
  bad.nx = 16384;
  EXPECT(ekpnp_create(&bad, 1, &o, &ctx) == EKPNP_ERR_UNSUPPORTED);
  bad = p[1];
  bad.ny = 3000;
  EXPECT(ekpnp_create(&bad, 1, &o, &ctx) == EKPNP_ERR_UNSUPPORTED);
  bad = p[1];
  bad.nx = 3000;  // not a power of two and above the direct-DFT limit
  EXPECT(ekpnp_create(&bad, 1, &o, &ctx) == EKPNP_ERR_UNSUPPORTED);
  o.reject_unresolved = 1;
  EXPECT(ekpnp_create(p, 4, &o, &ctx) == EKPNP_ERR_UNSTABLE);
  EXPECT(std::strstr(ekpnp_create_error(), "channel 2") != nullptr);
  o.reject_unresolved = 0;
  o.precision = 7;
  EXPECT(ekpnp_create(p, 2, &o, &ctx) == EKPNP_ERR_INVALID);

  // every entry point taking a context refuses a null one
  double d[8] = {0};
  int it = 0;
  EXPECT(std::strcmp(ekpnp_last_error(nullptr), "null context") == 0);
  ekpnp_destroy(nullptr);
  EXPECT(ekpnp_initialization(nullptr, 0, 0) == EKPNP_ERR_INVALID);
  EXPECT(ekpnp_init_equilibrium(nullptr) == EKPNP_ERR_INVALID);
  EXPECT(ekpnp_stream_collide_save(nullptr) == EKPNP_ERR_INVALID);
  EXPECT(ekpnp_fast_poisson(nullptr) == EKPNP_ERR_INVALID);
  EXPECT(ekpnp_step(nullptr, 1) == EKPNP_ERR_INVALID);
  EXPECT(ekpnp_rebase(nullptr) == EKPNP_ERR_INVALID);
  EXPECT(ekpnp_run_to_residual(nullptr, 1e-8, 100, 10, nullptr) == EKPNP_ERR_INVALID);
  EXPECT(ekpnp_get_field(nullptr, 0, 0, d) == EKPNP_ERR_INVALID);
  EXPECT(ekpnp_get_populations(nullptr, 0, 0, d) == EKPNP_ERR_INVALID);
  EXPECT(ekpnp_reductions(nullptr, d) == EKPNP_ERR_INVALID);
  EXPECT(ekpnp_poisson_iterative(nullptr, 0, 1.0, 1e-9, 10, 1, &it, d) == EKPNP_ERR_INVALID);
  EXPECT(ekpnp_save_checkpoint(nullptr, "x") == EKPNP_ERR_INVALID);
  EXPECT(ekpnp_load_checkpoint(nullptr, "x") == EKPNP_ERR_INVALID);
  EXPECT(ekpnp_autotune(nullptr, 1, d) == EKPNP_ERR_INVALID);
  EXPECT(ekpnp_set_poisson_plan(nullptr, 2048, 16, 0) == EKPNP_ERR_INVALID);
  EXPECT(ekpnp_time_stages(nullptr, 1, d) == EKPNP_ERR_INVALID);
  EXPECT(ekpnp_num_channels(nullptr) == 0 && ekpnp_steps_done(nullptr) == 0 && ekpnp_device_bytes(nullptr) == 0);
  EXPECT(ekpnp_copy_bandwidth(0, 100, 1, 1, d, nullptr) == EKPNP_ERR_INVALID);  // not a multiple of 16

  if (failures) {
    std::fprintf(stderr, "hostcheck: %d failure(s)\n", failures);
    return 1;
  }
  std::printf("hostcheck ok\n");
  return 0;
}
