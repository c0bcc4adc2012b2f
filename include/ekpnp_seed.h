/*
 * ekpnp_seed.h -- seeded synthetic inputs shared by every consumer of the
 * C ABI (header-only, plain C).
 *
 * The cold-start concentration of species s (0 = positive, 1 = negative) of
 * channel ch at interior node (x, y) is
 *     c_bulk * (1 + noise * (2u - 1)),  u in [0, 1),
 * where u is drawn from a splitmix64 stream keyed by (seed, ch, s) and
 * consumed in row-major order over all nodes (wall nodes draw too, then take
 * their wall value).  The reference seeds nothing: its start is the uniform
 * gpu_initialization state (LBM.cu:147-160).
 */
#ifndef EKPNP_SEED_H
#define EKPNP_SEED_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

static inline uint64_t ekpnp_splitmix64_next(uint64_t* state) {
  uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

/* state of the stream for (seed, channel, species) */
static inline uint64_t ekpnp_stream_key(uint64_t seed, uint64_t channel,
                                        uint64_t species) {
  uint64_t s = seed ^ 0x243F6A8885A308D3ull;
  uint64_t k = ekpnp_splitmix64_next(&s);
  s = k ^ (channel * 0xD1B54A32D192ED03ull);
  k = ekpnp_splitmix64_next(&s);
  s = k ^ ((species + 1) * 0x8CB92BA72F3D8DD7ull);
  return ekpnp_splitmix64_next(&s);
}

/* uniform double in [0, 1) with 53 random bits */
static inline double ekpnp_uniform01(uint64_t* state) {
  return (double)(ekpnp_splitmix64_next(state) >> 11) * (1.0 / 9007199254740992.0);
}

/*
 * Fill one species' cold-start field, [ny][nx] row-major.  wall_lo/wall_hi
 * are the per-x wall concentrations (length nx).
 */
static inline void ekpnp_cold_field(double* out, int nx, int ny, double c_bulk,
                                    double noise, const double* wall_lo,
                                    const double* wall_hi, uint64_t seed,
                                    int channel, int species) {
  uint64_t st = ekpnp_stream_key(seed, (uint64_t)channel, (uint64_t)species);
  for (int y = 0; y < ny; ++y) {
    for (int x = 0; x < nx; ++x) {
      double u = ekpnp_uniform01(&st);
      double v = c_bulk * (1.0 + noise * (2.0 * u - 1.0));
      if (y == 0) v = wall_lo[x];
      if (y == ny - 1) v = wall_hi[x];
      out[(long)y * nx + x] = v;
    }
  }
}

#ifdef __cplusplus
}
#endif

#endif /* EKPNP_SEED_H */
