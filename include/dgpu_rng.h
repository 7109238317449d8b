/*
 * dgpu_rng.h -- counter-based random numbers for the path tracer.
 *
 * Philox4x32-10 (Salmon, Moraes, Dror, Shaw: "Parallel random numbers: as easy
 * as 1, 2, 3", SC'11).  A path's random numbers are a pure function of
 * (seed, scene, pixel, sample, bounce, dimension), so a path can be replayed
 * exactly (adjoint rendering) and results do not depend on how paths are
 * scheduled onto threads.  dort itself draws from std::mt19937 in thread-order
 * dependent streams (src/dort/rng.hpp:6-28, renderer.cpp:55-75), so only the
 * distribution of samples, not the stream, is shared with dort.
 *
 * Dimension layout per (pixel, sample, bounce): three Philox blocks of 4 words.
 *   block 0: [0] light pick, [1..3] BSDF-strategy BsdfSample (u, v, component)
 *   block 1: [0..1] LightSample (u, v), [2..3] continuation BsdfSample (u, v)
 *   block 2: [0] continuation component, [1] Russian roulette,
 *            [2..3] film jitter (used at bounce 0 only)
 * path_renderer.cpp:17,73,89,104,125,156,191 draw these quantities.
 *
 * Header-only, usable from host C/C++ and HIP device code.
 */
#ifndef DGPU_RNG_H
#define DGPU_RNG_H

#include <stdint.h>

#if defined(__HIPCC__)
#define DGPU_HD __host__ __device__ __forceinline__
#else
#define DGPU_HD static inline
#endif

typedef struct { uint32_t v[4]; } dgpu_u32x4;

DGPU_HD uint32_t dgpu_mulhilo32(uint32_t a, uint32_t b, uint32_t* hi) {
  uint64_t p = (uint64_t)a * (uint64_t)b;
  *hi = (uint32_t)(p >> 32);
  return (uint32_t)p;
}

/* Philox4x32 with 10 rounds; constants from the Random123 specification. */
DGPU_HD dgpu_u32x4 dgpu_philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                      uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
  for (int r = 0; r < 10; ++r) {
    uint32_t hi0, hi1;
    uint32_t lo0 = dgpu_mulhilo32(M0, c0, &hi0);
    uint32_t lo1 = dgpu_mulhilo32(M1, c2, &hi1);
    uint32_t n0 = hi1 ^ c1 ^ k0;
    uint32_t n1 = lo1;
    uint32_t n2 = hi0 ^ c3 ^ k1;
    uint32_t n3 = lo0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += W0; k1 += W1;
  }
  dgpu_u32x4 out;
  out.v[0] = c0; out.v[1] = c1; out.v[2] = c2; out.v[3] = c3;
  return out;
}

/* Uniform float in [0, 1): the top 24 bits scaled by 2^-24 (exactly representable). */
DGPU_HD float dgpu_u01(uint32_t x) { return (float)(x >> 8) * (1.0f / 16777216.0f); }

/* The key of a render: seed and scene id. */
DGPU_HD void dgpu_rng_key(uint64_t seed, uint32_t scene_id, uint32_t* k0, uint32_t* k1) {
  *k0 = (uint32_t)seed;
  *k1 = (uint32_t)(seed >> 32) ^ (scene_id * 0x85EBCA6Bu);
}

/* Four uniforms of block `blk` (0..2) of (pixel, sample, bounce). */
DGPU_HD void dgpu_rng_block(uint32_t k0, uint32_t k1, uint32_t pixel, uint32_t sample,
                            uint32_t bounce, uint32_t blk, float out[4]) {
  dgpu_u32x4 r = dgpu_philox4x32_10(pixel, sample, (bounce << 2) | blk, 0x64677075u, k0, k1);
  out[0] = dgpu_u01(r.v[0]);
  out[1] = dgpu_u01(r.v[1]);
  out[2] = dgpu_u01(r.v[2]);
  out[3] = dgpu_u01(r.v[3]);
}

enum {
  DGPU_DIM_LIGHT_PICK = 0,
  DGPU_DIM_NEE_BSDF_U = 1, DGPU_DIM_NEE_BSDF_V = 2, DGPU_DIM_NEE_BSDF_C = 3,
  DGPU_DIM_LIGHT_U = 4, DGPU_DIM_LIGHT_V = 5,
  DGPU_DIM_BSDF_U = 6, DGPU_DIM_BSDF_V = 7,
  DGPU_DIM_BSDF_C = 8, DGPU_DIM_RR = 9,
  DGPU_DIM_FILM_X = 10, DGPU_DIM_FILM_Y = 11,
  DGPU_DIMS_PER_BOUNCE = 12
};

#endif /* DGPU_RNG_H */
