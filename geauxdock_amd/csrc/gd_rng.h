// gd_rng.h -- counter-based random numbers shared by the host and gfx950 code.
//
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
// as 1, 2, 3", SC'11): 10 rounds of two 32x32->64 multiplies with a Weyl key
// schedule.  Stateless: a draw is a pure function of (key, counter), which is
// what lets every workgroup re-derive a replica's move without shared state and
// lets a job replay bit-identically alone or inside a batch.  It replaces the
// reference's XORWOW states (backend_gpu/kernel_cuda_l1_initcurand.cu:4-10) and
// the time-seeded rand_r of the CPU backend (kernel_cpp_montecarlo_soa.C:8-15).
//
// Stream layout used throughout the library:
//   key     = (seed low 32 bits, seed high 32 bits)
//   counter = (step, replica, job_id, slot)
//   slot 0  -> move draws 0..3, slot 1 -> move draws 4..5, accept draw, spare
//   slot 2  -> temperature-exchange draw (counter word 0 = round, word 1 = pair)
//   slot 3  -> ligand-conformer swap draw, slot 4 -> protein-conformer swap draw
//              (counter word 0 = round, word 1 = lower-conformer replica)
#pragma once
#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define GD_HD __host__ __device__ __forceinline__
#else
#define GD_HD static inline
#endif

struct gd_u32x4 {
  uint32_t v[4];
};

GD_HD void gd_mulhilo32(uint32_t a, uint32_t b, uint32_t &hi, uint32_t &lo) {
  const uint64_t p = (uint64_t)a * (uint64_t)b;
  hi = (uint32_t)(p >> 32);
  lo = (uint32_t)p;
}

// Philox4x32 with R rounds; R = 10 is the standard generator.
GD_HD gd_u32x4 gd_philox4x32(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                             uint32_t k1, int rounds = 10) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
  for (int r = 0; r < rounds; ++r) {
    uint32_t hi0, lo0, hi1, lo1;
    gd_mulhilo32(M0, c0, hi0, lo0);
    gd_mulhilo32(M1, c2, hi1, lo1);
    const uint32_t n0 = hi1 ^ c1 ^ k0;
    const uint32_t n2 = hi0 ^ c3 ^ k1;
    c0 = n0;
    c1 = lo1;
    c2 = n2;
    c3 = lo0;
    k0 += W0;
    k1 += W1;
  }
  gd_u32x4 out;
  out.v[0] = c0;
  out.v[1] = c1;
  out.v[2] = c2;
  out.v[3] = c3;
  return out;
}

// 24-bit uniform in [0, 1): exact in FP32, identical on host and device.
GD_HD float gd_u01(uint32_t x) { return (float)(x >> 8) * (1.0f / 16777216.0f); }

// The eight draws of one replica-step: w[0..5] moves, w[6] accept, w[7] spare.
GD_HD void gd_step_draws(uint64_t seed, uint32_t job, uint32_t replica, uint32_t step,
                         uint32_t w[8]) {
  const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
  const gd_u32x4 a = gd_philox4x32(step, replica, job, 0u, k0, k1);
  const gd_u32x4 b = gd_philox4x32(step, replica, job, 1u, k0, k1);
  w[0] = a.v[0];
  w[1] = a.v[1];
  w[2] = a.v[2];
  w[3] = a.v[3];
  w[4] = b.v[0];
  w[5] = b.v[1];
  w[6] = b.v[2];
  w[7] = b.v[3];
}

// Accept draw of a replica-step alone (word 6 of gd_step_draws: slot 1 only).
GD_HD float gd_accept_u01(uint64_t seed, uint32_t job, uint32_t replica, uint32_t step) {
  const gd_u32x4 b = gd_philox4x32(step, replica, job, 1u, (uint32_t)seed, (uint32_t)(seed >> 32));
  return gd_u01(b.v[2]);
}

// Exchange draw for pair `pair` in exchange round `round`.
GD_HD float gd_exchange_u01(uint64_t seed, uint32_t job, uint32_t pair, uint32_t round) {
  const gd_u32x4 a = gd_philox4x32(round, pair, job, 2u, (uint32_t)seed, (uint32_t)(seed >> 32));
  return gd_u01(a.v[0]);
}

// Draw of a conformer pose swap in exchange round `round` along `axis`
// (0: ligand conformer, 1: protein conformer); the pair is named by its member
// on the lower conformer index.
GD_HD float gd_swap_u01(uint64_t seed, uint32_t job, uint32_t replica, uint32_t round,
                        uint32_t axis) {
  const gd_u32x4 a =
      gd_philox4x32(round, replica, job, 3u + axis, (uint32_t)seed, (uint32_t)(seed >> 32));
  return gd_u01(a.v[0]);
}

// Move increments of one replica-step from its draws w[0..5] (the reference
// move is scale * U[0,1), kernel_cpp_montecarlo_soa.C:242-250; mode 1 draws
// scale * N(0,1) instead).  Gaussian pairs by Box-Muller: with u1, u2 uniform
// on [0,1), r = sqrt(-2 ln(1 - u1)) (1 - u1 is in (0, 1], so r is finite) and
// the pair (r cos 2 pi u2, r sin 2 pi u2) is two independent N(0,1) draws.
// Compiled for the host (libm) and for the device (the device math library);
// the two agree to a few ulps, not bit for bit.
GD_HD void gd_move_incr(int mode, const float scale[6], const uint32_t w[8], float inc[6]) {
  if (mode == 1) {
    for (int k = 0; k < 3; ++k) {
      const float u1 = gd_u01(w[2 * k]), u2 = gd_u01(w[2 * k + 1]);
      const float rad = sqrtf(-2.0f * logf(1.0f - u1));
      float sn, cs;
#if defined(__HIP_DEVICE_COMPILE__)
      sincosf(6.28318530717958647692f * u2, &sn, &cs);
#else
      sn = sinf(6.28318530717958647692f * u2);
      cs = cosf(6.28318530717958647692f * u2);
#endif
      inc[2 * k] = scale[2 * k] * (rad * cs);
      inc[2 * k + 1] = scale[2 * k + 1] * (rad * sn);
    }
  } else {
    for (int i = 0; i < 6; ++i) inc[i] = scale[i] * gd_u01(w[i]);
  }
}
