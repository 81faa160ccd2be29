// qt_philox.h -- Philox4x64-10 counter-based generator (Salmon, Moraes, Dror,
// Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC'11), usable from host
// and device code.  Every MCWF step of trajectory `traj` consumes exactly one
// uniform draw, keyed (seed, traj) with counter (step ordinal, 0, 0, 0), so a
// trajectory's jump record does not depend on how trajectories are batched or
// scheduled.
#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define QT_HD __host__ __device__ inline
#else
#define QT_HD inline
#endif

namespace qt {

QT_HD void qt_mulhilo64(uint64_t a, uint64_t b, uint64_t* hi, uint64_t* lo) {
#if defined(__HIP_DEVICE_COMPILE__)
  *lo = a * b;
  *hi = __umul64hi(a, b);
#else
  unsigned __int128 p = (unsigned __int128)a * b;
  *hi = (uint64_t)(p >> 64);
  *lo = (uint64_t)p;
#endif
}

QT_HD void philox4x64_10(uint64_t c[4], uint64_t k0, uint64_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) { k0 += 0x9E3779B97F4A7C15ULL; k1 += 0xBB67AE8584CAA73BULL; }
    uint64_t hi0, lo0, hi1, lo1;
    qt_mulhilo64(0xD2E7470EE14C6C93ULL, c[0], &hi0, &lo0);
    qt_mulhilo64(0xCA5A826395121157ULL, c[2], &hi1, &lo1);
    uint64_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
  }
}

// uniform double in [0,1) with 53 random bits: (word0 >> 11) * 2^-53
QT_HD double jump_uniform(uint64_t seed, uint64_t traj, uint64_t step) {
  uint64_t c[4] = {step, 0, 0, 0};
  philox4x64_10(c, seed, traj);
  return (double)(c[0] >> 11) * (1.0 / 9007199254740992.0);
}

}  // namespace qt
