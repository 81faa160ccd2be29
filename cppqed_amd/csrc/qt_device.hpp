// qt_device.hpp -- device helpers shared by the MCWF kernels (complex FMA, DPP lane
// reductions, the Cash-Karp tableau, the step-size controller's power function).
#pragma once

#include <hip/hip_runtime.h>

namespace qt {

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------

template <bool F32> struct CT;
template <> struct CT<false> { using T = double2; using R = double; };
template <> struct CT<true> { using T = float2; using R = float; };

__device__ __forceinline__ double2 cfma(double2 a, double2 b, double2 c) {
  c.x = fma(a.x, b.x, c.x); c.x = fma(-a.y, b.y, c.x);
  c.y = fma(a.x, b.y, c.y); c.y = fma(a.y, b.x, c.y);
  return c;
}
__device__ __forceinline__ float2 cfma(float2 a, float2 b, float2 c) {
  c.x = fmaf(a.x, b.x, c.x); c.x = fmaf(-a.y, b.y, c.x);
  c.y = fmaf(a.x, b.y, c.y); c.y = fmaf(a.y, b.x, c.y);
  return c;
}
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ double2 to_d(double2 a) { return a; }
__device__ __forceinline__ double2 to_d(float2 a) { return make_double2((double)a.x, (double)a.y); }
template <typename T> __device__ __forceinline__ T from_d(double2 a);
template <> __device__ __forceinline__ double2 from_d<double2>(double2 a) { return a; }
template <> __device__ __forceinline__ float2 from_d<float2>(double2 a) { return make_float2((float)a.x, (float)a.y); }
// FP64 copy of a stage derivative made at its point of use.  For FP32 derivatives the empty asm
// hides the operand from common-subexpression elimination, so the FP32 values stay live across
// the RK stages instead of hoisted FP64 copies (half the registers).
__device__ __forceinline__ double2 to_d_use(double2 a) { return a; }
__device__ __forceinline__ double2 to_d_use(float2 a) {
  asm volatile("" : "+v"(a.x), "+v"(a.y));
  return make_double2((double)a.x, (double)a.y);
}
__device__ __forceinline__ double2 axpy(double s, double2 k, double2 acc) {  // acc + s*k
  acc.x = fma(s, k.x, acc.x); acc.y = fma(s, k.y, acc.y); return acc;
}
__device__ __forceinline__ double cabs2(double2 a) { return a.x * a.x + a.y * a.y; }

// Cross-lane moves through DPP (VALU latency, no LDS round trip).  Controls (GFX9 encoding):
// quad_perm [1,0,3,2] = 0xB1, quad_perm [2,3,0,1] = 0x4E, row_half_mirror = 0x141, row_mirror = 0x140.
// Every lane of these controls reads a lane of its own row, so no lane needs an old value
// (mov_dpp: no initialising move of the destination).
template <int CTRL>
__device__ __forceinline__ double dpp(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, true);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double readlane(double v, int l) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l), __builtin_amdgcn_readlane(__double2loint(v), l));
}

// A wavefront-uniform value read from LDS lands in vector registers; these put it back in scalar
// registers so that addresses built from it stay scalar base + vector offset.
__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }
template <typename P>
__device__ __forceinline__ P* uniform(P* p) {
  const unsigned long long v = reinterpret_cast<unsigned long long>(p);
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
  return reinterpret_cast<P*>(((unsigned long long)hi << 32) | lo);
}

// Orders the LDS accesses of one wavefront around a hand-over between its lanes.  The LDS
// executes one wave's instructions in issue order, so only the compiler has to be kept from
// moving accesses across (wavefront-scope fences emit no hardware wait).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// barrier of a team of NW wavefronts: a whole work group for NW > 1, one wavefront otherwise
template <int NW>
__device__ __forceinline__ void team_sync() {
  if constexpr (NW == 1) wave_sync();
  else __syncthreads();
}

// Row broadcasts across the 16-lane rows (GFX9 DPP): row_bcast:15 (0x142) hands lane 15 of each
// row to the next row, row_bcast:31 (0x143) hands lane 31 to rows 2 and 3.  Only the rows in
// ROWMASK are written; the others get 0.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_bcast(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROWMASK, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROWMASK, 0xF, false);
  return __hiloint2double(hi, lo);
}

// Sum / max over the 64 lanes of a wavefront; every lane gets the bitwise-identical result.
// Four butterfly steps leave every lane of row i with the row's sum r_i (each step adds two
// equal-order partial sums, addition being commutative); two row broadcasts then form
// (r3 + r2) + (r1 + r0) in lane 63, which is read out as a wavefront-uniform value (the same bits
// as (r0 + r1) + (r2 + r3), again by commutativity).  All 64 lanes must be active.
__device__ __forceinline__ double wave_sum(double v) {
  v += dpp<0xB1>(v); v += dpp<0x4E>(v); v += dpp<0x141>(v); v += dpp<0x140>(v);
  v += dpp_bcast<0x142, 0xA>(v);
  v += dpp_bcast<0x143, 0xC>(v);
  return readlane(v, 63);
}
__device__ __forceinline__ double wave_max(double v) {
  v = fmax(v, dpp<0xB1>(v)); v = fmax(v, dpp<0x4E>(v)); v = fmax(v, dpp<0x141>(v)); v = fmax(v, dpp<0x140>(v));
  // rows outside the mask receive 0 and are not read
  v = fmax(v, dpp_bcast<0x142, 0xA>(v));
  v = fmax(v, dpp_bcast<0x143, 0xC>(v));
  return readlane(v, 63);
}

// three sums over the 64 lanes at once (interleaved butterflies: one latency for all three)
__device__ __forceinline__ void wave_sum3(double v[3]) {
#pragma unroll
  for (int i = 0; i < 3; ++i) v[i] += dpp<0xB1>(v[i]);
#pragma unroll
  for (int i = 0; i < 3; ++i) v[i] += dpp<0x4E>(v[i]);
#pragma unroll
  for (int i = 0; i < 3; ++i) v[i] += dpp<0x141>(v[i]);
#pragma unroll
  for (int i = 0; i < 3; ++i) v[i] += dpp<0x140>(v[i]);
#pragma unroll
  for (int i = 0; i < 3; ++i) v[i] += dpp_bcast<0x142, 0xA>(v[i]);
#pragma unroll
  for (int i = 0; i < 3; ++i) v[i] += dpp_bcast<0x143, 0xC>(v[i]);
#pragma unroll
  for (int i = 0; i < 3; ++i) v[i] = readlane(v[i], 63);
}

template <int NW> struct GroupReduce {
  // sum / max over all lanes of the group; every lane gets the result.
  __device__ static double sum(double v, double* red) {
    v = wave_sum(v);
    if constexpr (NW == 1) { (void)red; return v; }
    else {
      const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
      __syncthreads();
      if (lane == 0) red[w] = v;
      __syncthreads();
      double s = 0.;
#pragma unroll
      for (int i = 0; i < NW; ++i) s += red[i];
      return s;
    }
  }
  __device__ static double max(double v, double* red) {
    v = wave_max(v);
    if constexpr (NW == 1) { (void)red; return v; }
    else {
      const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
      __syncthreads();
      if (lane == 0) red[w] = v;
      __syncthreads();
      double s = red[0];
#pragma unroll
      for (int i = 1; i < NW; ++i) s = fmax(s, red[i]);
      return s;
    }
  }
};

// x^y for the step-size controller, 0 < x, as exp(y log x): a few ulp from pow, which
// moves dt by a few ulp (Boost's default_step_adjuster calls std::pow).
__device__ __forceinline__ double ctrl_pow(double x, double y) { return exp(y * log(x)); }

// x^(-1/N) for N = 3, 5 (the controller's only exponents): an FP32 estimate refined by two
// Newton steps on r^N x = 1 in FP64 (quadratic convergence: 1e-7 -> 1e-14 -> rounding level),
// a few ulp from std::pow and far cheaper than FP64 log/exp.  Valid for 1e-30 < x < 1e30,
// which covers the controller's range (errors in [5^-5, 1e30]).
template <int N>
__device__ __forceinline__ double inv_root(double x) {
  double r = (double)exp2f(-__log2f((float)x) * (1.0f / N));
  #pragma unroll
  for (int it = 0; it < 2; ++it) {
    double rn = r;
    #pragma unroll
    for (int k = 1; k < N; ++k) rn *= r;        // r^N
    const double e = fma(-rn, x, 1.0);         // 1 - x r^N
    r = fma(r * (1.0 / N), e, r);              // r (1 + e/N)
  }
  return r;
}

// x^N for a compile-time N by repeated squaring
template <int N>
__device__ __forceinline__ double ipow(double r) {
  if constexpr (N == 1) return r;
  else if constexpr (N % 2 == 0) { const double h = ipow<N / 2>(r); return h * h; }
  else return r * ipow<N - 1>(r);
}

// m^(-1/N) for the controller on the SQUARED error norm m = err^2: err^(-1/3) = m^(-1/6),
// err^(-1/5) = m^(-1/10).  FP32 estimate, two Newton steps in FP64 (as inv_root).  For m beyond
// the FP32 range the estimate is 0, which the callers' clamps (dt *= max(..., 1/5)) absorb.
template <int N>
__device__ __forceinline__ double inv_root_sq(double m) {
  double r = (double)exp2f(-__log2f((float)m) * (1.0f / N));
#pragma unroll
  for (int it = 0; it < 2; ++it) {
    const double e = fma(-ipow<N>(r), m, 1.0);  // 1 - m r^N
    r = fma(r * (1.0 / N), e, r);
  }
  return r;
}

// sqrt(x), x >= 0: the hardware reciprocal square root refined by a coupled Newton step and a
// final residual correction (about an ulp from the IEEE result for 1e-300 < x < 1e300; 0 below)
__device__ __forceinline__ double fast_sqrt(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  double s = x * y, h = 0.5 * y;
  const double r = fma(-s, h, 0.5);
  s = fma(s, r, s);
  h = fma(h, r, h);
  s = fma(fma(-s, s, x), h, s);
  return x > 1e-300 ? s : 0.;
}

// 1/sqrt(x), x > 0 in the normal range: hardware estimate and two Newton steps
__device__ __forceinline__ double fast_rsqrt(double x) {
  double y = __builtin_amdgcn_rsq(x);
  const double h = 0.5 * x;
  y = y * fma(-h * y, y, 1.5);
  y = y * fma(-h * y, y, 1.5);
  return y;
}

// A read-only pointer into device (global) memory: loads go through the vector memory path
// (global_load) instead of the flat path that generic pointers take.
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(1))) T* gbl(const T* p) {
  return (const __attribute__((address_space(1))) T*)p;
}

// element i of a double2 / float2 array in global memory, loaded through the global path
__device__ __forceinline__ double2 gld2(const double2* p, int i) {
  const __attribute__((address_space(1))) double* q = (const __attribute__((address_space(1))) double*)p;
  return make_double2(q[2 * i], q[2 * i + 1]);
}
__device__ __forceinline__ double2 gld2(const double2* p, unsigned i) {
  const __attribute__((address_space(1))) double* q = (const __attribute__((address_space(1))) double*)p;
  return make_double2(q[2 * i], q[2 * i + 1]);
}
__device__ __forceinline__ float2 gld2(const float2* p, unsigned i) {
  const __attribute__((address_space(1))) float* q = (const __attribute__((address_space(1))) float*)p;
  return make_float2(q[2 * i], q[2 * i + 1]);
}

// Stream averages are summed over the trajectories of an ensemble as fixed-point integers:
// v = H 2^-20 + L 2^-60 with H = floor(v 2^20) and 0 <= L <= 2^40 (the remainder, exact in
// double, rounded to 2^-61).  Integer addition is associative, so the sums -- and the ensemble
// means made from them -- do not depend on the order in which trajectories reach a stream point
// (bit-identical for any layout, batching or mix).  |v| < vmax keeps the H sum within 2^63
// (vmax = 2^43 / n_traj); the L sums stay below 2^63 for up to 2^23 trajectories.  A value out
// of range sets bit 1 (non-finite) or bit 2 (too large) of the launch's error flag instead.
__device__ __forceinline__ void acc_add(unsigned long long* p, double v, double vmax, int* err) {
  if (!(fabs(v) < vmax)) {
    atomicOr(err, isfinite(v) ? 4 : 2);
    return;
  }
  const double h = floor(v * 0x1p20);
  const double r = v - h * 0x1p-20;  // exact: h 2^-20 is a double within 2^-20 of v
  const unsigned long long H = (unsigned long long)(long long)h;
  const unsigned long long L = (unsigned long long)rint(r * 0x1p60);
  // through the global path (the sums live in device memory), not the flat one
  __attribute__((address_space(1))) unsigned long long* g = (__attribute__((address_space(1))) unsigned long long*)p;
  __hip_atomic_fetch_add(g, H, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_fetch_add(g + 1, L, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// the mean of n values accumulated by acc_add (sums of H and of L, two's complement)
__device__ __forceinline__ double acc_mean(unsigned long long H, unsigned long long L, int n) {
  return ((double)(long long)H * 0x1p-20 + (double)L * 0x1p-60) / (double)n;
}

// Cash-Karp RK4(5) tableau (as in Boost.Odeint's runge_kutta_cash_karp54)
struct CK {
  static constexpr double c2 = 1. / 5, c3 = 3. / 10, c4 = 3. / 5, c5 = 1., c6 = 7. / 8;
  static constexpr double a21 = 1. / 5;
  static constexpr double a31 = 3. / 40, a32 = 9. / 40;
  static constexpr double a41 = 3. / 10, a42 = -9. / 10, a43 = 6. / 5;
  static constexpr double a51 = -11. / 54, a52 = 5. / 2, a53 = -70. / 27, a54 = 35. / 27;
  static constexpr double a61 = 1631. / 55296, a62 = 175. / 512, a63 = 575. / 13824, a64 = 44275. / 110592, a65 = 253. / 4096;
  static constexpr double b1 = 37. / 378, b3 = 250. / 621, b4 = 125. / 594, b6 = 512. / 1771;
  static constexpr double e1 = 37. / 378 - 2825. / 27648, e3 = 250. / 621 - 18575. / 48384,
                          e4 = 125. / 594 - 13525. / 55296, e5 = -277. / 14336, e6 = 512. / 1771 - 1. / 4;
};

}  // namespace qt
