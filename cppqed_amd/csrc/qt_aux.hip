// qt_aux.hip -- jump draws, ensemble means, the plain one-trajectory form, the addContribution
// drop-in (qt_apply_heff) and small fill kernels.
#include "qt_engine.hpp"

namespace qt {

// jump draws for the next `cap` MCWF steps of every trajectory, one Philox evaluation per lane:
// draws[traj][i] = jump_uniform(seed, traj0 + traj, ordinal[traj] + i)
__global__ void jump_draws_kernel(const unsigned long long* ordinal, unsigned long long seed, unsigned long long traj0,
                                  int n_traj, int cap, double* draws) {
  const size_t total = (size_t)n_traj * cap;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
    const int traj = (int)(i / cap), k = (int)(i % cap);
    draws[i] = jump_uniform(seed, traj0 + (unsigned long long)traj, ordinal[traj] + (unsigned long long)k);
  }
}

hipError_t launch_jump_draws(const unsigned long long* ordinal, unsigned long long seed, unsigned long long traj0,
                             int n_traj, int cap, double* draws, hipStream_t st) {
  const size_t total = (size_t)n_traj * cap;
  const int grid = (int)std::min<size_t>((total + 255) / 256, 16384);
  hipLaunchKernelGGL(jump_draws_kernel, dim3(grid), dim3(256), 0, st, ordinal, seed, traj0, n_traj, cap, draws);
  return hipGetLastError();
}

// The streamed rows of a launch from its fixed-point stream sums (AverageTrajectories and
// DensityOperatorStreamer, MCWF_Trajectory.h:406-420; the row format of Trajectory.h:428-437):
// rows[k] = (t_start + intervals[0] + ... + intervals[k], mean dtDid, processed mean averages).
// One thread per stream point; the time is summed in the host's order, the processing runs in
// place in average order (Averaged::process), as the host form does.
__global__ void stream_rows_kernel(const unsigned long long* acc, int n_int, int no, int n_traj, double t_start,
                                   const double* intervals, const int* pvar, const int* psqrt, double* rows) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n_int) return;
  double tt = t_start;
  for (int j = 0; j <= k; ++j) tt += intervals[j];
  double* r = rows + (size_t)k * (2 + no);
  r[0] = tt;
  for (int q = 0; q <= no; ++q) {
    unsigned long long H = 0, L = 0;
    for (int st = 0; st < kAccStripes; ++st) {
      const unsigned long long* a = acc + (((size_t)st * n_int + k) * (no + 1) + q) * 2;
      H += a[0];
      L += a[1];
    }
    const double m = acc_mean(H, L, n_traj);
    if (q == no) r[1] = m;
    else r[2 + q] = m;
  }
  for (int q = 0; q < no; ++q)
    if (pvar[q] >= 0) {
      r[2 + q] -= r[2 + pvar[q]] * r[2 + pvar[q]];
      if (psqrt[q]) r[2 + q] = sqrt(r[2 + q]);
    }
}

// ---------------------------------------------------------------------------
// plain form: one trajectory, one launch per RKCK try / per post-step phase
// ---------------------------------------------------------------------------

template <int ROWS, int NW, int WH>
__global__ void __launch_bounds__(64 * NW) plain_try_kernel(SysDev S, PlainTry P) {
  using Eng = Engine<ROWS, NW, WH, false>;
  constexpr int G = Eng::G;
  __shared__ double2 xs[ROWS * G];
  __shared__ double2 ps[ROWS * G];
  __shared__ double red[NW > 1 ? NW : 1];
  Eng eng(S, xs, ps, red, (int)threadIdx.x);
  double2 x[ROWS], xn[ROWS], k1[ROWS];
#pragma unroll
  for (int r = 0; r < ROWS; ++r) x[r] = eng.valid(r) ? P.x[r * G + threadIdx.x] : make_double2(0., 0.);
  long long derivs = 0;
  eng.derivative(x, k1);  // controlled_runge_kutta::try_step_v1 recomputes dxdt on every try
  const double err = sqrt(eng.tryStep(x, k1, P.dt, P.epsAbs, P.epsRel, xn, derivs));  // tryStep: squared norm
#pragma unroll
  for (int r = 0; r < ROWS; ++r) if (eng.valid(r)) P.xn[r * G + threadIdx.x] = xn[r];
  if (threadIdx.x == 0) *P.err = err;
}

// renormalise psi in place and compute the jump rates: out[0] = norm, out[1..nj] = rates
template <int ROWS, int NW, int WH>
__global__ void __launch_bounds__(64 * NW) plain_post_kernel(SysDev S, double2* psi, double* out) {
  using Eng = Engine<ROWS, NW, WH, false>;
  constexpr int G = Eng::G;
  __shared__ double2 xs[ROWS * G];
  __shared__ double2 ps[ROWS * G];
  __shared__ double red[NW > 1 ? NW : 1];
  Eng eng(S, xs, ps, red, (int)threadIdx.x);
  double2 x[ROWS];
  double s = 0.;
#pragma unroll
  for (int r = 0; r < ROWS; ++r) { x[r] = eng.valid(r) ? psi[r * G + threadIdx.x] : make_double2(0., 0.); s += cabs2(x[r]); }
  const double nrm = sqrt(eng.gsum(s));
#pragma unroll
  for (int r = 0; r < ROWS; ++r) { x[r].x /= nrm; x[r].y /= nrm; if (eng.valid(r)) psi[r * G + threadIdx.x] = x[r]; }
  eng.stagePsi(x);
  double rt[kMaxJumps];
  eng.rates(rt);
  if (threadIdx.x == 0) { out[0] = nrm; for (int m = 0; m < S.nj; ++m) out[1 + m] = rt[m]; }
}

// psi <- J_m psi / nf
template <int ROWS, int NW, int WH>
__global__ void __launch_bounds__(64 * NW) plain_jump_kernel(SysDev S, double2* psi, int m, double nf) {
  using Eng = Engine<ROWS, NW, WH, false>;
  constexpr int G = Eng::G;
  __shared__ double2 xs[ROWS * G];
  __shared__ double2 ps[ROWS * G];
  __shared__ double red[NW > 1 ? NW : 1];
  Eng eng(S, xs, ps, red, (int)threadIdx.x);
  double2 x[ROWS];
#pragma unroll
  for (int r = 0; r < ROWS; ++r) x[r] = eng.valid(r) ? psi[r * G + threadIdx.x] : make_double2(0., 0.);
  eng.stagePsi(x);
#pragma unroll
  for (int r = 0; r < ROWS; ++r) if (eng.valid(r)) { double2 y = eng.jumpRow(m, r); psi[r * G + threadIdx.x] = make_double2(y.x / nf, y.y / nf); }
}

// observables of one state: out[q]
template <int ROWS, int NW, int WH>
__global__ void __launch_bounds__(64 * NW) plain_obs_kernel(SysDev S, const double2* psi, double* out) {
  using Eng = Engine<ROWS, NW, WH, false>;
  constexpr int G = Eng::G;
  __shared__ double2 xs[ROWS * G];
  __shared__ double2 ps[ROWS * G];
  __shared__ double red[NW > 1 ? NW : 1];
  Eng eng(S, xs, ps, red, (int)threadIdx.x);
  double2 x[ROWS];
#pragma unroll
  for (int r = 0; r < ROWS; ++r) x[r] = eng.valid(r) ? psi[r * G + threadIdx.x] : make_double2(0., 0.);
  eng.stagePsi(x);
  for (int q = 0; q < S.no; ++q) { double v = eng.observable(q, x); if (threadIdx.x == 0) out[q] = v; }
}

// dpsidt += (H_nH/i)(t - t0) psi for a batch of state vectors, one group per vector (qt_apply_heff)
template <int ROWS, int NW, int WH, bool F32, bool TD>
__global__ void __launch_bounds__(64 * NW) apply_heff_kernel(SysDev S, const double2* psi, double2* dpsidt, int n,
                                                             double tau) {
  using Eng = Engine<ROWS, NW, WH, F32, false, TD>;
  using T = typename Eng::T;
  constexpr int G = Eng::G;
  __shared__ T xs[ROWS * G];
  __shared__ double2 ps[1];
  __shared__ double red[NW > 1 ? NW : 1];
  __shared__ T ph[kMaxFreqs];
  Eng eng(S, xs, ps, red, (int)threadIdx.x, ph);
  for (int v = blockIdx.x; v < n; v += gridDim.x) {
    double2 x[ROWS];
    T k[ROWS];
#pragma unroll
    for (int r = 0; r < ROWS; ++r) x[r] = eng.valid(r) ? psi[(size_t)v * S.dim + r * G + threadIdx.x] : make_double2(0., 0.);
    eng.derivative(x, k, tau);
#pragma unroll
    for (int r = 0; r < ROWS; ++r)
      if (eng.valid(r)) {
        double2& o = dpsidt[(size_t)v * S.dim + r * G + threadIdx.x];
        const double2 kd = to_d(k[r]);
        o.x += kd.x; o.y += kd.y;
      }
  }
}


// p[0..n) = v  (a kernel rather than a memset node, so that the launch sequence can be captured)
__global__ void fill_int_kernel(int* p, int n, int v) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

hipError_t launch_fill_int(int* p, int n, int v, hipStream_t st) {
  hipLaunchKernelGGL(fill_int_kernel, dim3(1), dim3(64), 0, st, p, n, v);
  return hipGetLastError();
}

// p[0..n) = 0 for 64-bit words
__global__ void zero_u64_kernel(unsigned long long* p, size_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0ull;
}

hipError_t launch_zero_u64(unsigned long long* p, size_t n, hipStream_t st) {
  const int grid = (int)std::min<size_t>((n + 255) / 256, 1024);
  hipLaunchKernelGGL(zero_u64_kernel, dim3(grid > 0 ? grid : 1), dim3(256), 0, st, p, n);
  return hipGetLastError();
}

hipError_t launch_stream_rows(const unsigned long long* acc, int n_int, int no, int n_traj, double t_start,
                              const double* intervals, const int* pvar, const int* psqrt, double* rows, hipStream_t st) {
  hipLaunchKernelGGL(stream_rows_kernel, dim3((n_int + 63) / 64), dim3(64), 0, st, acc, n_int, no, n_traj, t_start,
                     intervals, pvar, psqrt, rows);
  return hipGetLastError();
}

hipError_t launch_plain_try(int rows, int nw, int wh, const SysDev& S, const PlainTry& P, hipStream_t st) {
#define QT_CASE(R, N, W) \
  if (rows == R && nw == N && wh == W) { hipLaunchKernelGGL((plain_try_kernel<R, N, W>), dim3(1), dim3(64 * N), 0, st, S, P); return hipGetLastError(); }
  QT_LAYOUTS(QT_CASE)
#undef QT_CASE
  return hipErrorInvalidValue;
}

hipError_t launch_plain_post(int rows, int nw, int wh, const SysDev& S, double2* psi, double* out, hipStream_t st) {
#define QT_CASE(R, N, W) \
  if (rows == R && nw == N && wh == W) { hipLaunchKernelGGL((plain_post_kernel<R, N, W>), dim3(1), dim3(64 * N), 0, st, S, psi, out); return hipGetLastError(); }
  QT_LAYOUTS(QT_CASE)
#undef QT_CASE
  return hipErrorInvalidValue;
}

hipError_t launch_plain_jump(int rows, int nw, int wh, const SysDev& S, double2* psi, int m, double nf, hipStream_t st) {
#define QT_CASE(R, N, W) \
  if (rows == R && nw == N && wh == W) { hipLaunchKernelGGL((plain_jump_kernel<R, N, W>), dim3(1), dim3(64 * N), 0, st, S, psi, m, nf); return hipGetLastError(); }
  QT_LAYOUTS(QT_CASE)
#undef QT_CASE
  return hipErrorInvalidValue;
}

hipError_t launch_plain_obs(int rows, int nw, int wh, const SysDev& S, const double2* psi, double* out, hipStream_t st) {
#define QT_CASE(R, N, W) \
  if (rows == R && nw == N && wh == W) { hipLaunchKernelGGL((plain_obs_kernel<R, N, W>), dim3(1), dim3(64 * N), 0, st, S, psi, out); return hipGetLastError(); }
  QT_LAYOUTS(QT_CASE)
#undef QT_CASE
  return hipErrorInvalidValue;
}

hipError_t launch_apply_heff(int rows, int nw, int wh, bool f32, bool td, const SysDev& S, const double2* psi,
                             double2* dpsidt, int n, double tau, hipStream_t st) {
  const int grid = n < 4096 ? n : 4096;
#define QT_CASE(R, N, W)                                                                                      \
  if (rows == R && nw == N && wh == W) {                                                                      \
    if (td) {                                                                                                 \
      if (f32) hipLaunchKernelGGL((apply_heff_kernel<R, N, W, true, true>), dim3(grid), dim3(64 * N), 0, st, S, psi, dpsidt, n, tau); \
      else hipLaunchKernelGGL((apply_heff_kernel<R, N, W, false, true>), dim3(grid), dim3(64 * N), 0, st, S, psi, dpsidt, n, tau);    \
    } else {                                                                                                  \
      if (f32) hipLaunchKernelGGL((apply_heff_kernel<R, N, W, true, false>), dim3(grid), dim3(64 * N), 0, st, S, psi, dpsidt, n, tau); \
      else hipLaunchKernelGGL((apply_heff_kernel<R, N, W, false, false>), dim3(grid), dim3(64 * N), 0, st, S, psi, dpsidt, n, tau);    \
    }                                                                                                         \
    return hipGetLastError();                                                                                 \
  }
  QT_LAYOUTS(QT_CASE)
#undef QT_CASE
  return hipErrorInvalidValue;
}

}  // namespace qt
