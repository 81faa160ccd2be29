// qt_aux.hip -- jump draws, the stream rows of the fixed-point sums, the addContribution drop-in
// (qt_apply_heff) and small fill kernels.
#include "qt_engine.hpp"

namespace qt {

// jump draws for the next `cap` MCWF steps of every trajectory, one Philox evaluation per lane:
// draws[traj][i] = jump_uniform(seed, traj0 + traj, ordinal[traj] + i); base[traj] = ordinal[traj]
// (when base is given: the ordinal of draws[traj][0], for kernels that resume a trajectory
// mid-launch)
__global__ void jump_draws_kernel(const unsigned long long* ordinal, unsigned long long seed, unsigned long long traj0,
                                  int n_traj, int cap, double* draws, unsigned long long* base) {
  const size_t total = (size_t)n_traj * cap;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
    const int traj = (int)(i / cap), k = (int)(i % cap);
    const unsigned long long o = ordinal[traj];
    draws[i] = jump_uniform(seed, traj0 + (unsigned long long)traj, o + (unsigned long long)k);
    if (base && k == 0) base[traj] = o;
  }
}

hipError_t launch_jump_draws(const unsigned long long* ordinal, unsigned long long seed, unsigned long long traj0,
                             int n_traj, int cap, double* draws, hipStream_t st, unsigned long long* base) {
  const size_t total = (size_t)n_traj * cap;
  const int grid = (int)std::min<size_t>((total + 255) / 256, 16384);
  hipLaunchKernelGGL(jump_draws_kernel, dim3(grid), dim3(256), 0, st, ordinal, seed, traj0, n_traj, cap, draws, base);
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
  const int grid = std::max(1, std::min((n + 255) / 256, 1024));
  hipLaunchKernelGGL(fill_int_kernel, dim3(grid), dim3(256), 0, st, p, n, v);
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
