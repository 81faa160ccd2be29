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
  using Eng = Engine<ROWS, NW, WH, F32, TD>;
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


// The ensemble's density operator on request (AverageTrajectories<MCWF_Trajectory>,
// MCWF_Trajectory.h:406-420, with StateVector::addTo, StateVector.h:173-180):
// rho_ij = (1/N) sum_k psi_k[i] conj(psi_k[j]).  A work group owns a 64 x 64 tile of rho and a
// range of `per` trajectories (split blockIdx.z); 16 trajectories' tile segments are staged in LDS
// at a time and each thread accumulates a 4 x 4 block (rows ty + 16 r, columns tx + 16 c).  With one
// split the scaled sums go straight to rho; otherwise each split writes its partial sums to
// part[split] and rho_reduce_kernel adds them in split order (the result does not depend on
// scheduling).
__global__ void __launch_bounds__(256) ensemble_rho_kernel(const double2* __restrict__ psi, int n, int d, int per,
                                                           double scale, double2* __restrict__ out) {
  __shared__ double2 A[16][64], B[16][64];
  const int ti = blockIdx.y * 64, tj = blockIdx.x * 64;
  const int kBeg = blockIdx.z * per, kEnd = min(n, kBeg + per);
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  double2 acc[4][4];
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) acc[r][c] = make_double2(0., 0.);
  for (int k0 = kBeg; k0 < kEnd; k0 += 16) {
    for (int e = threadIdx.x; e < 16 * 64; e += 256) {
      const int kk = e >> 6, ii = e & 63, k = k0 + kk;
      const bool in = k < kEnd;
      A[kk][ii] = in && ti + ii < d ? psi[(size_t)k * d + ti + ii] : make_double2(0., 0.);
      B[kk][ii] = in && tj + ii < d ? psi[(size_t)k * d + tj + ii] : make_double2(0., 0.);
    }
    __syncthreads();
#pragma unroll 4
    for (int kk = 0; kk < 16; ++kk) {
      double2 a[4], b[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) { a[r] = A[kk][ty + 16 * r]; b[r] = B[kk][tx + 16 * r]; }
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c) {  // a conj(b)
          acc[r][c].x = fma(a[r].x, b[c].x, fma(a[r].y, b[c].y, acc[r][c].x));
          acc[r][c].y = fma(a[r].y, b[c].x, fma(-a[r].x, b[c].y, acc[r][c].y));
        }
    }
    __syncthreads();
  }
  double2* o = out + (size_t)blockIdx.z * d * d;
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int i = ti + ty + 16 * r, j = tj + tx + 16 * c;
      if (i < d && j < d) o[(size_t)i * d + j] = make_double2(acc[r][c].x * scale, acc[r][c].y * scale);
    }
}

__global__ void rho_reduce_kernel(const double2* __restrict__ part, int splits, size_t dd, double scale,
                                  double2* __restrict__ rho) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < dd; i += (size_t)gridDim.x * blockDim.x) {
    double2 s = part[i];
    for (int q = 1; q < splits; ++q) { s.x += part[q * dd + i].x; s.y += part[q * dd + i].y; }
    rho[i] = make_double2(s.x * scale, s.y * scale);
  }
}

// rho (d x d, row-major) of the n states psi[n][d]; part: splits * d * d scratch when splits > 1
hipError_t launch_ensemble_rho(const double2* psi, int n, int d, int splits, double2* part, double2* rho,
                               hipStream_t st) {
  const int per = ((n + splits - 1) / splits + 15) / 16 * 16;
  splits = (n + per - 1) / per;
  const int tiles = (d + 63) / 64;
  if (splits <= 1) {
    hipLaunchKernelGGL(ensemble_rho_kernel, dim3(tiles, tiles, 1), dim3(256), 0, st, psi, n, d, n, 1.0 / n, rho);
    return hipGetLastError();
  }
  hipLaunchKernelGGL(ensemble_rho_kernel, dim3(tiles, tiles, splits), dim3(256), 0, st, psi, n, d, per, 1.0, part);
  const size_t dd = (size_t)d * d;
  const int grid = (int)std::min<size_t>((dd + 255) / 256, 4096);
  hipLaunchKernelGGL(rho_reduce_kernel, dim3(grid), dim3(256), 0, st, part, splits, dd, 1.0 / n, rho);
  return hipGetLastError();
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
