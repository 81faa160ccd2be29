// qt_plain.hip -- the plain one-trajectory form: the MCWF step's control flow on the host, one
// launch per RKCK try and per post-step phase (the shape of the reference's serial driver, kept
// as the baseline the batched kernels are measured against).
#include "qt_engine.hpp"

namespace qt {

// One try from P.x: P.xn = the 5th-order proposal, *P.err = odeint's error norm, or GSL's error
// ratio with GSL (ODE_GSL.h:44-84).  Interaction picture (TD): stage times t + c dt - t0.
template <int ROWS, int NW, int WH, bool TD, bool GSL>
__global__ void __launch_bounds__(64 * NW) plain_try_kernel(SysDev S, PlainTry P) {
  using Eng = Engine<ROWS, NW, WH, false, TD>;
  constexpr int G = Eng::G;
  __shared__ double2 xs[ROWS * G];
  __shared__ double2 ps[ROWS * G];
  __shared__ double red[NW > 1 ? NW : 1];
  __shared__ double2 ph[TD ? kMaxFreqs : 1];
  Eng eng(S, xs, ps, red, (int)threadIdx.x, TD ? ph : nullptr);
  double2 x[ROWS], xn[ROWS], k1[ROWS];
#pragma unroll
  for (int r = 0; r < ROWS; ++r) x[r] = eng.valid(r) ? P.x[r * G + threadIdx.x] : make_double2(0., 0.);
  long long derivs = 0;
  eng.derivative(x, k1, P.t - P.t0);  // controlled_runge_kutta::try_step_v1 recomputes dxdt on every try
  double err;
  if constexpr (GSL) err = eng.template tryStep<true>(x, k1, P.dt, P.epsAbs, P.epsRel, xn, derivs, P.t, P.t0);
  else err = sqrt(eng.tryStep(x, k1, P.dt, P.epsAbs, P.epsRel, xn, derivs, P.t, P.t0));  // squared norm
#pragma unroll
  for (int r = 0; r < ROWS; ++r) if (eng.valid(r)) P.xn[r * G + threadIdx.x] = xn[r];
  if (threadIdx.x == 0) *P.err = err;
}

// After a step: Exact::actWithU over dtau (interaction picture, S.exz set and dtau != 0), then
// renormalise psi in place and compute the jump rates: out[0] = norm, out[1..nj] = rates
template <int ROWS, int NW, int WH>
__global__ void __launch_bounds__(64 * NW) plain_post_kernel(SysDev S, double2* psi, double* out, double dtau) {
  using Eng = Engine<ROWS, NW, WH, false>;
  constexpr int G = Eng::G;
  __shared__ double2 xs[ROWS * G];
  __shared__ double2 ps[ROWS * G];
  __shared__ double red[NW > 1 ? NW : 1];
  Eng eng(S, xs, ps, red, (int)threadIdx.x);
  double2 x[ROWS];
  double s = 0.;
#pragma unroll
  for (int r = 0; r < ROWS; ++r) {
    x[r] = eng.valid(r) ? psi[r * G + threadIdx.x] : make_double2(0., 0.);
    if (S.exz && dtau != 0. && eng.valid(r)) {  // MCWF_Trajectory.h:269-272
      const double2 z = S.exz[r * G + threadIdx.x];
      const double e = exp(-z.x * dtau);
      double sn, cs;
      sincos(-z.y * dtau, &sn, &cs);
      x[r] = cmul(x[r], make_double2(e * cs, e * sn));
    }
    s += cabs2(x[r]);
  }
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

hipError_t launch_plain_try(int rows, int nw, int wh, bool td, bool gsl, const SysDev& S, const PlainTry& P,
                            hipStream_t st) {
#define QT_LAUNCH(R, N, W, A, B) hipLaunchKernelGGL((plain_try_kernel<R, N, W, A, B>), dim3(1), dim3(64 * N), 0, st, S, P)
#define QT_CASE(R, N, W)                                          \
  if (rows == R && nw == N && wh == W) {                          \
    if (td && gsl) QT_LAUNCH(R, N, W, true, true);                \
    else if (td) QT_LAUNCH(R, N, W, true, false);                 \
    else if (gsl) QT_LAUNCH(R, N, W, false, true);                \
    else QT_LAUNCH(R, N, W, false, false);                        \
    return hipGetLastError();                                     \
  }
  QT_LAYOUTS(QT_CASE)
#undef QT_CASE
#undef QT_LAUNCH
  return hipErrorInvalidValue;
}

hipError_t launch_plain_post(int rows, int nw, int wh, const SysDev& S, double2* psi, double* out, double dtau,
                             hipStream_t st) {
#define QT_CASE(R, N, W) \
  if (rows == R && nw == N && wh == W) { hipLaunchKernelGGL((plain_post_kernel<R, N, W>), dim3(1), dim3(64 * N), 0, st, S, psi, out, dtau); return hipGetLastError(); }
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

}  // namespace qt
