// qt_subgroup.hip -- batched MCWF kernel with several trajectories per wavefront.
//
// A wavefront is split into 64/G groups of G lanes (G = 16, 32 or 64); each group owns
// one trajectory at a time, lane j of a group holding rows {j, j+G, ..., j+(ROWS-1)G} of
// psi and of the RKCK stage derivatives in registers.  Every iteration of the main loop
// performs exactly one RKCK try for every group (six H_nH/i applications, the
// reference's count per try) and then the post-step logic of MCWF_Trajectory::step
// (MCWF_Trajectory.h:347-374) for the groups whose try was accepted.  Control flow that
// differs between groups (accept / reject, step-back, jump, interval end) is expressed
// with per-group selects on uniform code, so all 64 lanes stay active and the group
// reductions can use DPP lane moves.  Rare events (a stream point, a finished trajectory)
// are handled in a wave-uniform branch taken when any group needs it.
//
// Step-size control, tolerances and the order of operations follow ode_engine::Base::step
// (CPPQEDutils/ODE.h:234-262) with Boost.Odeint's controlled RKCK54, exactly as the
// one-trajectory-per-group kernel in qt_mcwf.hip does.

#include "qt_device.hpp"
#include "qt_kernels.hpp"
#include "qt_philox.h"

namespace qt {

// sum / max over the G lanes of a group; all 64 lanes of the wavefront must be active
template <int G>
__device__ __forceinline__ double gsum(double v) {
  v += dpp<0xB1>(v); v += dpp<0x4E>(v); v += dpp<0x141>(v); v += dpp<0x140>(v);
  if constexpr (G == 32) v += __shfl_xor(v, 16, 64);
  if constexpr (G == 64) v = (readlane(v, 0) + readlane(v, 16)) + (readlane(v, 32) + readlane(v, 48));
  return v;
}
template <int G>
__device__ __forceinline__ double gmax(double v) {
  v = fmax(v, dpp<0xB1>(v)); v = fmax(v, dpp<0x4E>(v)); v = fmax(v, dpp<0x141>(v)); v = fmax(v, dpp<0x140>(v));
  if constexpr (G == 32) v = fmax(v, __shfl_xor(v, 16, 64));
  if constexpr (G == 64) v = fmax(fmax(readlane(v, 0), readlane(v, 16)), fmax(readlane(v, 32), readlane(v, 48)));
  return v;
}

template <int G, int ROWS, int WH, bool F32>
__global__ void __launch_bounds__(64) mcwf_sub_kernel(SysDev S, EnsDev E) {
  using T = typename CT<F32>::T;
  constexpr int TPW = 64 / G;        // trajectories per wavefront
  constexpr int NR = ROWS * G;       // rows per group slot
  constexpr bool kRegs = ROWS * WH <= 24;
  constexpr int RR = kRegs ? ROWS : 1, RW = kRegs ? WH : 1;

  __shared__ T xs[TPW * NR];             // RK stage vectors (derivative precision)
  __shared__ double2 ps[TPW * NR];       // FP64 state (rates, jumps, observables)
  __shared__ double2 pc[TPW * NR];       // psi at the start of the MCWF step
  __shared__ double rts[TPW][kMaxJumps]; // jump rates of the step just finished
  __shared__ long long cnt[TPW][kCounters];
  __shared__ int claim[TPW];

  const int lane = threadIdx.x, g = lane / G, j = lane % G;
  T* const xg = xs + g * NR;
  double2* const pg = ps + g * NR;
  double2* const cg = pc + g * NR;
  const int dim = S.dim, nj = S.nj;
  const bool hasLi = nj > 0;

  int hcol[RR][RW];
  T hval[RR][RW];
  if constexpr (kRegs) {
#pragma unroll
    for (int r = 0; r < ROWS; ++r)
#pragma unroll
      for (int w = 0; w < WH; ++w) {
        const int row = r * G + j;
        const bool ok = row < dim && w < S.wh;
        hcol[r][w] = ok ? S.hcol[w * dim + row] : 0;
        if constexpr (F32) hval[r][w] = ok ? S.hval32[w * dim + row] : make_float2(0.f, 0.f);
        else hval[r][w] = ok ? S.hval[w * dim + row] : make_double2(0., 0.);
      }
  }
  auto valid = [&](int r) { return r * G + j < dim; };

  // k = (H_nH/i) v for this lane's rows; v is written to the group's stage slot first
  auto derivative = [&](const double2 (&v)[ROWS], T (&kk)[ROWS]) {
    __syncthreads();
#pragma unroll
    for (int r = 0; r < ROWS; ++r) xg[r * G + j] = from_d<T>(v[r]);
    __syncthreads();
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
      T a0 = from_d<T>(make_double2(0., 0.)), a1 = a0;
      if constexpr (kRegs) {
#pragma unroll
        for (int w = 0; w < WH; w += 2) {
          a0 = cfma(hval[r][w], xg[hcol[r][w]], a0);
          if (w + 1 < WH) a1 = cfma(hval[r][w + 1], xg[hcol[r][w + 1]], a1);
        }
      } else {
        const int row = r * G + j;
        if (row < dim) {
          for (int w = 0; w < S.wh; ++w) {
            T c;
            if constexpr (F32) c = S.hval32[w * dim + row]; else c = S.hval[w * dim + row];
            a0 = cfma(c, xg[S.hcol[w * dim + row]], a0);
          }
        }
      }
      kk[r].x = a0.x + a1.x; kk[r].y = a0.y + a1.y;
    }
  };
  auto stagePsi = [&](const double2 (&v)[ROWS]) {
    __syncthreads();
#pragma unroll
    for (int r = 0; r < ROWS; ++r) pg[r * G + j] = v[r];
    __syncthreads();
  };
  // (J_m psi) for row r, psi staged in pg; m may differ between groups
  auto jumpRow = [&](int m, int r) {
    const int row = r * G + j;
    double2 acc = make_double2(0., 0.);
    if (row < dim) {
      const int base = m * S.wj * dim + row;
      for (int w = 0; w < S.wj; ++w) acc = cfma(S.jval[base + w * dim], pg[S.jcol[base + w * dim]], acc);
    }
    return acc;
  };

  // per-group state (identical in the G lanes of a group)
  int traj = 0, k = 0;
  bool live = false, stepBegin = false, setup = false;
  double t = 0., dtDid = 0., dtTry = 0., endTime = 0., tCache = 0., dtDidCache = 0., timeBefore = 0., nextDtTry = 0.;
  unsigned long long ord = 0, key = 0;
  double2 x[ROWS];
#pragma unroll
  for (int r = 0; r < ROWS; ++r) x[r] = make_double2(0., 0.);

  auto fetch = [&](bool need) {
    __syncthreads();
    if (need && j == 0) claim[g] = atomicAdd(E.work, 1);
    __syncthreads();
    if (need) {
      traj = claim[g];
      live = traj < E.n_traj;
      if (live) {
#pragma unroll
        for (int r = 0; r < ROWS; ++r) x[r] = valid(r) ? E.psi[(size_t)traj * dim + r * G + j] : make_double2(0., 0.);
        t = E.t[traj]; dtDid = E.dtDid[traj]; dtTry = E.dtTry[traj]; ord = E.ordinal[traj];
        key = E.traj0 + (unsigned long long)traj;
        if (j < kCounters) cnt[g][j] = E.counters[(size_t)traj * kCounters + j];
        k = 0;
        endTime = t + E.intervals[0];
        stepBegin = true;
      }
    }
    __syncthreads();
  };

  fetch(true);

  for (;;) {
    // ---- stream points and trajectory turnover (wave-uniform branch) ----
    for (;;) {
      const bool atEnd = live && stepBegin && (endTime - t == 0.);
      if (!__any(atEnd)) break;
      stagePsi(x);
      for (int q = 0; q < S.no; ++q) {
        const bool im = S.oimag[q] != 0;
        double s = 0.;
#pragma unroll
        for (int r = 0; r < ROWS; ++r) {
          const int row = r * G + j;
          if (row >= dim) continue;
          double2 acc = make_double2(0., 0.);
          const int base = q * S.wo * dim + row;
          const int wq = S.owid[q];
          for (int w = 0; w < wq; ++w) acc = cfma(S.oval[base + w * dim], pg[S.ocol[base + w * dim]], acc);
          s += im ? (x[r].x * acc.y - x[r].y * acc.x) : (x[r].x * acc.x + x[r].y * acc.y);
        }
        s = gsum<G>(s);
        if (atEnd && j == 0) {
          acc_add(E.acc + (((size_t)(traj % kAccStripes) * E.n_intervals + k) * (S.no + 1) + q) * 2, s, E.acc_max,
                  E.err_flag);
          if (k == E.n_intervals - 1) E.last_obs[(size_t)traj * S.no + q] = s;
        }
      }
      bool finished = false;
      if (atEnd) {
        if (j == 0)
          acc_add(E.acc + (((size_t)(traj % kAccStripes) * E.n_intervals + k) * (S.no + 1) + S.no) * 2, dtDid,
                  E.acc_max, E.err_flag);
        ++k;
        if (k == E.n_intervals) finished = true;
        else endTime = t + E.intervals[k];
      }
      if (finished) {
#pragma unroll
        for (int r = 0; r < ROWS; ++r) if (valid(r)) E.psi[(size_t)traj * dim + r * G + j] = x[r];
        if (j == 0) { E.t[traj] = t; E.dtDid[traj] = dtDid; E.dtTry[traj] = dtTry; E.ordinal[traj] = ord; }
      }
      __syncthreads();
      if (finished && j < kCounters) E.counters[(size_t)traj * kCounters + j] = cnt[g][j];
      if (finished) live = false;
      fetch(finished);
    }
    if (!__any(live)) return;

    // ---- a new MCWF step caches psi; a new step or a step-back restarts ode_engine::Base::step ----
    if (live && stepBegin) {
#pragma unroll
      for (int r = 0; r < ROWS; ++r) cg[r * G + j] = x[r];
      tCache = t; dtDidCache = dtDid;
      setup = true;
    }
    if (live && setup) {
      const double Dt = endTime - t;
      timeBefore = t;
      nextDtTry = (fabs(Dt) < fabs(dtTry / 10.) ? dtTry : 0.);
      if ((Dt > 0.) != (dtTry > 0.)) dtTry = -dtDid;
      if (fabs(dtTry) > fabs(Dt)) dtTry = Dt;
    }
    stepBegin = false; setup = false;

    // ---- one RKCK54 try for every group ----
    T k1[ROWS], k2[ROWS], k3[ROWS], k4[ROWS], k5[ROWS], k6[ROWS];
    double2 v[ROWS], xn[ROWS];
    const double dt = dtTry;
    derivative(x, k1);
    {
      const double s21 = CK::a21 * dt;
#pragma unroll
      for (int r = 0; r < ROWS; ++r) v[r] = axpy(s21, to_d(k1[r]), x[r]);
    }
    derivative(v, k2);
    {
      const double q1 = CK::a31 * dt, q2 = CK::a32 * dt;
#pragma unroll
      for (int r = 0; r < ROWS; ++r) v[r] = axpy(q2, to_d(k2[r]), axpy(q1, to_d(k1[r]), x[r]));
    }
    derivative(v, k3);
    {
      const double q1 = CK::a41 * dt, q2 = CK::a42 * dt, q3 = CK::a43 * dt;
#pragma unroll
      for (int r = 0; r < ROWS; ++r) v[r] = axpy(q3, to_d(k3[r]), axpy(q2, to_d(k2[r]), axpy(q1, to_d(k1[r]), x[r])));
    }
    derivative(v, k4);
    {
      const double q1 = CK::a51 * dt, q2 = CK::a52 * dt, q3 = CK::a53 * dt, q4 = CK::a54 * dt;
#pragma unroll
      for (int r = 0; r < ROWS; ++r)
        v[r] = axpy(q4, to_d(k4[r]), axpy(q3, to_d(k3[r]), axpy(q2, to_d(k2[r]), axpy(q1, to_d(k1[r]), x[r]))));
    }
    derivative(v, k5);
    {
      const double q1 = CK::a61 * dt, q2 = CK::a62 * dt, q3 = CK::a63 * dt, q4 = CK::a64 * dt, q5 = CK::a65 * dt;
#pragma unroll
      for (int r = 0; r < ROWS; ++r)
        v[r] = axpy(q5, to_d(k5[r]), axpy(q4, to_d(k4[r]), axpy(q3, to_d(k3[r]), axpy(q2, to_d(k2[r]), axpy(q1, to_d(k1[r]), x[r])))));
    }
    derivative(v, k6);
    double e = 0., n2 = 0.;
    {
      const double q1 = CK::b1 * dt, q3 = CK::b3 * dt, q4 = CK::b4 * dt, q6 = CK::b6 * dt;
      const double f1 = CK::e1 * dt, f3 = CK::e3 * dt, f4 = CK::e4 * dt, f5 = CK::e5 * dt, f6 = CK::e6 * dt;
#pragma unroll
      for (int r = 0; r < ROWS; ++r) {
        const double2 d1 = to_d(k1[r]), d3 = to_d(k3[r]), d4 = to_d(k4[r]), d5 = to_d(k5[r]), d6 = to_d(k6[r]);
        xn[r] = axpy(q6, d6, axpy(q4, d4, axpy(q3, d3, axpy(q1, d1, x[r]))));
        double2 xe = make_double2(f1 * d1.x, f1 * d1.y);
        xe = axpy(f6, d6, axpy(f5, d5, axpy(f4, d4, axpy(f3, d3, xe))));
        if (valid(r)) {
          const double ax = sqrt(cabs2(x[r])), ak = sqrt(cabs2(d1)), ae = sqrt(cabs2(xe));
          e = fmax(e, ae / (E.epsAbs + E.epsRel * (ax + dt * ak)));
          n2 += cabs2(xn[r]);
        }
      }
    }
    const double err = gmax<G>(e);
    const double nrm = sqrt(gsum<G>(n2));
    const bool accepted = live && !(err > 1.0);
    if (live && j == 0) {
      cnt[g][kTries] += 1; cnt[g][kDerivs] += 6;
      if (!accepted) cnt[g][kFailed] += 1;
    }
    if (live) {
      if (!accepted) {
        dtTry *= fmax(9. / 10. * ctrl_pow(err, -1. / (4 - 1)), 1. / 5.);
      } else {
        t += dtTry;
        if (err < 0.5) dtTry *= 9. / 10. * ctrl_pow(fmax(0.00032, err), -1. / 5);  // max(5^-5, err)
#pragma unroll
        for (int r = 0; r < ROWS; ++r) { x[r].x = xn[r].x / nrm; x[r].y = xn[r].y / nrm; }
        dtDid = t - timeBefore;
        if (nextDtTry != 0.) dtTry = nextDtTry;
        if (j == 0) cnt[g][kOdeSteps] += 1;
      }
    }

    // ---- post-step: rates, dpLimit management, jump (MCWF_Trajectory.h:294-344) ----
    bool done = accepted;
    if (hasLi && __any(accepted)) {
      stagePsi(x);
      for (int m = 0; m < nj; ++m) {
        double s = 0.;
#pragma unroll
        for (int r = 0; r < ROWS; ++r) s += cabs2(jumpRow(m, r));
        s = gsum<G>(s);
        if (j == 0) rts[g][m] = s;
      }
      __syncthreads();
      double totalRate = 0.;
      for (int m = 0; m < nj; ++m) totalRate += rts[g][m];
      if (accepted && !isfinite(totalRate) && j == 0) atomicOr(E.err_flag, 2);  // LiouvilleanAveragedCommon.h:82-93
      const double sugg = E.dpLimit / totalRate;
      const bool stepBack = accepted && totalRate * dtDid > E.overshootTol * E.dpLimit;
      if (stepBack) {
        t = tCache; dtDid = dtDidCache; dtTry = sugg;
#pragma unroll
        for (int r = 0; r < ROWS; ++r) x[r] = cg[r * G + j];
        if (j == 0) cnt[g][kStepBacks] += 1;
        setup = true;
      }
      done = accepted && !stepBack;
      int no = 0;
      bool jump = false;
      if (done) {
        if (totalRate * dtTry > E.dpLimit && j == 0) cnt[g][kOvershoots] += 1;
        dtTry = fmin(dtTry, sugg);
        double random = jump_uniform(E.seed, key, ord++) / dtDid;
        for (; random > 0. && no != nj; random -= rts[g][no++]) ;
        jump = random < 0.;
        if (jump) --no;
      }
      if (__any(jump)) {
        const double nf = sqrt(rts[g][no]);
        double2 y[ROWS];
#pragma unroll
        for (int r = 0; r < ROWS; ++r) y[r] = jumpRow(no, r);
        if (jump) {
          if (!isfinite(nf) && j == 0) atomicOr(E.err_flag, 1);
#pragma unroll
          for (int r = 0; r < ROWS; ++r) x[r] = make_double2(y[r].x / nf, y[r].y / nf);
          if (j == 0) {
            const long long nJ = cnt[g][kJumps];
            if (nJ < E.jcap) { E.jt[(size_t)traj * E.jcap + nJ] = t; E.jch[(size_t)traj * E.jcap + nJ] = no; }
            cnt[g][kJumps] = nJ + 1;
          }
        }
      }
    }
    if (done && j == 0) cnt[g][kMcwfSteps] += 1;
    stepBegin = done;
  }
}

// ---------------------------------------------------------------------------
// variants and launch table
// ---------------------------------------------------------------------------

#define QT_SUB(X) \
  X(16, 2, 4) X(16, 4, 4) X(32, 1, 4) X(32, 2, 4) X(64, 1, 4) \
  X(16, 2, 8) X(16, 4, 8) X(32, 1, 8) X(32, 2, 8) X(64, 1, 8) \
  X(32, 1, 16) X(64, 1, 16)

bool sub_supported(int g, int rows, int wh) {
#define QT_CASE(GG, R, W) if (g == GG && rows == R && wh == W) return true;
  QT_SUB(QT_CASE)
#undef QT_CASE
  return false;
}

int sub_blocks_per_cu(int g, int rows, int wh, bool f32) {
  int n = 0;
#define QT_CASE(GG, R, W)                                                                                \
  if (g == GG && rows == R && wh == W) {                                                                 \
    if (f32) (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, mcwf_sub_kernel<GG, R, W, true>, 64, 0);  \
    else (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, mcwf_sub_kernel<GG, R, W, false>, 64, 0);    \
    return n;                                                                                            \
  }
  QT_SUB(QT_CASE)
#undef QT_CASE
  return 0;
}

hipError_t launch_sub(int g, int rows, int wh, bool f32, int grid, const SysDev& S, const EnsDev& E, hipStream_t st) {
#define QT_CASE(GG, R, W)                                                                                  \
  if (g == GG && rows == R && wh == W) {                                                                   \
    if (f32) hipLaunchKernelGGL((mcwf_sub_kernel<GG, R, W, true>), dim3(grid), dim3(64), 0, st, S, E);     \
    else hipLaunchKernelGGL((mcwf_sub_kernel<GG, R, W, false>), dim3(grid), dim3(64), 0, st, S, E);        \
    return hipGetLastError();                                                                              \
  }
  QT_SUB(QT_CASE)
#undef QT_CASE
  return hipErrorInvalidValue;
}

}  // namespace qt
