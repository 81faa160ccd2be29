// qt_mcwf.hip -- MCWF kernels for gfx950 (CDNA4).
//
// Batched persistent kernel: one work group ("group") owns one trajectory at a
// time and advances it through every requested stream interval.  A group is
// NW wavefronts (64 lanes each); lane l of the group owns rows
// {l, l+G, l+2G, ...} (ROWS of them, G = 64*NW) of psi and of the six RKCK
// stage derivatives, all held in registers.  The H_nH/i row coefficients of
// those rows are loaded into registers once per group (the system is shared by
// every trajectory).  A derivative evaluation stages the RK stage vector
// through LDS and gathers each row's ELL neighbours from it; norms, the error
// norm and the jump rates are group reductions (wavefront butterflies, then
// LDS across the NW waves).  All step-size control is FP64 and follows
// ode_engine::Base::step + Boost.Odeint's controlled RKCK54 (CPPQEDutils/ODE.h
// :234-262) and MCWF_Trajectory::step (MCWF_Trajectory.h:347-374) exactly; with
// F32 = true only the derivative evaluations H_nH/i . x run in FP32 complex.
//
// Work distribution: a global atomic counter hands trajectories to groups, so
// uneven trajectory costs (jumps, step-backs, rejected tries) balance out.

#include "qt_kernels.hpp"
#include "qt_philox.h"

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
__device__ __forceinline__ double2 to_d(double2 a) { return a; }
__device__ __forceinline__ double2 to_d(float2 a) { return make_double2((double)a.x, (double)a.y); }
template <typename T> __device__ __forceinline__ T from_d(double2 a);
template <> __device__ __forceinline__ double2 from_d<double2>(double2 a) { return a; }
template <> __device__ __forceinline__ float2 from_d<float2>(double2 a) { return make_float2((float)a.x, (float)a.y); }
__device__ __forceinline__ double2 axpy(double s, double2 k, double2 acc) {  // acc + s*k
  acc.x = fma(s, k.x, acc.x); acc.y = fma(s, k.y, acc.y); return acc;
}
__device__ __forceinline__ double cabs2(double2 a) { return a.x * a.x + a.y * a.y; }

template <int NW> struct GroupReduce {
  // sum / max over all lanes of the group; every lane gets the result.
  __device__ static double sum(double v, double* red) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
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
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
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

// ---------------------------------------------------------------------------
// the per-group trajectory engine
// ---------------------------------------------------------------------------

template <int ROWS, int NW, int WH, bool F32>
struct Engine {
  static constexpr int G = 64 * NW;
  // row coefficients live in registers when they fit; otherwise they are read from
  // the (L2-resident, shared by every trajectory) column-major ELL table
  static constexpr bool kRegs = ROWS * WH <= 24;
  static constexpr int RR = kRegs ? ROWS : 1;
  static constexpr int RW = kRegs ? WH : 1;
  using T = typename CT<F32>::T;

  const SysDev& S;
  T* xs;            // LDS stage vector [ROWS*G] in derivative precision
  double2* ps;      // LDS FP64 state copy [ROWS*G] (rates, jumps, observables)
  double* red;      // LDS reduction scratch
  int lane;
  int hcol[RR][RW];
  T hval[RR][RW];

  __device__ __forceinline__ T coef(int w, int row) const {
    if constexpr (F32) return S.hval32[w * S.dim + row];
    else return S.hval[w * S.dim + row];
  }

  __device__ Engine(const SysDev& s, T* xs_, double2* ps_, double* red_) : S(s), xs(xs_), ps(ps_), red(red_) {
    lane = threadIdx.x;
    if constexpr (kRegs) {
#pragma unroll
      for (int r = 0; r < ROWS; ++r) {
        const int row = r * G + lane;
#pragma unroll
        for (int w = 0; w < WH; ++w) {
          if (row < S.dim && w < S.wh) {
            hcol[r][w] = S.hcol[w * S.dim + row];
            hval[r][w] = coef(w, row);
          } else {
            hcol[r][w] = 0;
            hval[r][w] = from_d<T>(make_double2(0., 0.));
          }
        }
      }
    }
  }

  __device__ __forceinline__ bool valid(int r) const { return r * G + lane < S.dim; }

  // k = (H_nH/i) v, with v already staged in xs
  __device__ __forceinline__ void gather(T k[ROWS]) const {
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
      T acc = from_d<T>(make_double2(0., 0.));
      if constexpr (kRegs) {
#pragma unroll
        for (int w = 0; w < WH; ++w) acc = cfma(hval[r][w], xs[hcol[r][w]], acc);
      } else {
        const int row = r * G + lane;
        if (row < S.dim) {
#pragma unroll 4
          for (int w = 0; w < S.wh; ++w) acc = cfma(coef(w, row), xs[S.hcol[w * S.dim + row]], acc);
        }
      }
      k[r] = acc;
    }
  }

  __device__ __forceinline__ void stage(const double2 v[ROWS]) const {
    __syncthreads();  // previous readers of xs are done
#pragma unroll
    for (int r = 0; r < ROWS; ++r) if (valid(r)) xs[r * G + lane] = from_d<T>(v[r]);
    __syncthreads();
  }

  __device__ __forceinline__ void stagePsi(const double2 v[ROWS]) const {
    __syncthreads();
#pragma unroll
    for (int r = 0; r < ROWS; ++r) if (valid(r)) ps[r * G + lane] = v[r];
    __syncthreads();
  }

  // One RKCK54 try: returns the odeint error norm; xn is the 5th-order proposal.
  // k1 must hold (H/i) x on entry.
  __device__ double tryStep(const double2 x[ROWS], const T k1[ROWS], double dt, double epsAbs, double epsRel,
                            double2 xn[ROWS], long long& derivs) const {
    T k2[ROWS], k3[ROWS], k4[ROWS], k5[ROWS], k6[ROWS];
    double2 v[ROWS];
    const double s21 = CK::a21 * dt;
#pragma unroll
    for (int r = 0; r < ROWS; ++r) v[r] = axpy(s21, to_d(k1[r]), x[r]);
    stage(v); gather(k2);
    { const double q1 = CK::a31 * dt, q2 = CK::a32 * dt;
#pragma unroll
      for (int r = 0; r < ROWS; ++r) v[r] = axpy(q2, to_d(k2[r]), axpy(q1, to_d(k1[r]), x[r])); }
    stage(v); gather(k3);
    { const double q1 = CK::a41 * dt, q2 = CK::a42 * dt, q3 = CK::a43 * dt;
#pragma unroll
      for (int r = 0; r < ROWS; ++r) v[r] = axpy(q3, to_d(k3[r]), axpy(q2, to_d(k2[r]), axpy(q1, to_d(k1[r]), x[r]))); }
    stage(v); gather(k4);
    { const double q1 = CK::a51 * dt, q2 = CK::a52 * dt, q3 = CK::a53 * dt, q4 = CK::a54 * dt;
#pragma unroll
      for (int r = 0; r < ROWS; ++r)
        v[r] = axpy(q4, to_d(k4[r]), axpy(q3, to_d(k3[r]), axpy(q2, to_d(k2[r]), axpy(q1, to_d(k1[r]), x[r])))); }
    stage(v); gather(k5);
    { const double q1 = CK::a61 * dt, q2 = CK::a62 * dt, q3 = CK::a63 * dt, q4 = CK::a64 * dt, q5 = CK::a65 * dt;
#pragma unroll
      for (int r = 0; r < ROWS; ++r)
        v[r] = axpy(q5, to_d(k5[r]), axpy(q4, to_d(k4[r]), axpy(q3, to_d(k3[r]), axpy(q2, to_d(k2[r]), axpy(q1, to_d(k1[r]), x[r]))))); }
    stage(v); gather(k6);
    derivs += 5;
    double e = 0.;
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
          e = fmax(e, ae / (epsAbs + epsRel * (ax + dt * ak)));
        }
      }
    }
    return GroupReduce<NW>::max(e, red);
  }

  __device__ __forceinline__ void derivative(const double2 x[ROWS], T k[ROWS]) const { stage(x); gather(k); }

  // y = J_m psi for the rows of this lane (psi staged in ps)
  __device__ __forceinline__ double2 jumpRow(int m, int r) const {
    const int row = r * G + lane;
    double2 acc = make_double2(0., 0.);
    const int base = m * S.wj * S.dim + row;
    for (int w = 0; w < S.wj; ++w) acc = cfma(S.jval[base + w * S.dim], ps[S.jcol[base + w * S.dim]], acc);
    return acc;
  }

  __device__ void rates(double rt[kMaxJumps]) const {
    for (int m = 0; m < S.nj; ++m) {
      double s = 0.;
#pragma unroll
      for (int r = 0; r < ROWS; ++r) if (valid(r)) s += cabs2(jumpRow(m, r));
      rt[m] = GroupReduce<NW>::sum(s, red);
    }
  }

  __device__ double observable(int q, const double2 x[ROWS]) const {
    double s = 0.;
    const bool im = S.oimag[q] != 0;
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
      if (!valid(r)) continue;
      const int row = r * G + lane;
      double2 acc = make_double2(0., 0.);
      const int base = q * S.wo * S.dim + row;
      for (int w = 0; w < S.wo; ++w) acc = cfma(S.oval[base + w * S.dim], ps[S.ocol[base + w * S.dim]], acc);
      // conj(x) * acc
      s += im ? (x[r].x * acc.y - x[r].y * acc.x) : (x[r].x * acc.x + x[r].y * acc.y);
    }
    return GroupReduce<NW>::sum(s, red);
  }
};

// ---------------------------------------------------------------------------
// batched persistent MCWF kernel
// ---------------------------------------------------------------------------

template <int ROWS, int NW, int WH, bool F32>
__global__ void __launch_bounds__(64 * NW) mcwf_batch_kernel(SysDev S, EnsDev E) {
  using Eng = Engine<ROWS, NW, WH, F32>;
  using T = typename Eng::T;
  constexpr int G = Eng::G;
  __shared__ T xs[ROWS * G];
  __shared__ double2 ps[ROWS * G];
  // psi at the start of the MCWF step (step-back restore): LDS when it fits, else registers
  static constexpr bool kPcLds = (size_t)ROWS * G * (sizeof(T) + 2 * sizeof(double2)) <= 150 * 1024;
  __shared__ double2 pc[kPcLds ? ROWS * G : 1];
  double2 pcr[kPcLds ? 1 : ROWS];
  __shared__ double red[NW > 1 ? NW : 1];
  __shared__ double rts[kMaxJumps];        // jump rates of the current step
  __shared__ long long cnt[kCounters];     // this trajectory's counters (lane 0 updates)
  __shared__ int shTraj;

  Eng eng(S, xs, ps, red);
  const int lane = threadIdx.x;
  auto bump = [&](int c, long long v) { if (lane == 0) cnt[c] += v; };

  for (;;) {
    __syncthreads();
    if (lane == 0) shTraj = atomicAdd(E.work, 1);
    __syncthreads();
    const int traj = shTraj;
    if (traj >= E.n_traj) return;

    double2 x[ROWS], xn[ROWS];
    T k1[ROWS];
#pragma unroll
    for (int r = 0; r < ROWS; ++r) x[r] = eng.valid(r) ? E.psi[(size_t)traj * S.dim + r * G + lane] : make_double2(0., 0.);
    double t = E.t[traj], dtDid = E.dtDid[traj], dtTry = E.dtTry[traj];
    unsigned long long ord = E.ordinal[traj];
    if (lane < kCounters) cnt[lane] = E.counters[(size_t)traj * kCounters + lane];
    __syncthreads();
    const unsigned long long key = E.traj0 + (unsigned long long)traj;
    const int nj = S.nj;

    // coherent part of one MCWF step (MCWF_Trajectory.h:250-276 with ODE.h:234-262)
    auto coherent = [&](double Dt) {
      const double timeBefore = t;
      const double nextDtTry = (fabs(Dt) < fabs(dtTry / 10.) ? dtTry : 0.);
      if ((Dt > 0.) != (dtTry > 0.)) dtTry = -dtDid;
      if (fabs(dtTry) > fabs(Dt)) dtTry = Dt;
      eng.derivative(x, k1);
      long long derivs = 1, tries = 0, failed = 0;
      for (;;) {
        ++tries;
        const double err = eng.tryStep(x, k1, dtTry, E.epsAbs, E.epsRel, xn, derivs);
        if (err > 1.0) {
          dtTry *= fmax(9. / 10. * pow(err, -1. / (4 - 1)), 1. / 5.);
          ++failed;
          continue;
        }
        t += dtTry;
        if (err < 0.5) {
          const double e = fmax(pow(5.0, -5.0), err);
          dtTry *= 9. / 10. * pow(e, -1. / 5);
        }
        break;
      }
#pragma unroll
      for (int r = 0; r < ROWS; ++r) x[r] = xn[r];
      dtDid = t - timeBefore;
      if (nextDtTry != 0.) dtTry = nextDtTry;
      if (lane == 0) { cnt[kOdeSteps] += 1; cnt[kTries] += tries; cnt[kFailed] += failed; cnt[kDerivs] += derivs; }
      // renorm (StateVector.h:148-153)
      double s = 0.;
#pragma unroll
      for (int r = 0; r < ROWS; ++r) if (eng.valid(r)) s += cabs2(x[r]);
      const double nrm = sqrt(GroupReduce<NW>::sum(s, red));
#pragma unroll
      for (int r = 0; r < ROWS; ++r) { x[r].x /= nrm; x[r].y /= nrm; }
    };

    auto computeRates = [&]() {
      eng.stagePsi(x);
      for (int m = 0; m < nj; ++m) {
        double sacc = 0.;
#pragma unroll
        for (int r = 0; r < ROWS; ++r) if (eng.valid(r)) sacc += cabs2(eng.jumpRow(m, r));
        const double v = GroupReduce<NW>::sum(sacc, red);
        if (lane == 0) rts[m] = v;
      }
      __syncthreads();
    };

    for (int k = 0; k < E.n_intervals; ++k) {
      const double endTime = t + E.intervals[k];
      for (;;) {
        const double Dt = endTime - t;
        if (Dt == 0.) break;
        // MCWF_Trajectory::step (MCWF_Trajectory.h:347-374)
#pragma unroll
        for (int r = 0; r < ROWS; ++r) {
          if constexpr (kPcLds) { if (eng.valid(r)) pc[r * G + lane] = x[r]; }
          else pcr[r] = x[r];
        }
        const double tCache = t, dtDidCache = dtDid;
        coherent(Dt);
        if (nj > 0) {
          computeRates();
          for (;;) {  // manageTimeStep (MCWF_Trajectory.h:294-317)
            double totalRate = 0.;
            for (int m = 0; m < nj; ++m) totalRate += rts[m];
            const double sugg = E.dpLimit / totalRate;
            if (totalRate * dtDid > E.overshootTol * E.dpLimit) {
              t = tCache; dtDid = dtDidCache; dtTry = sugg;
              bump(kStepBacks, 1);
#pragma unroll
              for (int r = 0; r < ROWS; ++r) {
                if constexpr (kPcLds) x[r] = eng.valid(r) ? pc[r * G + lane] : make_double2(0., 0.);
                else x[r] = pcr[r];
              }
              coherent(Dt);
              computeRates();
              continue;
            } else if (totalRate * dtTry > E.dpLimit) bump(kOvershoots, 1);
            dtTry = fmin(dtTry, sugg);
            break;
          }
          // performJump (MCWF_Trajectory.h:320-344)
          double random = jump_uniform(E.seed, key, ord++) / dtDid;
          int no = 0;
          for (; random > 0. && no != nj; random -= rts[no++]) ;
          if (random < 0.) {
            --no;
            const double nf = sqrt(rts[no]);
            if (!isfinite(nf) && lane == 0) atomicExch(E.err_flag, 1);
#pragma unroll
            for (int r = 0; r < ROWS; ++r) {
              if (!eng.valid(r)) continue;
              const double2 y = eng.jumpRow(no, r);
              x[r] = make_double2(y.x / nf, y.y / nf);
            }
            if (lane == 0) {
              const long long nJ = cnt[kJumps];
              if (nJ < E.jcap) {
                E.jt[(size_t)traj * E.jcap + nJ] = t;
                E.jch[(size_t)traj * E.jcap + nJ] = no;
              }
              cnt[kJumps] = nJ + 1;
            }
          }
        }
        bump(kMcwfSteps, 1);
      }
      // stream point: raw linear averages of the normalised state
      eng.stagePsi(x);
      for (int q = 0; q < S.no; ++q) {
        const double v = eng.observable(q, x);
        if (lane == 0) E.obs_out[((size_t)traj * E.n_intervals + k) * S.no + q] = v;
      }
      if (lane == 0) E.dtdid_out[(size_t)traj * E.n_intervals + k] = dtDid;
    }

#pragma unroll
    for (int r = 0; r < ROWS; ++r) if (eng.valid(r)) E.psi[(size_t)traj * S.dim + r * G + lane] = x[r];
    __syncthreads();
    if (lane == 0) { E.t[traj] = t; E.dtDid[traj] = dtDid; E.dtTry[traj] = dtTry; E.ordinal[traj] = ord; }
    if (lane < kCounters) E.counters[(size_t)traj * kCounters + lane] = cnt[lane];
  }
}

// ensemble mean over trajectories of the streamed raw averages (deterministic order):
// out[k][q] = (1/N) sum_traj obs[traj][k][q], out_dt[k] = (1/N) sum dtDid
__global__ void ensemble_mean_kernel(const double* obs, const double* dtd, int n_traj, int n_int, int no,
                                     double* out, double* out_dt) {
  const int k = blockIdx.x;
  const int q = blockIdx.y;  // q == no: dtDid
  __shared__ double red[256];
  double s = 0.;
  for (int i = threadIdx.x; i < n_traj; i += blockDim.x)
    s += q < no ? obs[((size_t)i * n_int + k) * no + q] : dtd[(size_t)i * n_int + k];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (q < no) out[k * no + q] = red[0] / (double)n_traj;
    else out_dt[k] = red[0] / (double)n_traj;
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
  Eng eng(S, xs, ps, red);
  double2 x[ROWS], xn[ROWS], k1[ROWS];
#pragma unroll
  for (int r = 0; r < ROWS; ++r) x[r] = eng.valid(r) ? P.x[r * G + threadIdx.x] : make_double2(0., 0.);
  long long derivs = 0;
  eng.derivative(x, k1);  // controlled_runge_kutta::try_step_v1 recomputes dxdt on every try
  const double err = eng.tryStep(x, k1, P.dt, P.epsAbs, P.epsRel, xn, derivs);
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
  Eng eng(S, xs, ps, red);
  double2 x[ROWS];
  double s = 0.;
#pragma unroll
  for (int r = 0; r < ROWS; ++r) { x[r] = eng.valid(r) ? psi[r * G + threadIdx.x] : make_double2(0., 0.); s += cabs2(x[r]); }
  const double nrm = sqrt(GroupReduce<NW>::sum(s, red));
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
  Eng eng(S, xs, ps, red);
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
  Eng eng(S, xs, ps, red);
  double2 x[ROWS];
#pragma unroll
  for (int r = 0; r < ROWS; ++r) x[r] = eng.valid(r) ? psi[r * G + threadIdx.x] : make_double2(0., 0.);
  eng.stagePsi(x);
  for (int q = 0; q < S.no; ++q) { double v = eng.observable(q, x); if (threadIdx.x == 0) out[q] = v; }
}

// dpsidt += (H_nH/i) psi for a batch of state vectors, one group per vector (qt_apply_heff)
template <int ROWS, int NW, int WH, bool F32>
__global__ void __launch_bounds__(64 * NW) apply_heff_kernel(SysDev S, const double2* psi, double2* dpsidt, int n) {
  using Eng = Engine<ROWS, NW, WH, F32>;
  using T = typename Eng::T;
  constexpr int G = Eng::G;
  __shared__ T xs[ROWS * G];
  __shared__ double2 ps[1];
  __shared__ double red[NW > 1 ? NW : 1];
  Eng eng(S, xs, ps, red);
  for (int v = blockIdx.x; v < n; v += gridDim.x) {
    double2 x[ROWS];
    T k[ROWS];
#pragma unroll
    for (int r = 0; r < ROWS; ++r) x[r] = eng.valid(r) ? psi[(size_t)v * S.dim + r * G + threadIdx.x] : make_double2(0., 0.);
    eng.derivative(x, k);
#pragma unroll
    for (int r = 0; r < ROWS; ++r)
      if (eng.valid(r)) {
        double2& o = dpsidt[(size_t)v * S.dim + r * G + threadIdx.x];
        const double2 kd = to_d(k[r]);
        o.x += kd.x; o.y += kd.y;
      }
  }
}

// ---------------------------------------------------------------------------
// launch table
// ---------------------------------------------------------------------------

struct Layout { int rows, nw, wh; };

// Every (rows, waves, ELL width) combination compiled into the library.
#define QT_LAYOUTS(X) \
  X(1, 1, 4) X(1, 1, 8) X(1, 1, 16) \
  X(2, 1, 4) X(2, 1, 8) X(2, 1, 16) \
  X(4, 1, 8) X(4, 1, 16) \
  X(1, 4, 8) X(1, 4, 16) X(2, 4, 8) X(2, 4, 16) X(4, 4, 16) \
  X(4, 8, 16) X(2, 16, 16) X(4, 16, 16)

bool layout_supported(int rows, int nw, int wh) {
#define QT_CASE(R, N, W) if (rows == R && nw == N && wh == W) return true;
  QT_LAYOUTS(QT_CASE)
#undef QT_CASE
  return false;
}

hipError_t launch_batch(int rows, int nw, int wh, bool f32, int grid, const SysDev& S, const EnsDev& E, hipStream_t st) {
#define QT_CASE(R, N, W)                                                                         \
  if (rows == R && nw == N && wh == W) {                                                         \
    if (f32) hipLaunchKernelGGL((mcwf_batch_kernel<R, N, W, true>), dim3(grid), dim3(64 * N), 0, st, S, E);  \
    else hipLaunchKernelGGL((mcwf_batch_kernel<R, N, W, false>), dim3(grid), dim3(64 * N), 0, st, S, E);    \
    return hipGetLastError();                                                                    \
  }
  QT_LAYOUTS(QT_CASE)
#undef QT_CASE
  return hipErrorInvalidValue;
}

hipError_t launch_ensemble_mean(const double* obs, const double* dtd, int n_traj, int n_int, int no, double* out,
                                double* out_dt, hipStream_t st) {
  hipLaunchKernelGGL(ensemble_mean_kernel, dim3(n_int, no + 1), dim3(256), 0, st, obs, dtd, n_traj, n_int, no, out, out_dt);
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

hipError_t launch_apply_heff(int rows, int nw, int wh, bool f32, const SysDev& S, const double2* psi, double2* dpsidt, int n,
                             hipStream_t st) {
  const int grid = n < 4096 ? n : 4096;
#define QT_CASE(R, N, W)                                                                                      \
  if (rows == R && nw == N && wh == W) {                                                                      \
    if (f32) hipLaunchKernelGGL((apply_heff_kernel<R, N, W, true>), dim3(grid), dim3(64 * N), 0, st, S, psi, dpsidt, n); \
    else hipLaunchKernelGGL((apply_heff_kernel<R, N, W, false>), dim3(grid), dim3(64 * N), 0, st, S, psi, dpsidt, n);    \
    return hipGetLastError();                                                                                 \
  }
  QT_LAYOUTS(QT_CASE)
#undef QT_CASE
  return hipErrorInvalidValue;
}

}  // namespace qt
