// qt_master_sparse.hip -- batched Master-equation sweep beyond the MFMA kernel's dimensions (gfx950).
//
// The MFMA kernel stages a point's whole stage matrix in LDS, which caps it at d = 96.  This kernel
// keeps every matrix of a point in HBM -- in the MFMA kernel's tile order, so uploads, state reads
// and checkpoints are the same for both -- and applies the Master superoperator from the sparse
// (ELL) rows of H_nH/i and of the jump operators, term by term as Master::step's derivative does
// (Master.h:157-184):
//     drho = 2 Re[(H_nH/i) S]  +  sum_m J_m S^dagger J_m^dagger,
//     drho_ij = sum_w h_iw S_{c_iw, j} + conj(sum_w h_jw S_{c_jw, i})
//             + sum_m sum_ab J_ia conj(S_{c_jb, c_ia}) conj(J_jb),
// where 2 Re[A] = A + A^dagger (calculateTwoTimesRealPartOfSelf, CMatrix.cc:14-30) and the jump term
// is actWithJ on S's columns, the Hermitian conjugate, actWithJ again.  In the interaction picture an
// entry of H_nH/i at picture time tau is h exp(f tau) with the entry's own frequency f (the
// tables' frequency index; the Exact frees' rotation E H E^-1 of the MFMA kernel is the special
// case f = exz_r - exz_c), and each step ends with rho <- U rho U^dagger (Master.h:186-193).
//
// One work group of 16 wavefronts per sweep point.  Thread `tid` owns elements tid + 1024 n of
// the point's matrices, so the element-wise streams (stage sums, the error norm, the step's end)
// are coalesced; the gathers read the stage matrix from L2.  A step's six derivatives alternate
// between two stage buffers, so one barrier per stage separates the writers of the next stage
// matrix from the readers of the last.  HBM per point: rho, K1..K6, the 5th-order proposal and
// the two stage buffers, nine d^2 complex FP64 matrices (d padded to a multiple of 16).

#include "qt_master_common.hpp"

namespace qt {

constexpr int kSparseThreads = 1024;
constexpr int kSparseMaxDim = 4096;  // the interaction picture's U table lives in LDS

// index of element (i, j) in tile order (tile t = (i / 16) NB + j / 16, index (4 t + v) 64 + l
// for row 16 (i / 16) + 4 v + l / 16, column 16 (j / 16) + l % 16)
__device__ __forceinline__ int tile_index(int i, int j, int NB) {
  return ((4 * ((i >> 4) * NB + (j >> 4)) + ((i & 15) >> 2)) << 6) + ((i & 3) << 4) + (j & 15);
}

template <bool TD>
__global__ void __launch_bounds__(kSparseThreads) master_sparse_kernel(MasterDev M) {
  constexpr int NW = kSparseThreads / 64;
  const int p = blockIdx.x;
  if (p >= M.n_points) return;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int d = M.dim, NB = (d + 15) >> 4, NN = 256 * NB * NB;
  __shared__ double red[NW];
  __shared__ long long cnt[kCounters];
  __shared__ double2 fph[kMaxFreqs];             // exp(f tau) of the point's frequencies, current stage
  __shared__ double2 u[TD ? kSparseMaxDim : 1];  // exp(-exz (t - t0)) at a step's end
  __shared__ int asStop;

  double2* const X = M.x + (size_t)p * NN;
  double2* const K = M.k + (size_t)p * 9 * NN;   // K1..K6, the proposal, two stage buffers
  double2* const XN = K + 6 * (size_t)NN;
  const int wr = M.wr, nj = M.nj, wj = M.wj;
  const size_t hb = (size_t)p * wr * d;
  const int* const hc = M.rcol + hb;
  const double2* const hv = M.rval + hb;
  const int* const hf = M.rfi + hb;
  const int* const jc = M.jcol + (size_t)p * nj * wj * d;
  const double2* const jv = M.jval + (size_t)p * nj * wj * d;

  auto bsum = [&](double v) {
    v = wave_sum(v);
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    double s = red[0];
#pragma unroll
    for (int i = 1; i < NW; ++i) s += red[i];
    return s;
  };
  auto bmax = [&](double v) {
    v = wave_max(v);
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    double s = red[0];
#pragma unroll
    for (int i = 1; i < NW; ++i) s = fmax(s, red[i]);
    return s;
  };
  // f(e, i, j) for this thread's elements inside the d x d block (the padding is never touched)
  auto forElems = [&](auto&& f) {
    for (int e = tid; e < NN; e += kSparseThreads) {
      const int t = e >> 8, v = (e >> 6) & 3, l = e & 63;
      const int i = 16 * (t / NB) + 4 * v + (l >> 4), j = 16 * (t % NB) + (l & 15);
      if (i < d && j < d) f(e, i, j);
    }
  };
  auto coef = [&](int o) {  // entry o of the H_nH/i table at the current stage's picture time
    double2 h = hv[o];
    if constexpr (TD) h = cmul(h, fph[hf[o]]);
    return h;
  };
  // drho_ij for the stage matrix Sm
  auto superop = [&](const double2* Sm, int i, int j) {
    double2 a = make_double2(0., 0.), b = make_double2(0., 0.);
    for (int q = 0; q < wr; ++q) {
      const int oi = q * d + i, oj = q * d + j;
      a = cfma(coef(oi), Sm[tile_index(hc[oi], j, NB)], a);
      b = cfma(coef(oj), Sm[tile_index(hc[oj], i, NB)], b);
    }
    double2 r = make_double2(a.x + b.x, a.y - b.y);
    for (int m = 0; m < nj; ++m) {
      const int base = m * wj * d;
      for (int x = 0; x < wj; ++x) {
        const double2 ja = jv[base + x * d + i];
        if (ja.x == 0. && ja.y == 0.) continue;
        const int ca = jc[base + x * d + i];
        for (int y = 0; y < wj; ++y) {
          const double2 jb = jv[base + y * d + j];
          if (jb.x == 0. && jb.y == 0.) continue;
          const double2 q = cmul(Sm[tile_index(jc[base + y * d + j], ca, NB)], jb);  // S_{cb,ca} J_jb
          r = cfma(ja, make_double2(q.x, -q.y), r);
        }
      }
    }
    return r;
  };
  // Kout = the superoperator on Sm at picture time tau (Sm's writers are past the barrier)
  auto derivative = [&](const double2* Sm, double2* Kout, double tau) {
    __syncthreads();
    if constexpr (TD) {
      if (tid < M.nrf) {
        const double2 f = M.rfreq[(size_t)p * M.nrf + tid];
        double sn, cs;
        sincos(f.y * tau, &sn, &cs);
        const double e = exp(f.x * tau);
        fph[tid] = make_double2(e * cs, e * sn);
      }
      __syncthreads();
    } else {
      (void)tau;
    }
    forElems([&](int e, int i, int j) { Kout[e] = superop(Sm, i, j); });
  };

  double t = M.t[p], dtDid = M.dtDid[p], dtTry = M.dtTry[p];
  if (tid < kCounters) cnt[tid] = M.counters[(size_t)p * kCounters + tid];
  __syncthreads();
  int done = M.done[p];
  const double2* const exzP = TD ? M.exz + (size_t)p * d : nullptr;

  for (int kI = 0; kI < M.n_intervals; ++kI) {
    // Dt mode: advance(intervals[kI]); dc mode: up to dc steps step(t_end - t) per stream point
    const double endTime = M.dc > 0 ? M.t_end : t + M.intervals[kI];
    const int maxSteps = M.dc > 0 ? M.dc : -1;
    if (!done) {
      for (int steps = 0;; ++steps) {
        const double Dt = endTime - t;
        if (Dt == 0. || steps == maxSteps) break;
        // ode_engine::Base::step(Dt) (ODE.h:234-262) with the pictures coinciding at the step's start
        const double timeBefore = t;
        const double t0 = t;
        const double nextDtTry = (fabs(Dt) < fabs(dtTry / 10.) ? dtTry : 0.);
        if ((Dt > 0.) != (dtTry > 0.)) dtTry = -dtDid;
        if (fabs(dtTry) > fabs(Dt)) dtTry = Dt;
        derivative(X, K, t - t0);  // k1 at (x, t); unchanged on retries
        long long tries = 0, failed = 0;
        for (;;) {
          ++tries;
          const double dt = dtTry;
#pragma unroll 1
          for (int st = 1; st < 6; ++st) {
            double2* const Sb = K + (size_t)(7 + (st & 1)) * NN;
            forElems([&](int e, int, int) {
              double2 acc = X[e];
              for (int q = 0; q < st; ++q) acc = axpy(ck_a(st, q) * dt, K[(size_t)q * NN + e], acc);
              Sb[e] = acc;
            });
            derivative(Sb, K + (size_t)st * NN, (t + ck_c(st) * dt) - t0);
          }
          double emax = 0.;
          forElems([&](int e, int, int) {
            const double2 x = X[e], k1 = K[e], k3 = K[2 * (size_t)NN + e], k4 = K[3 * (size_t)NN + e],
                          k5 = K[4 * (size_t)NN + e], k6 = K[5 * (size_t)NN + e];
            XN[e] = axpy(CK::b6 * dt, k6, axpy(CK::b4 * dt, k4, axpy(CK::b3 * dt, k3, axpy(CK::b1 * dt, k1, x))));
            double2 xe = make_double2(CK::e1 * dt * k1.x, CK::e1 * dt * k1.y);
            xe = axpy(CK::e6 * dt, k6, axpy(CK::e5 * dt, k5, axpy(CK::e4 * dt, k4, axpy(CK::e3 * dt, k3, xe))));
            const double D = M.epsAbs + M.epsRel * (sqrt(cabs2(x)) + dt * sqrt(cabs2(k1)));
            emax = fmax(emax, cabs2(xe) / (D * D));
          });
          const double err = sqrt(bmax(emax));
          if (err > 1.0) {
            dtTry *= fmax(9. / 10. * inv_root<3>(err), 1. / 5.);
            ++failed;
            continue;
          }
          t += dtTry;
          if (err < 0.5) dtTry *= 9. / 10. * inv_root<5>(fmax(0.00032, err));
          break;
        }
        dtDid = t - timeBefore;
        if (nextDtTry != 0.) dtTry = nextDtTry;
        // Exact part (Master.h:186-193): rho_ij <- u_i rho_ij conj(u_j), u = exp(-exz (t - t0)); then
        // calculateTwoTimesRealPartOfSelf (CMatrix.cc:14-30) and the trace renormalisation
        if constexpr (TD) {
          for (int i = tid; i < d; i += kSparseThreads) {
            const double2 z = exzP[i];
            double sn, cs;
            sincos(-z.y * (t - t0), &sn, &cs);
            const double e = exp(-z.x * (t - t0));
            u[i] = make_double2(e * cs, e * sn);
          }
          __syncthreads();
        }
        auto ui = [&](int i) {
          if constexpr (TD) return u[i];
          else { (void)i; return make_double2(1., 0.); }
        };
        double tr = 0.;
        for (int i = tid; i < d; i += kSparseThreads) tr += 2. * cabs2(ui(i)) * XN[tile_index(i, i, NB)].x;
        const double inv = 1.0 / bsum(tr);
        forElems([&](int e, int i, int j) {
          const double2 a = ui(i), b = ui(j);
          const double2 yij = cmul(cmul(a, XN[e]), make_double2(b.x, -b.y));
          const double2 yji = cmul(cmul(b, XN[tile_index(j, i, NB)]), make_double2(a.x, -a.y));
          X[e] = make_double2((yij.x + yji.x) * inv, (yij.y - yji.y) * inv);
        });
        if (tid == 0) {
          cnt[kMcwfSteps] += 1; cnt[kOdeSteps] += 1; cnt[kTries] += tries; cnt[kFailed] += failed;
          cnt[kDerivs] += 6 * tries;
        }
      }
    }
    // stream point: averages Tr(rho O_q), ||drho/dt||_F, steady-state check
    __syncthreads();
    for (int q = 0; q < M.no; ++q) {
      double part = 0.;
      const bool im = M.oimag[q] != 0;
      for (int i = tid; i < d; i += kSparseThreads) {
        for (int a = 0; a < M.owid[q]; ++a) {
          const double2 o = M.oval[((size_t)q * M.wo + a) * d + i];
          const int c = M.ocol[((size_t)q * M.wo + a) * d + i];
          const double2 r = X[tile_index(c, i, NB)];  // Tr(rho O) = sum_i sum_a O_{i,c} rho_{c,i}
          part += im ? (o.x * r.y + o.y * r.x) : (o.x * r.x - o.y * r.y);
        }
      }
      const double v = bsum(part);
      if (tid == 0) M.obs_out[((size_t)p * M.n_intervals + kI) * M.no + q] = v;
    }
    derivative(X, K, 0.);  // t0 = t at a stream point
    double n2 = 0.;
    forElems([&](int e, int, int) { n2 += cabs2(K[e]); });
    const double dn = sqrt(bsum(n2));
    if (tid == 0) {
      if (M.time_out) M.time_out[(size_t)p * M.n_intervals + kI] = t;
      M.dtdid_out[(size_t)p * M.n_intervals + kI] = dtDid;
      M.dnorm_out[(size_t)p * M.n_intervals + kI] = dn;
    }
    if (!done && M.ss_tol > 0. && dn < M.ss_tol) {
      done = 1;
      if (tid == 0) { M.done[p] = 1; M.tss[p] = t; }
    }
    if (!done && M.as_rep > 0) {
      if (tid == 0) asStop = autostop(M, p, M.obs_out + ((size_t)p * M.n_intervals + kI) * M.no);
      __syncthreads();
      if (asStop) {
        done = 1;
        if (tid == 0) { M.done[p] = 1; M.tss[p] = t; }
      }
      __syncthreads();
    }
  }
  __syncthreads();
  if (tid == 0) { M.t[p] = t; M.dtDid[p] = dtDid; M.dtTry[p] = dtTry; }
  if (tid < kCounters) M.counters[(size_t)p * kCounters + tid] = cnt[tid];
}

bool master_sparse_supported(int d) { return d >= 1 && d <= kSparseMaxDim; }

hipError_t launch_master_sparse(const MasterDev& M, hipStream_t st) {
  if (!master_sparse_supported(M.dim)) return hipErrorInvalidValue;
  if (M.exz) hipLaunchKernelGGL((master_sparse_kernel<true>), dim3(M.n_points), dim3(kSparseThreads), 0, st, M);
  else hipLaunchKernelGGL((master_sparse_kernel<false>), dim3(M.n_points), dim3(kSparseThreads), 0, st, M);
  return hipGetLastError();
}

}  // namespace qt
