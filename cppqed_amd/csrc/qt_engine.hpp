// qt_engine.hpp -- MCWF trajectory engines and persistent kernels (templates) for gfx950 (CDNA4).
//
// Included by the translation units that instantiate them (qt_batch*.hip, qt_mixed.hip,
// qt_aux.hip), which are compiled in parallel.
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

#pragma once

#include "qt_device.hpp"
#include "qt_kernels.hpp"
#include "qt_philox.h"

#include <algorithm>

namespace qt {

// ---------------------------------------------------------------------------
// the per-group trajectory engine
// ---------------------------------------------------------------------------

// A team of NW wavefronts is a whole work group (its hand-overs use the work-group barrier) or,
// for NW = 1, one wavefront.  TD = interaction-picture system (H entries
// rotate with the picture, psi gets the Exact U after each ODE step); the Schroedinger-picture
// instantiations carry none of that code.  OPT: optional code paths, compiled only where they
// pay (each adds registers to a kernel): kOffsets = gathers use the constant-offset form of the
// H table when the system has one; kTwoStage = two stage buffers (multi-wavefront teams).
enum { kOffsets = 1, kTwoStage = 2 };
template <int ROWS, int NW, int WH, bool F32, bool TD = false, int OPT = 0>
struct Engine {
  static constexpr int kRows = ROWS, kNW = NW;
  static constexpr bool kTD = TD;
  static constexpr int G = 64 * NW;
  // row coefficients live in registers when they fit; otherwise they are read from
  // the (L2-resident, shared by every trajectory) column-major ELL table
  static constexpr bool kRegs = ROWS * WH <= 24;
  static constexpr int RR = kRegs ? ROWS : 1;
  static constexpr int RW = kRegs ? WH : 1;
  using T = typename CT<F32>::T;

  const SysDev& S;
  mutable T* xs;    // LDS stage vector [ROWS*G] in derivative precision (the one being read)
  T* xsA;           // its storage
  T* xsB = nullptr; // a second stage buffer (multi-wavefront teams, when LDS allows): stages
                    // alternate between the two and a hand-over needs one barrier, not two
  double2* ps;      // LDS FP64 state copy [ROWS*G] (rates, jumps, observables)
  double* red;      // LDS reduction scratch
  const int lane;
  int hcol[RR][RW];
  T hval[RR][RW];
  // scalar copies of the H table's base addresses, width and dimension
  const int* hcolU;
  const T* hvalU;
  const int* hoffU;  // constant-offset form of the table (SysDev::hoff), or null
  int whU, dimU;
  // dictionary form of the table (SysDev::hdict): the team's LDS copies of the dictionary and of the
  // slots' bases (loadDict), and the rows' index words; null without one
  const T* dictL = nullptr;
  const unsigned char* idxL = nullptr;
  const int* dbL = nullptr;
  // the slots' column offsets (LDS copy of SysDev::hoff, with a dictionary) and whether the
  // slot-major gather applies: constant offsets, the state fills the team (dim == ROWS * G) and
  // no |offset| exceeds G (gatherC)
  const int* doffL = nullptr;
  bool offC = false;
  // interaction picture: the H entry (w, row) at picture time tau is hval * ph[f] with
  // ph[f] = exp(freq[f] tau), the phasor table in LDS rebuilt for each stage time; the
  // register-cached column carries f in its top byte
  bool td = false;
  T* ph = nullptr;
  const double2* freqU = nullptr;
  const int* hfidxU = nullptr;
  int nfU = 0;

  // unsigned element offsets from a scalar base: loads address as scalar base + 32-bit vector offset
  __device__ __forceinline__ T coef(int w, int row) const { return gld2(hvalU, (unsigned)(w * dimU + row)); }
  __device__ __forceinline__ int col(int w, int row) const { return gbl(hcolU)[(unsigned)(w * dimU + row)]; }

  // lane_ = index of this thread within its team of NW wavefronts; ph_ = LDS phasor table
  // [kMaxFreqs] (needed for interaction-picture systems only)
  __device__ __forceinline__ Engine(const SysDev& s, T* xs_, double2* ps_, double* red_, int lane_, T* ph_ = nullptr)
      : S(s), xs(xs_), xsA(xs_), ps(ps_), red(red_), lane(lane_) {
    hcolU = uniform(s.hcol);
    if constexpr (kRegs || !(OPT & kOffsets)) {
      hoffU = nullptr;  // (register-cached columns otherwise)
    } else {
      hoffU = uniform(s.hoff);
      if (hoffU && uniform(s.dim) == ROWS * G) {
        int mx = 0;
        for (int w = 0; w < uniform(s.wh); ++w) mx = max(mx, abs(gbl(hoffU)[w]));
        offC = mx <= G;
      }
    }
    if constexpr (F32) hvalU = uniform(s.hval32);
    else hvalU = uniform(s.hval);
    whU = uniform(s.wh);
    dimU = uniform(s.dim);
    if constexpr (TD) {
      nfU = uniform(s.nfreq);
      td = nfU > 0 && ph_ != nullptr;
      if (td) { ph = ph_; freqU = uniform(s.freq); hfidxU = uniform(s.hfidx); }
    } else {
      (void)ph_;
    }
    if constexpr (kRegs) {
#pragma unroll
      for (int r = 0; r < ROWS; ++r) {
        const int row = r * G + lane;
#pragma unroll
        for (int w = 0; w < WH; ++w) {
          if (row < dimU && w < whU) {
            hcol[r][w] = col(w, row);
            if constexpr (TD) if (td) hcol[r][w] |= hfidxU[(unsigned)(w * dimU + row)] << 24;
            hval[r][w] = coef(w, row);
          } else {
            hcol[r][w] = 0;
            hval[r][w] = from_d<T>(make_double2(0., 0.));
          }
        }
      }
    }
  }

  // Copies the system's coefficient dictionary into the team's LDS buffers -- dict [kMaxDict],
  // base [kMaxDictSlots], idx [idxCap bytes] -- and switches the gathers to it; ends with a team
  // barrier.  Without a dictionary, when its index does not fit idxCap, or for an engine without
  // the offset form, nothing changes.  off [kMaxDictSlots] (may be null): LDS for the slots'
  // column offsets, which the slot-major gather needs.
  __device__ __forceinline__ void loadDict(T* dict, int* base, unsigned char* idx, int idxCap, int* off = nullptr) {
    if constexpr (kRegs || !(OPT & kOffsets)) {
      (void)dict; (void)base; (void)idx; (void)idxCap; (void)off;
    } else {
      const int n = uniform(S.ndict);
      if (!hoffU || n <= 0 || n > kMaxDict || whU * dimU > idxCap) return;
      const T* src;
      if constexpr (F32) src = uniform(S.hdict32);
      else src = uniform(S.hdict);
      for (int i = lane; i < n; i += G) dict[i] = src[i];
      if (lane < kMaxDictSlots) base[lane] = S.hdbase[lane];
      // the index bytes in 4-byte words (whU * dimU is a multiple of 4: dim is, for every
      // layout with a dictionary)
      const unsigned* isrc = reinterpret_cast<const unsigned*>(uniform(S.hidx));
      unsigned* idst = reinterpret_cast<unsigned*>(idx);
      for (int i = lane; i < whU * dimU / 4; i += G) idst[i] = isrc[i];
      dictL = dict; dbL = base; idxL = idx;
      if (off) {
        if (lane < whU) off[lane] = gbl(hoffU)[lane];
        doffL = off;
      }
      sync();
    }
  }

  // FP64 value of a stage derivative at its point of use.  With several rows per lane the FP32
  // values stay live and are converted at each use (half the registers); with one or two rows
  // the compiler may keep converted copies (fewer instructions).
  static __device__ __forceinline__ double2 kd(const T& k) {
    if constexpr (ROWS >= 4) return to_d_use(k);
    else return to_d(k);
  }

  __device__ __forceinline__ bool valid(int r) const { return r * G + lane < dimU; }
  // row block r holds part of the state for some lane of the team (uniform over the team)
  __device__ __forceinline__ bool live(int r) const { return r == 0 || r * G < dimU; }
  __device__ __forceinline__ void sync() const {
    if constexpr (NW == 1) wave_sync();
    else __syncthreads();
  }

  // sum / max over the team's lanes; every lane gets the bitwise-same result
  __device__ __forceinline__ double gsum(double v) const {
    v = wave_sum(v);
    if constexpr (NW == 1) return v;
    else {
      sync();
      if ((lane & 63) == 0) red[lane >> 6] = v;
      sync();
      double s = 0.;
#pragma unroll
      for (int i = 0; i < NW; ++i) s += red[i];
      return s;
    }
  }
  __device__ __forceinline__ double gmax(double v) const {
    v = wave_max(v);
    if constexpr (NW == 1) return v;
    else {
      sync();
      if ((lane & 63) == 0) red[lane >> 6] = v;
      sync();
      double s = red[0];
#pragma unroll
      for (int i = 1; i < NW; ++i) s = fmax(s, red[i]);
      return s;
    }
  }

  // Constant-offset table when the state fills the team and no |offset| exceeds G, slot-major: a
  // slot's offset and coefficient source -- its part of the dictionary and row of index bytes in
  // LDS (DICT), else its column of the global table -- are set up once for all of the lane's
  // rows, and the column row + offset needs the clamp of the general path only in the first row
  // block (it can fall below 0) and the last (beyond dim - 1); in between it stays inside the
  // state.  Each row still sums its slots in order with the same operands, so k is bitwise the
  // general path's.
  template <bool DICT>
  __device__ __forceinline__ void gatherC(T k[ROWS]) const {
    T acc[ROWS];
#pragma unroll
    for (int r = 0; r < ROWS; ++r) acc[r] = from_d<T>(make_double2(0., 0.));
    const int wh = whU, dim = dimU;
#pragma unroll 3
    for (int w = 0; w < wh; ++w) {
      const int off = DICT && doffL ? doffL[w] : gbl(hoffU)[w];
      const T* dw = DICT ? dictL + dbL[w] : nullptr;
      const unsigned char* iw = DICT ? idxL + (unsigned)(w * dim + lane) : nullptr;
      const unsigned cw = (unsigned)(w * dim + lane);
      const T* xw = xs + (lane + off);
#pragma unroll
      for (int r = 0; r < ROWS; ++r) {
        T xv;
        if (r == 0 || r == ROWS - 1) {
          int c = r * G + lane + off;
          if (r == 0) c = max(c, 0);
          if (r == ROWS - 1) c = min(c, dim - 1);
          xv = xs[c];
        } else {
          xv = xw[r * G];
        }
        const T cf = DICT ? dw[iw[r * G]] : gld2(hvalU, cw + (unsigned)(r * G));
        acc[r] = cfma(cf, xv, acc[r]);
      }
    }
#pragma unroll
    for (int r = 0; r < ROWS; ++r) k[r] = acc[r];
  }

  // k = (H_nH/i) v, with v already staged in xs
  __device__ __forceinline__ void gather(T k[ROWS]) const {
    if constexpr (TD) if (td) { gatherTD(k); return; }
    if constexpr (!kRegs && (OPT & kOffsets) != 0)
      if (offC) {
        if (dictL) gatherC<true>(k);
        else gatherC<false>(k);
        return;
      }
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
      T acc = from_d<T>(make_double2(0., 0.));
      if (!live(r)) { k[r] = acc; continue; }
      if constexpr (kRegs) {
#pragma unroll
        for (int w = 0; w < WH; ++w) acc = cfma(hval[r][w], xs[hcol[r][w]], acc);
      } else {
        const int row = r * G + lane;
        if ((OPT & kOffsets) && row < dimU && hoffU && dictL) {
          // dictionary form: the coefficient is the entry of slot w's part of the dictionary named
          // by the row's index byte, both read from the team's LDS copies
#pragma unroll 4
          for (int w = 0; w < whU; ++w) {
            const int c = min(max(row + gbl(hoffU)[w], 0), dimU - 1);
            acc = cfma(dictL[dbL[w] + idxL[(unsigned)(w * dimU + row)]], xs[c], acc);
          }
        } else if ((OPT & kOffsets) && row < dimU && hoffU) {
          // constant offsets: the column is computed, only the coefficient is loaded
#pragma unroll 4
          for (int w = 0; w < whU; ++w) {
            const int c = min(max(row + gbl(hoffU)[w], 0), dimU - 1);
            acc = cfma(coef(w, row), xs[c], acc);
          }
        } else if (row < dimU) {
#pragma unroll 4
          for (int w = 0; w < whU; ++w) acc = cfma(coef(w, row), xs[col(w, row)], acc);
        }
      }
      k[r] = acc;
    }
  }

  // the same at the picture time of the last setTime: entry = hval * ph[f]
  __device__ __forceinline__ void gatherTD(T k[ROWS]) const {
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
      T acc = from_d<T>(make_double2(0., 0.));
      if (!live(r)) { k[r] = acc; continue; }
      if constexpr (kRegs) {
#pragma unroll
        for (int w = 0; w < WH; ++w) {
          const T c = cmul(hval[r][w], ph[(unsigned)hcol[r][w] >> 24]);
          acc = cfma(c, xs[hcol[r][w] & 0xFFFFFF], acc);
        }
      } else {
        const int row = r * G + lane;
        if (row < dimU) {
          for (int w = 0; w < whU; ++w) {
            const unsigned o = (unsigned)(w * dimU + row);
            acc = cfma(cmul(gld2(hvalU, o), ph[gbl(hfidxU)[o]]), xs[gbl(hcolU)[o]], acc);
          }
        }
      }
      k[r] = acc;
    }
  }

  // rebuilds the phasor table for picture time tau = t_stage - t0 (interaction picture only)
  __device__ __forceinline__ void setTime(double tau) const {
    if constexpr (!TD) { (void)tau; return; }
    if (!td) return;
    sync();  // previous readers of the table are done
    if (lane < nfU) {
      const double2 f = freqU[lane];
      const double e = exp(f.x * tau);
      double sn, cs;
      sincos(f.y * tau, &sn, &cs);
      ph[lane] = from_d<T>(make_double2(e * cs, e * sn));
    }
    sync();
  }

  // xs and ps may share storage: both hand-overs start with a barrier, so the readers of the
  // other buffer are done before it is overwritten
  __device__ __forceinline__ void stage(const double2 v[ROWS]) const {
    if constexpr ((OPT & kTwoStage) != 0 && NW > 1) {  // xsB is set with this option
      // the buffer nobody reads now: its last readers (two hand-overs ago) have all passed the
      // barrier of the previous hand-over
      T* nb = xs == xsA ? xsB : xsA;
#pragma unroll
      for (int r = 0; r < ROWS; ++r) if (live(r)) nb[r * G + lane] = from_d<T>(v[r]);
      sync();
      xs = nb;
      return;
    }
    sync();  // previous readers of xs / ps are done
#pragma unroll
    for (int r = 0; r < ROWS; ++r) if (live(r)) xs[r * G + lane] = from_d<T>(v[r]);
    sync();
  }

  __device__ __forceinline__ void stagePsi(const double2 v[ROWS]) const {
    sync();
#pragma unroll
    for (int r = 0; r < ROWS; ++r) if (live(r)) ps[r * G + lane] = v[r];
    sync();
  }

  // One RKCK54 try: returns the SQUARE of the odeint error norm; xn is the 5th-order proposal.
  // k1 must hold (H/i) x on entry.
  // t, t0: time at the start of the try and picture origin (stage times t + c_i dt - t0 set the
  // interaction-picture phasors)
  // GSL: GSL's controller instead (returns GSL's error ratio, not squared; see the end)
  template <bool GSL = false>
  __device__ double tryStep(const double2 x[ROWS], const T k1[ROWS], double dt, double epsAbs, double epsRel,
                            double2 xn[ROWS], long long& derivs, double t = 0., double t0 = 0.) const {
    T k2[ROWS], k3[ROWS], k4[ROWS], k5[ROWS], k6[ROWS];
    double2 xeg[GSL ? ROWS : 1];
    double2 v[ROWS];
    const double s21 = CK::a21 * dt;
#pragma unroll
    for (int r = 0; r < ROWS; ++r) if (live(r)) v[r] = axpy(s21, kd(k1[r]), x[r]);
    stage(v); setTime((t + CK::c2 * dt) - t0); gather(k2);
    { const double q1 = CK::a31 * dt, q2 = CK::a32 * dt;
#pragma unroll
      for (int r = 0; r < ROWS; ++r) if (live(r)) v[r] = axpy(q2, kd(k2[r]), axpy(q1, kd(k1[r]), x[r])); }
    stage(v); setTime((t + CK::c3 * dt) - t0); gather(k3);
    { const double q1 = CK::a41 * dt, q2 = CK::a42 * dt, q3 = CK::a43 * dt;
#pragma unroll
      for (int r = 0; r < ROWS; ++r) if (live(r)) v[r] = axpy(q3, kd(k3[r]), axpy(q2, kd(k2[r]), axpy(q1, kd(k1[r]), x[r]))); }
    stage(v); setTime((t + CK::c4 * dt) - t0); gather(k4);
    { const double q1 = CK::a51 * dt, q2 = CK::a52 * dt, q3 = CK::a53 * dt, q4 = CK::a54 * dt;
#pragma unroll
      for (int r = 0; r < ROWS; ++r)
        if (live(r)) v[r] = axpy(q4, kd(k4[r]), axpy(q3, kd(k3[r]), axpy(q2, kd(k2[r]), axpy(q1, kd(k1[r]), x[r])))); }
    stage(v); setTime((t + CK::c5 * dt) - t0); gather(k5);
    { const double q1 = CK::a61 * dt, q2 = CK::a62 * dt, q3 = CK::a63 * dt, q4 = CK::a64 * dt, q5 = CK::a65 * dt;
#pragma unroll
      for (int r = 0; r < ROWS; ++r)
        if (live(r)) v[r] = axpy(q5, kd(k5[r]), axpy(q4, kd(k4[r]), axpy(q3, kd(k3[r]), axpy(q2, kd(k2[r]), axpy(q1, kd(k1[r]), x[r]))))); }
    stage(v); setTime((t + CK::c6 * dt) - t0); gather(k6);
    derivs += 5;
    double e = 0.;
    {
      const double q1 = CK::b1 * dt, q3 = CK::b3 * dt, q4 = CK::b4 * dt, q6 = CK::b6 * dt;
      const double f1 = CK::e1 * dt, f3 = CK::e3 * dt, f4 = CK::e4 * dt, f5 = CK::e5 * dt, f6 = CK::e6 * dt;
#pragma unroll
      for (int r = 0; r < ROWS; ++r) {
        if (!live(r)) { xn[r] = make_double2(0., 0.); continue; }
        const double2 d1 = kd(k1[r]), d3 = kd(k3[r]), d4 = kd(k4[r]), d5 = kd(k5[r]), d6 = kd(k6[r]);
        xn[r] = axpy(q6, d6, axpy(q4, d4, axpy(q3, d3, axpy(q1, d1, x[r]))));
        double2 xe = make_double2(f1 * d1.x, f1 * d1.y);
        xe = axpy(f6, d6, axpy(f5, d5, axpy(f4, d4, axpy(f3, d3, xe))));
        if constexpr (GSL) { xeg[r] = xe; continue; }
        if (valid(r)) {
          // odeint's rel_error: |xerr| / (epsAbs + epsRel (|x| + dt |dxdt|)); the maximum is taken
          // over squared ratios and returned squared (the controller works on the square)
          const double ax = fast_sqrt(cabs2(x[r])), ak = fast_sqrt(cabs2(d1));
          const double D = epsAbs + epsRel * (ax + dt * ak);
          double rc = __builtin_amdgcn_rcp(D);
          rc = fma(rc, fma(-D, rc, 1.0), rc);
          rc = fma(rc, fma(-D, rc, 1.0), rc);
          e = fmax(e, cabs2(xe) * (rc * rc));
        }
      }
    }
    if constexpr (GSL) {
      // GSL's rkck + control_standard + hadjust (ODE_GSL.h:44-84): the derivative at the new point
      // (dydt_out, a seventh evaluation), then max over the real components of
      // |yerr| / (epsAbs + epsRel (|y| + |dt dydt|)) with the new state y (not squared)
      T kg[ROWS];
      derivative(xn, kg, (t + dt) - t0);
      ++derivs;
      double m = 2.2250738585072014e-308;  // DBL_MIN, gsl_odeiv2_control_hadjust's start
#pragma unroll
      for (int r = 0; r < ROWS; ++r) {
        if (!valid(r)) continue;
        const double2 g = to_d(kg[r]);
        const double Dx = epsAbs + epsRel * (fabs(xn[r].x) + fabs(dt * g.x));
        const double Dy = epsAbs + epsRel * (fabs(xn[r].y) + fabs(dt * g.y));
        m = fmax(m, fmax(fabs(xeg[r].x) / Dx, fabs(xeg[r].y) / Dy));
      }
      return gmax(m);
    }
    return gmax(e);
  }

  __device__ __forceinline__ void derivative(const double2 x[ROWS], T k[ROWS], double tau = 0.) const {
    stage(x);
    setTime(tau);
    gather(k);
  }

  // y = J_m psi for the rows of this lane (psi staged in ps)
  __device__ __forceinline__ double2 jumpRow(int m, int r) const {
    const int row = r * G + lane;
    double2 acc = make_double2(0., 0.);
    const int base = m * S.wj * S.dim + row;
    const auto jc = gbl(S.jcol);
    for (int w = 0; w < S.wj; ++w) acc = cfma(gld2(S.jval, base + w * S.dim), ps[jc[base + w * S.dim]], acc);
    return acc;
  }

  __device__ void rates(double rt[kMaxJumps]) const {
    for (int m = 0; m < S.nj; ++m) {
      double s = 0.;
#pragma unroll
      for (int r = 0; r < ROWS; ++r) if (valid(r)) s += cabs2(jumpRow(m, r));
      rt[m] = gsum(s);
    }
  }

  // LDS scratch [8][kOredRow] for the batched stream-point reduction (one-wavefront teams; may be
  // null).  Lane 8j + c reads doubles 8c .. 8c + 7 of row j as four 16-byte reads; rows of 66
  // doubles keep them 16-byte aligned and start each row 4 banks on, so the 64 lanes spread over
  // all banks (bank-conflict cycles per JC-60 launch: 4.3e8 with 64-double rows, 5.8e8 with 65,
  // whose misaligned rows split the reads, 2.7e8 with 66).
  static constexpr int kOredRow = 66;
  double* ored = nullptr;
  // LDS copy of the first two channels' rate weights, [2][ROWS * G] (SysDev::jw; may be null)
  const double* jwl = nullptr;

  // all observables of a stream point at once (one wavefront with ored): each lane writes its
  // partials of eight observables to LDS, lane 8j + c adds the partials of lanes 8c..8c+7 for
  // observable j, three DPP steps finish the sum over the eight lanes, and lane 8j hands it to
  // store(q, value) -- one cross-lane pass per eight observables instead of one each
  template <class Store>
  __device__ void observablesBatched(const double2 x[ROWS], int no, Store store) const {
    for (int q0 = 0; q0 < no; q0 += 8) {
      const int nq = no - q0 < 8 ? no - q0 : 8;
      for (int j = 0; j < 8; ++j) ored[j * kOredRow + lane] = j < nq ? obsPartial(q0 + j, x) : 0.;
      wave_sync();
      const int j = lane >> 3, c = lane & 7;
      double v = 0.;
#pragma unroll
      for (int i = 0; i < 8; ++i) v += ored[j * kOredRow + c * 8 + i];
      v += dpp<0xB1>(v); v += dpp<0x4E>(v); v += dpp<0x141>(v);
      if (c == 0 && j < nq) store(q0 + j, v);
      wave_sync();
    }
  }

  __device__ double observable(int q, const double2 x[ROWS]) const { return gsum(obsPartial(q, x)); }

  // this lane's part of <x|O_q|x> (real or imaginary part as the observable asks)
  __device__ double obsPartial(int q, const double2 x[ROWS]) const {
    double s = 0.;
    const bool im = S.oimag[q] != 0;
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
      if (!valid(r)) continue;
      const int row = r * G + lane;
      double2 acc = make_double2(0., 0.);
      const int base = q * S.wo * S.dim + row;
      const int wq = gbl(S.owid)[q];
      const auto oc = gbl(S.ocol);
      for (int w = 0; w < wq; ++w) acc = cfma(gld2(S.oval, base + w * S.dim), ps[oc[base + w * S.dim]], acc);
      // conj(x) * acc
      s += im ? (x[r].x * acc.y - x[r].y * acc.x) : (x[r].x * acc.x + x[r].y * acc.y);
    }
    return s;
  }
};

// ---------------------------------------------------------------------------
// LDS-state engine for multi-wavefront teams: psi lives in LDS (FP64, read by every lane for
// rates, jumps and observables without a copy), only the six stage derivatives stay in
// registers, the step-start copy for a step-back goes to the trajectory's slot in HBM.  This
// halves the register footprint of a lane's rows and keeps large states (dim up to 4096 in a
// 16-wavefront team) free of register spills.
// ---------------------------------------------------------------------------

template <int ROWS, int NW, int WH, bool F32, bool TD = false, int OPT = 0>
struct LdsEngine : Engine<ROWS, NW, WH, F32, TD, OPT> {
  using B = Engine<ROWS, NW, WH, F32, TD, OPT>;
  using T = typename B::T;
  using B::G; using B::S; using B::xs; using B::red; using B::lane; using B::kd; using B::gsum; using B::gmax;
  double2* X;  // LDS state [ROWS*G] (== B::ps: jumpRow / observable gather from it)

  // xs2: optional second stage buffer (see Engine::xsB)
  __device__ __forceinline__ LdsEngine(const SysDev& s, T* xs_, double2* X_, double* red_, int lane_, T* ph_ = nullptr,
                       T* xs2 = nullptr)
      : B(s, xs_, X_, red_, lane_, ph_), X(X_) { B::xsB = xs2; }

  // k = (H/i) X at picture time tau
  __device__ __forceinline__ void derivativeX(T k[ROWS], double tau = 0.) const {
    if constexpr ((OPT & kTwoStage) != 0) {  // B::xsB is set with this option
      T* nb = xs == B::xsA ? B::xsB : B::xsA;
#pragma unroll
      for (int r = 0; r < ROWS; ++r) if (B::live(r)) nb[r * G + lane] = from_d<T>(X[r * G + lane]);
      this->sync();
      xs = nb;
    } else {
      this->sync();
#pragma unroll
      for (int r = 0; r < ROWS; ++r) if (B::live(r)) xs[r * G + lane] = from_d<T>(X[r * G + lane]);
      this->sync();
    }
    this->setTime(tau);
    B::gather(k);
  }

  __device__ __forceinline__ void stageX(const double2 v[ROWS]) const { B::stage(v); }

  // One RKCK54 try from X with k1 = (H/i) X; returns the odeint error norm and, when the try is
  // accepted (norm <= 1), replaces the lane's rows of X by the 5th-order solution.
  // GSL: GSL's controller instead -- X always takes the 5th-order solution, and the return value
  // is GSL's error ratio (not squared), which needs the derivative at the new point
  template <bool GSL = false>
  __device__ double tryStepX(const T k1[ROWS], double dt, double epsAbs, double epsRel, long long& derivs,
                             double t = 0., double t0 = 0.) const {
    T k2[ROWS], k3[ROWS], k4[ROWS], k5[ROWS], k6[ROWS];
    double2 v[ROWS];
    auto xr = [&](int r) { return X[r * G + lane]; };
    const double s21 = CK::a21 * dt;
#pragma unroll
    for (int r = 0; r < ROWS; ++r) if (B::live(r)) v[r] = axpy(s21, kd(k1[r]), xr(r));
    B::stage(v); this->setTime((t + CK::c2 * dt) - t0); B::gather(k2);
    { const double q1 = CK::a31 * dt, q2 = CK::a32 * dt;
#pragma unroll
      for (int r = 0; r < ROWS; ++r) if (B::live(r)) v[r] = axpy(q2, kd(k2[r]), axpy(q1, kd(k1[r]), xr(r))); }
    B::stage(v); this->setTime((t + CK::c3 * dt) - t0); B::gather(k3);
    { const double q1 = CK::a41 * dt, q2 = CK::a42 * dt, q3 = CK::a43 * dt;
#pragma unroll
      for (int r = 0; r < ROWS; ++r)
        if (B::live(r)) v[r] = axpy(q3, kd(k3[r]), axpy(q2, kd(k2[r]), axpy(q1, kd(k1[r]), xr(r)))); }
    B::stage(v); this->setTime((t + CK::c4 * dt) - t0); B::gather(k4);
    { const double q1 = CK::a51 * dt, q2 = CK::a52 * dt, q3 = CK::a53 * dt, q4 = CK::a54 * dt;
#pragma unroll
      for (int r = 0; r < ROWS; ++r)
        if (B::live(r))
          v[r] = axpy(q4, kd(k4[r]), axpy(q3, kd(k3[r]), axpy(q2, kd(k2[r]), axpy(q1, kd(k1[r]), xr(r))))); }
    B::stage(v); this->setTime((t + CK::c5 * dt) - t0); B::gather(k5);
    { const double q1 = CK::a61 * dt, q2 = CK::a62 * dt, q3 = CK::a63 * dt, q4 = CK::a64 * dt, q5 = CK::a65 * dt;
#pragma unroll
      for (int r = 0; r < ROWS; ++r)
        if (B::live(r))
          v[r] = axpy(q5, kd(k5[r]), axpy(q4, kd(k4[r]), axpy(q3, kd(k3[r]),
                 axpy(q2, kd(k2[r]), axpy(q1, kd(k1[r]), xr(r)))))); }
    B::stage(v); this->setTime((t + CK::c6 * dt) - t0); B::gather(k6);
    derivs += 5;
    const double q1 = CK::b1 * dt, q3 = CK::b3 * dt, q4 = CK::b4 * dt, q6 = CK::b6 * dt;
    if constexpr (GSL) {
      // ODE_GSL.h:44-84: rkck's yerr, the new state, its derivative (dydt_out), then
      // max over the real components of |yerr| / (epsAbs + epsRel (|y| + |dt dydt|))
      const double f1 = CK::e1 * dt, f3 = CK::e3 * dt, f4 = CK::e4 * dt, f5 = CK::e5 * dt, f6 = CK::e6 * dt;
      double2 xe[ROWS];
#pragma unroll
      for (int r = 0; r < ROWS; ++r) {
        if (!B::live(r)) continue;
        const double2 d1 = kd(k1[r]), d3 = kd(k3[r]), d4 = kd(k4[r]), d5 = kd(k5[r]), d6 = kd(k6[r]);
        xe[r] = axpy(f6, d6, axpy(f5, d5, axpy(f4, d4, axpy(f3, d3, make_double2(f1 * d1.x, f1 * d1.y)))));
        X[r * G + lane] = axpy(q6, d6, axpy(q4, d4, axpy(q3, d3, axpy(q1, d1, xr(r)))));
      }
      T kg[ROWS];
      derivativeX(kg, (t + dt) - t0);
      ++derivs;
      double m = 2.2250738585072014e-308;  // DBL_MIN, gsl_odeiv2_control_hadjust's start
#pragma unroll
      for (int r = 0; r < ROWS; ++r) {
        if (!B::valid(r)) continue;
        const double2 g = to_d(kg[r]), y = xr(r);
        const double Dx = epsAbs + epsRel * (fabs(y.x) + fabs(dt * g.x));
        const double Dy = epsAbs + epsRel * (fabs(y.y) + fabs(dt * g.y));
        m = fmax(m, fmax(fabs(xe[r].x) / Dx, fabs(xe[r].y) / Dy));
      }
      return gmax(m);
    }
    double e = 0.;
    {
      const double f1 = CK::e1 * dt, f3 = CK::e3 * dt, f4 = CK::e4 * dt, f5 = CK::e5 * dt, f6 = CK::e6 * dt;
#pragma unroll
      for (int r = 0; r < ROWS; ++r) {
        if (!B::valid(r)) continue;
        const double2 d1 = kd(k1[r]);
        double2 xe = make_double2(f1 * d1.x, f1 * d1.y);
        xe = axpy(f6, kd(k6[r]), axpy(f5, kd(k5[r]), axpy(f4, kd(k4[r]), axpy(f3, kd(k3[r]), xe))));
        const double ax = fast_sqrt(cabs2(xr(r))), ak = fast_sqrt(cabs2(d1));
        const double D = epsAbs + epsRel * (ax + dt * ak);
        double rc = __builtin_amdgcn_rcp(D);
        rc = fma(rc, fma(-D, rc, 1.0), rc);
        rc = fma(rc, fma(-D, rc, 1.0), rc);
        e = fmax(e, cabs2(xe) * (rc * rc));
      }
    }
    const double err = gmax(e);  // squared error norm
    if (err <= 1.0) {
#pragma unroll
      for (int r = 0; r < ROWS; ++r)
        if (B::live(r))
          X[r * G + lane] = axpy(q6, kd(k6[r]), axpy(q4, kd(k4[r]), axpy(q3, kd(k3[r]), axpy(q1, kd(k1[r]), xr(r)))));
    }
    return err;
  }

  // <X|O_q|X>, real or imaginary part as the observable asks (inlined: an out-of-line member call
  // would put the engine object in private memory, behind a scratch load on every member access)
  __device__ __forceinline__ double observableX(int q) const {
    double s = 0.;
    const bool im = S.oimag[q] != 0;
#pragma unroll
    for (int r = 0; r < ROWS; ++r) {
      if (!B::valid(r)) continue;
      const int row = r * G + lane;
      double2 acc = make_double2(0., 0.);
      const int base = q * S.wo * S.dim + row;
      const int wq = S.owid[q];
      for (int w = 0; w < wq; ++w) acc = cfma(gld2(S.oval, base + w * S.dim), X[gbl(S.ocol)[base + w * S.dim]], acc);
      const double2 x = X[row];
      s += im ? (x.x * acc.y - x.y * acc.x) : (x.x * acc.x + x.y * acc.y);
    }
    return gsum(s);
  }
};

// advance_traj for the LDS-state engine: the same MCWF / ODE control flow, with psi in X.
template <class Eng, bool GSL = false>
__device__ __forceinline__ void advance_traj_lds(const SysDev& S, const EnsDev& E, const Eng& eng, int traj,
                                                 double* rts, long long* cnt) {
  constexpr int ROWS = Eng::kRows, G = Eng::G, NW = Eng::kNW;
  using T = typename Eng::T;
  const int lane = eng.lane;
  double2* X = eng.X;
  double2* slot = E.psi + (size_t)traj * S.dim;  // step-start copy (the launch's final state goes here too)
  auto bump = [&](int c, long long v) { if (lane == 0) cnt[c] += v; };

  T k1[ROWS];
  eng.sync();  // previous users of X / cnt are done
#pragma unroll
  for (int r = 0; r < ROWS; ++r) if (eng.live(r)) X[r * G + lane] = eng.valid(r) ? slot[r * G + lane] : make_double2(0., 0.);
  double t = E.t[traj], dtDid = E.dtDid[traj], dtTry = E.dtTry[traj], t0 = E.t0[traj];
  unsigned long long ord = E.ordinal[traj];
  // index of the draw of ordinal ord in the trajectory's row of E.draws: ord - ord0
  const unsigned long long ord0 = E.draw_base ? E.draw_base[traj] : ord;
  const double* draws = E.draws ? E.draws + (size_t)traj * E.draw_cap : nullptr;
  const double2* exz = S.exz;
  if (lane < kCounters) cnt[lane] = E.counters[(size_t)traj * kCounters + lane];
  eng.sync();
  const unsigned long long key = E.traj0 + (unsigned long long)traj;
  const int nj = S.nj;
  const double epsAbs = E.epsAbs, epsRel = E.epsRel;

  auto coherent = [&](double Dt) {  // MCWF_Trajectory.h:250-276 with ODE.h:234-262
    const double timeBefore = t;
    const double nextDtTry = (fabs(Dt) * 10. < fabs(dtTry) ? dtTry : 0.);  // |Dt| < |dtTry| / 10
    if ((Dt > 0.) != (dtTry > 0.)) dtTry = -dtDid;
    if (fabs(dtTry) > fabs(Dt)) dtTry = Dt;
    eng.derivativeX(k1, t - t0);
    long long derivs = 1, tries = 0, failed = 0;
    if constexpr (GSL) {  // ControlledErrorStepperGSL::try_step (ODE_GSL.h:44-84), as in advance_traj
      ++tries;
      const double rmax = eng.template tryStepX<true>(k1, dtTry, epsAbs, epsRel, derivs, t, t0);
      const double hOld = dtTry;
      t += dtTry;
      if (rmax > 1.1) {
        const double r = 0.9 * inv_root<5>(rmax);
        dtTry = (r < 0.2 ? 0.2 : r) * hOld;
      } else if (rmax < 0.5) {
        const double r = rmax < 3.4012224e-5 ? 5.0 : 0.9 * inv_root<6>(rmax);
        dtTry = (r > 5.0 ? 5.0 : r < 1.0 ? 1.0 : r) * hOld;
      }
    }
    for (; !GSL;) {
      ++tries;
      const double err = eng.tryStepX(k1, dtTry, epsAbs, epsRel, derivs, t, t0);
      // err is the squared norm m: err > 1 <=> m > 1, err < 0.5 <=> m < 0.25,
      // err^(-1/3) = m^(-1/6), max(5^-5, err)^(-1/5) = max(5^-10, m)^(-1/10)
      if (err > 1.0) {
        dtTry *= fmax(9. / 10. * inv_root_sq<6>(err), 1. / 5.);
        ++failed;
        continue;
      }
      t += dtTry;
      if (err < 0.25) dtTry *= 9. / 10. * inv_root_sq<10>(fmax(0.00032 * 0.00032, err));
      break;
    }
    dtDid = t - timeBefore;
    if (nextDtTry != 0.) dtTry = nextDtTry;
    if (lane == 0) { cnt[kOdeSteps] += 1; cnt[kTries] += tries; cnt[kFailed] += failed; cnt[kDerivs] += derivs; }
    if (Eng::kTD && exz) {  // Exact::actWithU on the lane's rows (MCWF_Trajectory.h:269-272)
      const double dtau = t - t0;
#pragma unroll
      for (int r = 0; r < ROWS; ++r)
        if (eng.valid(r)) {
          const double2 z = exz[r * G + lane];
          const double e = exp(-z.x * dtau);
          double sn, cs;
          sincos(-z.y * dtau, &sn, &cs);
          X[r * G + lane] = cmul(X[r * G + lane], make_double2(e * cs, e * sn));
        }
      t0 = t;
    }
    double s2 = 0.;  // renorm (StateVector.h:148-153)
#pragma unroll
    for (int r = 0; r < ROWS; ++r) if (eng.valid(r)) s2 += cabs2(X[r * G + lane]);
    const double inv = fast_rsqrt(eng.gsum(s2));
#pragma unroll
    for (int r = 0; r < ROWS; ++r) if (eng.live(r)) { double2& x = X[r * G + lane]; x.x *= inv; x.y *= inv; }
    eng.sync();  // X complete before other lanes gather from it
  };

  const double* jw = S.jw;
  // every channel's rate in one team hand-over: each wavefront's partial sums go to red[m][wave],
  // then lane m adds channel m's partials in wave order (the sums of separate gsum calls)
  auto computeRates = [&]() {
    eng.sync();  // previous readers of red are done
    for (int m = 0; m < nj; ++m) {
      double sacc = 0.;
      if (jw) {  // rate weights: no gather
#pragma unroll
        for (int r = 0; r < ROWS; ++r)
          if (eng.valid(r)) sacc += gbl(jw)[(unsigned)(m * S.dim + r * G + lane)] * cabs2(X[r * G + lane]);
      } else {
#pragma unroll
        for (int r = 0; r < ROWS; ++r) if (eng.valid(r)) sacc += cabs2(eng.jumpRow(m, r));
      }
      sacc = wave_sum(sacc);
      if ((lane & 63) == 0) eng.red[m * NW + (lane >> 6)] = sacc;
    }
    eng.sync();
    if (lane < nj) {
      double sum = 0.;
      for (int k = 0; k < NW; ++k) sum += eng.red[lane * NW + k];
      rts[lane] = sum;
    }
    eng.sync();
  };

  for (int k = 0; k < E.n_intervals; ++k) {
    // Dt mode: advance(intervals[k]); dc mode: up to dc MCWF steps step(t_end - t) per stream point
    // (trajectory::run with dc > 0, Trajectory.h:560-600)
    const double endTime = E.dc > 0 ? E.t_end : t + E.intervals[k];
    const int maxSteps = E.dc > 0 ? E.dc : -1;
    for (int steps = 0;; ++steps) {
      const double Dt = endTime - t;
      if (Dt == 0. || steps == maxSteps) break;
      const unsigned long long di = ord - ord0;
      const double uPre = (draws && di < (unsigned long long)E.draw_cap) ? gbl(draws)[di] : -1.0;
#pragma unroll
      for (int r = 0; r < ROWS; ++r) if (eng.valid(r)) slot[r * G + lane] = X[r * G + lane];
      const double tCache = t, dtDidCache = dtDid;
      coherent(Dt);
      if (nj > 0) {
        computeRates();
        for (;;) {  // manageTimeStep (MCWF_Trajectory.h:294-317)
          double totalRate = 0.;
          for (int m = 0; m < nj; ++m) totalRate += rts[m];
          // Liouvillean::rates goes through LiouvilleanAveragedCommon::average, which refuses
          // non-finite values (LiouvilleanAveragedCommon.h:82-93): flag the launch
          if (!isfinite(totalRate) && lane == 0) atomicOr(E.err_flag, 2);
          const double sugg = E.dpLimit / totalRate;
          if (totalRate * dtDid > E.overshootTol * E.dpLimit) {
            t = tCache; dtDid = dtDidCache; dtTry = sugg;
            bump(kStepBacks, 1);
#pragma unroll
            for (int r = 0; r < ROWS; ++r) if (eng.live(r)) X[r * G + lane] = eng.valid(r) ? slot[r * G + lane] : make_double2(0., 0.);
            coherent(Dt);
            computeRates();
            continue;
          } else if (totalRate * dtTry > E.dpLimit) bump(kOvershoots, 1);
          dtTry = fmin(dtTry, sugg);
          break;
        }
        if (uPre < 0. && E.draws_strict && lane == 0) atomicOr(E.err_flag, 8);  // host draws ran out
        const double u = uPre >= 0. ? uPre : jump_uniform(E.seed, key, ord);  // performJump (:320-344)
        ++ord;
        double random = u / dtDid;
        int no = 0;
        for (; random > 0. && no != nj; random -= rts[no++]) ;
        if (random < 0.) {
          --no;
          const double nf = sqrt(rts[no]);
          if (!isfinite(nf) && lane == 0) atomicOr(E.err_flag, 1);
          const double inv = 1.0 / nf;
          double2 y[ROWS];
#pragma unroll
          for (int r = 0; r < ROWS; ++r) {
            const double2 a = eng.valid(r) ? eng.jumpRow(no, r) : make_double2(0., 0.);
            y[r] = make_double2(a.x * inv, a.y * inv);
          }
          eng.sync();  // every lane has read X
#pragma unroll
          for (int r = 0; r < ROWS; ++r) if (eng.live(r)) X[r * G + lane] = y[r];
          eng.sync();
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
    unsigned long long* accRow = E.acc + ((size_t)(traj % kAccStripes) * E.n_intervals + k) * (S.no + 1) * 2;
    const bool lastPoint = k == E.n_intervals - 1;
    double* rowOut = E.rows_out ? E.rows_out + ((size_t)traj * E.n_intervals + k) * (2 + S.no) : nullptr;
    for (int q = 0; q < S.no; ++q) {
      const double v = eng.observableX(q);
      if (lane == 0) {
        acc_add(accRow + 2 * q, v, E.acc_max, E.err_flag);
        if (lastPoint) E.last_obs[(size_t)traj * S.no + q] = v;
        if (rowOut) rowOut[2 + q] = v;
      }
    }
    if (lane == 0) {
      acc_add(accRow + 2 * S.no, dtDid, E.acc_max, E.err_flag);
      if (rowOut) { rowOut[0] = t; rowOut[1] = dtDid; }
    }
  }

#pragma unroll
  for (int r = 0; r < ROWS; ++r) if (eng.valid(r)) slot[r * G + lane] = X[r * G + lane];
  eng.sync();
  if (lane == 0) { E.t[traj] = t; E.dtDid[traj] = dtDid; E.dtTry[traj] = dtTry; E.t0[traj] = t0; E.ordinal[traj] = ord; }
  if (lane < kCounters) E.counters[(size_t)traj * kCounters + lane] = cnt[lane];
}

// ---------------------------------------------------------------------------
// one trajectory through every stream interval of a launch
// ---------------------------------------------------------------------------

// Advances trajectory `traj` of ensemble E through E.intervals and writes its stream outputs.
// Called by every lane of a team (eng.lane); pc (LDS, when PC_LDS) keeps psi at the start of the
// MCWF step for a step-back, rts / cnt (LDS) hold the step's jump rates and the trajectory's
// counters.
// kFirst, kLast: the stream intervals [kFirst, kLast) of the launch this call covers (kLast < 0: to
// the last one);
// the trajectory's state is read from E at the start and written back at the end, so a later call
// continues where this one stopped.
template <class Eng, bool PC_LDS, bool GSL = false>
__device__ __forceinline__ void advance_traj(const SysDev& S, const EnsDev& E, const Eng& eng, int traj, double2* pc,
                                             double* rts, long long* cnt, int kFirst = 0, int kLast = -1) {
  constexpr int ROWS = Eng::kRows, G = Eng::G, NW = Eng::kNW;
  using T = typename Eng::T;
  const int lane = eng.lane;
  double2 pcr[PC_LDS ? 1 : ROWS];
  auto bump = [&](int c, long long v) { if (lane == 0) cnt[c] += v; };

  double2 x[ROWS], xn[ROWS];
  T k1[ROWS];
#pragma unroll
  for (int r = 0; r < ROWS; ++r) x[r] = eng.valid(r) ? E.psi[(size_t)traj * S.dim + r * G + lane] : make_double2(0., 0.);
  double t = E.t[traj], dtDid = E.dtDid[traj], dtTry = E.dtTry[traj], t0 = E.t0[traj];
  unsigned long long ord = E.ordinal[traj];
  // index of the draw of ordinal ord in the trajectory's row of E.draws: ord - ord0
  const unsigned long long ord0 = E.draw_base ? E.draw_base[traj] : ord;
  const double* draws = E.draws ? E.draws + (size_t)traj * E.draw_cap : nullptr;
  const double2* exz = S.exz;
  if (lane < kCounters) cnt[lane] = E.counters[(size_t)traj * kCounters + lane];
  eng.sync();
  const unsigned long long key = E.traj0 + (unsigned long long)traj;
  const int nj = S.nj;
  const double epsAbs = E.epsAbs, epsRel = E.epsRel;
  constexpr bool gsl = GSL;  // GSL controller (ODE_GSL.h) instead of odeint's

  // coherent part of one MCWF step (MCWF_Trajectory.h:250-276 with ODE.h:234-262)
  auto coherent = [&](double Dt) {
    const double timeBefore = t;
    const double nextDtTry = (fabs(Dt) * 10. < fabs(dtTry) ? dtTry : 0.);  // |Dt| < |dtTry| / 10
    if ((Dt > 0.) != (dtTry > 0.)) dtTry = -dtDid;
    if (fabs(dtTry) > fabs(Dt)) dtTry = Dt;
    eng.derivative(x, k1, t - t0);
    long long derivs = 1, tries = 0, failed = 0;
    if constexpr (gsl) {
      // ControlledErrorStepperGSL::try_step (ODE_GSL.h:44-84): every step is accepted, then
      // gsl_odeiv2_control_hadjust sets the next dt from the error ratio with the order GSL's rkck
      // reports, 5: dt *= max(0.9 rmax^(-1/5), 0.2) above 1.1, clamp(0.9 rmax^(-1/6), 1, 5) below 0.5
      ++tries;
      const double rmax = eng.template tryStep<true>(x, k1, dtTry, epsAbs, epsRel, xn, derivs, t, t0);
      const double hOld = dtTry;
      t += dtTry;
      if (rmax > 1.1) {
        const double r = 0.9 * inv_root<5>(rmax);
        dtTry = (r < 0.2 ? 0.2 : r) * hOld;
      } else if (rmax < 0.5) {
        // 0.9 rmax^(-1/6) exceeds the cap 5 below rmax = 0.18^6
        const double r = rmax < 3.4012224e-5 ? 5.0 : 0.9 * inv_root<6>(rmax);
        dtTry = (r > 5.0 ? 5.0 : r < 1.0 ? 1.0 : r) * hOld;
      }
    }
    for (; !gsl;) {
      ++tries;
      const double err = eng.tryStep(x, k1, dtTry, epsAbs, epsRel, xn, derivs, t, t0);
      // err is the squared error norm m (tryStep): err > 1 <=> m > 1, err < 0.5 <=> m < 0.25
      if (err > 1.0) {
        // default_step_adjuster::decrease_step: dt *= max(0.9 err^(-1/(error_order-1)), 1/5)
        dtTry *= fmax(9. / 10. * inv_root_sq<6>(err), 1. / 5.);
        ++failed;
        continue;
      }
      t += dtTry;
      // increase_step: if err < 0.5, dt *= 0.9 max(5^-stepper_order, err)^(-1/stepper_order)
      if (err < 0.25) dtTry *= 9. / 10. * inv_root_sq<10>(fmax(0.00032 * 0.00032, err));
      break;
    }
#pragma unroll
    for (int r = 0; r < ROWS; ++r) x[r] = xn[r];
    dtDid = t - timeBefore;
    if (nextDtTry != 0.) dtTry = nextDtTry;
    if (Eng::kTD && exz) {  // Exact::actWithU, then the pictures coincide again (MCWF_Trajectory.h:269-272)
      const double dtau = t - t0;
#pragma unroll
      for (int r = 0; r < ROWS; ++r)
        if (eng.valid(r)) {
          const double2 z = exz[r * G + lane];
          const double e = exp(-z.x * dtau);
          double sn, cs;
          sincos(-z.y * dtau, &sn, &cs);
          x[r] = cmul(x[r], make_double2(e * cs, e * sn));
        }
      t0 = t;
    }
    if (lane == 0) { cnt[kOdeSteps] += 1; cnt[kTries] += tries; cnt[kFailed] += failed; cnt[kDerivs] += derivs; }
  };

  // renorm (StateVector.h:148-153) and, when there are jumps, the rates of the renormalised
  // state (MCWF_Trajectory.h:353-356): ||J_m psi/|psi| ||^2 = ||J_m psi||^2 / |psi|^2, so the norm
  // and the rates come from one pass over the staged unnormalised state and one joint reduction.
  // psi stays staged unnormalised in ps; jumpScale = 1/|psi| converts J ps to J psi.
  // With rate weights (SysDev::jw, up to two channels) the rates come from the lane's own rows,
  // nothing is staged and the rates stay in registers (rw0, rw1); a jump then stages psi.
  double jumpScale = 1.;
  const bool useW = S.jw != nullptr && nj <= 2;
  // the weights of the lane's rows: from the team's LDS copy (eng.jwl, [2][ROWS * G]) or global memory
  auto weight = [&](int m, int r) {
    const int row = r * G + lane;
    if (eng.jwl) return eng.jwl[m * ROWS * G + row];
    return m < nj && row < S.dim ? gbl(S.jw)[(unsigned)(m * S.dim + row)] : 0.;
  };
  double rw0 = 0., rw1 = 0.;
  auto rate = [&](int m) { return useW ? (m == 0 ? rw0 : rw1) : rts[m]; };
  auto renormRates = [&](bool withRates) {
    double part[3] = {0., 0., 0.};
    if (withRates && useW) {
#pragma unroll
      for (int r = 0; r < ROWS; ++r) {
        const double a = cabs2(x[r]);
        part[0] += a;
        part[1] += weight(0, r) * a;
        part[2] += weight(1, r) * a;
      }
      wave_sum3(part);
      const double inv = fast_rsqrt(part[0]);
      const double inv2 = inv * inv;
#pragma unroll
      for (int r = 0; r < ROWS; ++r) { x[r].x *= inv; x[r].y *= inv; }
      rw0 = part[1] * inv2;
      rw1 = part[2] * inv2;
      jumpScale = 1.;  // a jump stages the renormalised psi
      return;
    }
    const bool fused = withRates && nj <= 2;
    if (withRates) eng.stagePsi(x);
#pragma unroll
    for (int r = 0; r < ROWS; ++r) if (eng.valid(r)) part[0] += cabs2(x[r]);
    if (fused) {
      for (int m = 0; m < nj; ++m) {
        double sacc = 0.;
#pragma unroll
        for (int r = 0; r < ROWS; ++r) if (eng.valid(r)) sacc += cabs2(eng.jumpRow(m, r));
        part[1 + m] = sacc;
      }
      wave_sum3(part);
    } else {
      part[0] = eng.gsum(part[0]);
    }
    const double inv = fast_rsqrt(part[0]);
    const double inv2 = inv * inv;
#pragma unroll
    for (int r = 0; r < ROWS; ++r) { x[r].x *= inv; x[r].y *= inv; }
    jumpScale = inv;
    if (fused) {
      if (lane == 0) for (int m = 0; m < nj; ++m) rts[m] = part[1 + m] * inv2;
    } else if (withRates) {
      for (int m = 0; m < nj; ++m) {
        double sacc = 0.;
#pragma unroll
        for (int r = 0; r < ROWS; ++r) if (eng.valid(r)) sacc += cabs2(eng.jumpRow(m, r));
        const double v = eng.gsum(sacc) * inv2;
        if (lane == 0) rts[m] = v;
      }
    }
    if (withRates) eng.sync();
  };

  const int kEnd = kLast < 0 ? E.n_intervals : kLast;
  for (int k = kFirst; k < kEnd; ++k) {
    // Dt mode: advance(intervals[k]); dc mode: up to dc MCWF steps step(t_end - t) per stream point
    // (trajectory::run with dc > 0, Trajectory.h:560-600)
    const double endTime = E.dc > 0 ? E.t_end : t + E.intervals[k];
    const int maxSteps = E.dc > 0 ? E.dc : -1;
    for (int steps = 0;; ++steps) {
      const double Dt = endTime - t;
      if (Dt == 0. || steps == maxSteps) break;
      // MCWF_Trajectory::step (MCWF_Trajectory.h:347-374)
      // the step's jump draw, fetched early so the load completes behind the RK work
      const unsigned long long di = ord - ord0;
      const double uPre = (draws && di < (unsigned long long)E.draw_cap) ? gbl(draws)[di] : -1.0;
#pragma unroll
      for (int r = 0; r < ROWS; ++r) {
        if constexpr (PC_LDS) pc[r * G + lane] = x[r];
        else pcr[r] = x[r];
      }
      const double tCache = t, dtDidCache = dtDid;
      coherent(Dt);
      renormRates(nj > 0);
      if (nj > 0) {
        for (;;) {  // manageTimeStep (MCWF_Trajectory.h:294-317)
          double totalRate = 0.;
          for (int m = 0; m < nj; ++m) totalRate += rate(m);
          // Liouvillean::rates goes through LiouvilleanAveragedCommon::average, which refuses
          // non-finite values (LiouvilleanAveragedCommon.h:82-93): flag the launch
          if (!isfinite(totalRate) && lane == 0) atomicOr(E.err_flag, 2);
          const double sugg = E.dpLimit / totalRate;
          if (totalRate * dtDid > E.overshootTol * E.dpLimit) {
            t = tCache; dtDid = dtDidCache; dtTry = sugg;
            bump(kStepBacks, 1);
#pragma unroll
            for (int r = 0; r < ROWS; ++r) {
              if constexpr (PC_LDS) x[r] = pc[r * G + lane];
              else x[r] = pcr[r];
            }
            coherent(Dt);
            renormRates(true);
            continue;
          } else if (totalRate * dtTry > E.dpLimit) bump(kOvershoots, 1);
          dtTry = fmin(dtTry, sugg);
          break;
        }
        // performJump (MCWF_Trajectory.h:320-344): one draw per step; host-supplied draws that ran
        // out fail the launch (strict), otherwise the counter RNG fills in
        if (uPre < 0. && E.draws_strict && lane == 0) atomicOr(E.err_flag, 8);
        const double u = uPre >= 0. ? uPre : jump_uniform(E.seed, key, ord);
        ++ord;
        double random = u / dtDid;
        int no = 0;
        for (; random > 0. && no != nj; random -= rate(no++)) ;
        if (random < 0.) {
          --no;
          const double nf = sqrt(rate(no));
          if (!isfinite(nf) && lane == 0) atomicOr(E.err_flag, 1);
          if (useW) eng.stagePsi(x);
          const double inv = jumpScale / nf;  // ps holds psi before the renorm (after it with weights)
#pragma unroll
          for (int r = 0; r < ROWS; ++r) {
            const double2 y = eng.valid(r) ? eng.jumpRow(no, r) : make_double2(0., 0.);
            x[r] = make_double2(y.x * inv, y.y * inv);
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
    unsigned long long* accRow = E.acc + ((size_t)(traj % kAccStripes) * E.n_intervals + k) * (S.no + 1) * 2;
    const bool lastPoint = k == E.n_intervals - 1;
    double* rowOut = E.rows_out ? E.rows_out + ((size_t)traj * E.n_intervals + k) * (2 + S.no) : nullptr;
    auto store = [&](int q, double v) {
      acc_add(accRow + 2 * q, v, E.acc_max, E.err_flag);
      if (lastPoint) E.last_obs[(size_t)traj * S.no + q] = v;
      if (rowOut) rowOut[2 + q] = v;
    };
    if (NW == 1 && eng.ored) {
      eng.observablesBatched(x, S.no, store);
    } else {
      for (int q = 0; q < S.no; ++q) {
        const double v = eng.observable(q, x);
        if (lane == 0) store(q, v);
      }
    }
    if (lane == 0) {
      acc_add(accRow + 2 * S.no, dtDid, E.acc_max, E.err_flag);
      if (rowOut) { rowOut[0] = t; rowOut[1] = dtDid; }
    }
  }

#pragma unroll
  for (int r = 0; r < ROWS; ++r) if (eng.valid(r)) E.psi[(size_t)traj * S.dim + r * G + lane] = x[r];
  eng.sync();
  if (lane == 0) { E.t[traj] = t; E.dtDid[traj] = dtDid; E.dtTry[traj] = dtTry; E.t0[traj] = t0; E.ordinal[traj] = ord; }
  if (lane < kCounters) E.counters[(size_t)traj * kCounters + lane] = cnt[lane];
}

// ---------------------------------------------------------------------------
// interleaved chunks of trajectories (EnsDev::chunk)
// ---------------------------------------------------------------------------

constexpr int kErrChunkWait = 16;  // err_flag bit: a chunk's wait for its predecessor timed out

// Waits until the trajectory's progress count *p reaches `target` (its previous chunk is done),
// then acquires that chunk's stores of the trajectory state.  The chunk was claimed earlier by a
// resident wavefront, which waits only on chunks claimed before its own, so the wait ends; as a
// guard it gives up after ~10 s (100 MHz real-time counter) or once another wait has given up,
// flagging the launch.
__device__ __forceinline__ bool chunk_wait(const int* p, int target, int* err) {
  const unsigned long long start = __builtin_amdgcn_s_memrealtime();
  while (__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
    if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & kErrChunkWait) return false;
    if (__builtin_amdgcn_s_memrealtime() - start > 1000000000ull) {
      if ((threadIdx.x & 63) == 0) atomicOr(err, kErrChunkWait);
      return false;
    }
    __builtin_amdgcn_s_sleep(8);
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  return true;
}

// Publishes the trajectory state written by this wavefront, then the new progress count.
__device__ __forceinline__ void chunk_publish(int* p, int count) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  wave_sync();
  if ((threadIdx.x & 63) == 0) __hip_atomic_store(p, count, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// psi at the start of the MCWF step goes to LDS when the stage vector and it fit
template <int ROWS, int NW>
constexpr bool pc_in_lds() { return (size_t)ROWS * 64 * NW * 2 * sizeof(double2) <= 150 * 1024; }

// ---------------------------------------------------------------------------
// batched persistent MCWF kernel: one system, one team = one work group
// ---------------------------------------------------------------------------

template <int ROWS, int NW, int WH, bool F32, int MINW, bool TD, bool GSL = false>
__global__ void __launch_bounds__(64 * NW, MINW) mcwf_batch_kernel(SysDev Sarg, EnsDev Earg) {
  using Eng = Engine<ROWS, NW, WH, F32, TD, kOffsets>;
  using T = typename Eng::T;
  constexpr int G = Eng::G;
  constexpr bool kPcLds = pc_in_lds<ROWS, NW>();
  // kernel arguments live in LDS: read where used instead of pinning ~80 scalar registers
  __shared__ SysDev S;
  __shared__ EnsDev E;
  __shared__ double red[NW > 1 ? NW * kMaxJumps : 1];
  __shared__ double rts[kMaxJumps];        // jump rates of the current step
  __shared__ long long cnt[kCounters];     // this trajectory's counters (lane 0 updates)
  __shared__ T ph[kMaxFreqs];              // interaction-picture phasors of the current stage
  __shared__ int shTraj;

  const int lane = threadIdx.x;
  if (lane == 0) { S = Sarg; E = Earg; }
  __syncthreads();
  auto claim = [&]() {
    __syncthreads();
    if (lane == 0) shTraj = atomicAdd(E.work, 1);
    __syncthreads();
    return shTraj;
  };
  if constexpr (NW > 1) {
    // multi-wavefront team: psi in LDS
    __shared__ double2 Xs[ROWS * G];
    __shared__ T xs[ROWS * G];
    // the H table's coefficient dictionary with its index bytes (Schroedinger picture), when LDS
    // allows, and a second stage buffer when it fits beside the state (saves a barrier per
    // derivative); the dictionary goes first (dim 4096: 297 -> 209 ms without the second buffer)
    constexpr size_t kBase = (size_t)ROWS * G * (sizeof(double2) + sizeof(T));
    constexpr size_t kIdx = (size_t)kMaxDictSlots * ROWS * G;
    constexpr bool kDict = !TD && kBase + kIdx + kMaxDict * sizeof(T) + 8 * 1024 <= 160 * 1024;
    constexpr size_t kDictBytes = kDict ? kIdx + kMaxDict * sizeof(T) : 0;
    constexpr bool kTwo = kBase + ROWS * G * sizeof(T) + kDictBytes <= 150 * 1024;
    __shared__ T xs2[kTwo ? ROWS * G : 1];
    LdsEngine<ROWS, NW, WH, F32, TD, kOffsets | (kTwo ? kTwoStage : 0)> eng(S, xs, Xs, red, lane, ph,
                                                                                 kTwo ? xs2 : nullptr);
    __shared__ T dict[kDict ? kMaxDict : 1];
    __shared__ int dbase[kMaxDictSlots], doff[kMaxDictSlots];
    __shared__ __align__(16) unsigned char didx[kDict ? kIdx : 4];
    if constexpr (kDict) eng.loadDict(dict, dbase, didx, (int)kIdx, doff);
    for (int traj; (traj = claim()) < E.n_traj;) advance_traj_lds<decltype(eng), GSL>(S, E, eng, traj, rts, cnt);
  } else {
    // one wavefront: psi in registers; the RK stage vector (derivative precision) and the FP64
    // state copy share storage
    __shared__ double2 ps[ROWS * G];
    __shared__ double2 pc[kPcLds ? ROWS * G : 1];
    __shared__ double ored[8 * Eng::kOredRow];
    __shared__ double jwl[2 * ROWS * G];  // rate weights of the system (the same for every trajectory)
    Eng eng(S, reinterpret_cast<T*>(ps), ps, red, lane, ph);
    eng.ored = ored;
    if (S.jw && S.nj <= 2) {
#pragma unroll
      for (int m = 0; m < 2; ++m)
#pragma unroll
        for (int r = 0; r < ROWS; ++r) {
          const int row = r * G + lane;
          jwl[m * ROWS * G + row] = m < S.nj && row < S.dim ? gbl(S.jw)[(unsigned)(m * S.dim + row)] : 0.;
        }
      eng.jwl = jwl;
      wave_sync();
    }
    if (E.chunk > 0) {
      // interleaved chunks: every trajectory's chunk c is claimed before any chunk c + 1, so all of
      // them start early instead of the last ones queueing behind whole trajectories (the costs of
      // trajectories are heavy-tailed: one that starts late and takes several times the mean
      // number of steps otherwise sets the length of the launch).  (Raising the issue priority of
      // chunks that fall behind was measured to slow the launch down and is not done.)
      const int nt = E.n_traj, ch = E.chunk;
      const int items = (E.n_intervals + ch - 1) / ch * nt;
      for (int item; (item = claim()) < items;) {
        const int c = item / nt, traj = item - c * nt;
        if (c > 0 && !chunk_wait(E.prog + traj, c, E.err_flag)) continue;
        advance_traj<Eng, kPcLds, GSL>(S, E, eng, traj, pc, rts, cnt, c * ch, min(E.n_intervals, (c + 1) * ch));
        chunk_publish(E.prog + traj, c + 1);
      }
    } else {
      for (int traj; (traj = claim()) < E.n_traj;) advance_traj<Eng, kPcLds, GSL>(S, E, eng, traj, pc, rts, cnt);
    }
  }
}

// One-wavefront trajectory runner of the mixed kernel, kept out of line: it gets its own
// register allocation, so the spills of the large-state paths do not reach it.  ph: the
// wavefront's phasor table (interaction-picture members), else null.
template <class Eng, bool GSL>
__device__ __noinline__ void run_wave_traj(const SysDev* S, const EnsDev* E, double2* ps, double2* pc, double* rts,
                                           long long* cnt, int lane, int traj, typename Eng::T* ph) {
  Eng eng(*S, reinterpret_cast<typename Eng::T*>(ps), ps, nullptr, lane, ph);
  advance_traj<Eng, true, GSL>(*S, *E, eng, traj, pc, rts, cnt);
}

// the members' controller picks the advance: GSL's (ODE_GSL.h) when the variant carries it
template <bool GSLC, class Eng>
__device__ __forceinline__ void advance_member_lds(const SysDev& S, const EnsDev& E, const Eng& eng, int traj,
                                                   double* rts, long long* cnt) {
  if (GSLC && E.gsl) advance_traj_lds<Eng, true>(S, E, eng, traj, rts, cnt);
  else advance_traj_lds<Eng, false>(S, E, eng, traj, rts, cnt);
}

// ---------------------------------------------------------------------------
// mixed batch: trajectories of up to kMaxMembers ensembles (different systems) in one batched
// launch of two persistent kernels.  Three team kinds:
//   kind 0: a work group of NWB wavefronts owns a trajectory (RB rows per lane, psi in LDS);
//   kind 2: a quarter team of NWB/4 wavefronts owns a trajectory (RB rows per lane, psi in LDS);
//   kind 1: a wavefront owns a trajectory (RW rows per lane, psi in registers).
// PH selects the kinds a kernel runs: PH = 1, kind 0 on NWB-wavefront work groups; PH = 14,
// kinds 2 and 1 on work groups of one quarter team each (bit 3), which first drain the kind-2
// queue together (work-group barrier) and then split into wavefronts for the kind-1 queue; the
// team's LDS region holds its wavefronts' kind-1 regions.  The host launches the two on two
// streams: each gets a register allocation of its own (the work-group engine's state and the
// smaller engines' do not spill into each other; a quarter team's lanes get the registers of a
// 64 NWB / 4 thread group), and the second fills compute units as the first drains.
// TDC: members may run in the interaction picture (engines with phasor tables; a member without
// frequencies skips them at run time); GSLC: members may use GSL's controller.  The host picks
// the plain variant unless some member needs one of them.
// ---------------------------------------------------------------------------

template <int NWB, int RB, int WHB, int RW, int WHW, bool F32, bool TDC, bool GSLC, int PH>
__global__ void __launch_bounds__(64 * ((PH & 8) ? NWB / 4 : NWB), (PH & 8) ? 2 : 1)
    mcwf_mixed_kernel(const MixDev* __restrict__ Mg) {
  constexpr bool kP1 = (PH & 1) != 0, kP2 = (PH & 2) != 0, kP3 = (PH & 4) != 0, kQG = (PH & 8) != 0;
  static_assert(!(kQG && kP1), "a quarter-team work group runs phases 2 and 3 only");
  static_assert(kQG || !(kP2 || kP3), "quarter and wavefront teams run on quarter-team work groups");
  constexpr int kThreads = 64 * (kQG ? NWB / 4 : NWB);
  constexpr int NWQ = NWB / 4;
  // (a second stage buffer for the work-group team measured 5% slower here: 404 vs 384 ms on
  // the 8192-trajectory four-system batch, same GPU and layout)
  using EngB = LdsEngine<RB, NWB, WHB, F32, TDC, kOffsets>;
  using EngQ = LdsEngine<RB, NWQ, WHB, F32, TDC, kOffsets>;
  using EngW = Engine<RW, 1, WHW, F32, TDC>;
  using TB = typename EngB::T;
  using TW = typename EngW::T;
  constexpr int GB = 64 * NWB, GQ = 64 * NWQ;
  // a team's phasor table (16-byte aligned) at the end of its LDS region
  constexpr size_t kPh = TDC ? kMaxFreqs * sizeof(TB) + 16 : 0;
  constexpr size_t kBlockBytes = (size_t)RB * GB * (sizeof(double2) + sizeof(TB));
  constexpr size_t kWaveEach = ((size_t)RW * 64 * sizeof(double2) * 2 + (kMaxJumps + kCounters) * 8 + kPh + 15) / 16 * 16;
  constexpr size_t kQuarterOwn =
      (size_t)RB * GQ * (sizeof(double2) + sizeof(TB)) + (NWQ * kMaxJumps + kMaxJumps + kCounters) * 8 + kPh;
  constexpr size_t kQuarterRaw = kQuarterOwn > kWaveEach * NWQ ? kQuarterOwn : kWaveEach * NWQ;
  constexpr size_t kQuarter = (kQuarterRaw + 15) / 16 * 16;
  constexpr size_t kTeams = kP2 || kP3 ? (kQG ? 1 : 4) * kQuarter : 0;
  constexpr size_t kBytes = kP1 && kBlockBytes > kTeams ? kBlockBytes : kTeams;
  // the work-group team's coefficient dictionary with its index bytes, where LDS allows
  // (the index bytes get what LDS has left, up to kMaxDictSlots slots; loadDict checks the fit)
  constexpr long kIdxAvail = 160L * 1024 - (long)kBytes - kMaxDict * (long)sizeof(TB) - (long)sizeof(MixDev) - 6 * 1024;
  constexpr size_t kIdxB = kIdxAvail < 4L * RB * GB ? 0
                           : (size_t)(kIdxAvail / 1024 * 1024 < (long)kMaxDictSlots * RB * GB ? kIdxAvail / 1024 * 1024
                                                                                              : (long)kMaxDictSlots * RB * GB);
  constexpr bool kDictB = kP1 && kIdxB > 0;
  __shared__ MixDev M;
  __shared__ __align__(16) unsigned char pool[kBytes];
  __shared__ TB dictS[kDictB ? kMaxDict : 1];
  __shared__ int dbaseS[kMaxDictSlots], doffS[kMaxDictSlots];
  __shared__ __align__(16) unsigned char didxS[kDictB ? kIdxB : 4];
  __shared__ double red[kP1 ? NWB * kMaxJumps : 1];
  __shared__ double rtsB[kMaxJumps];
  __shared__ long long cntB[kCounters];
  __shared__ int shItem;
  // the quarter team's claimed item, outside the pool (phase 3 reuses the pool)
  __shared__ int qslot[1];
  __shared__ TB phB[TDC ? kMaxFreqs : 1];
  auto align16 = [](void* p) {
    return reinterpret_cast<TB*>((reinterpret_cast<size_t>(p) + 15) / 16 * 16);
  };

  {  // copy the member descriptors to LDS (16-byte words)
    const int4* src = reinterpret_cast<const int4*>(Mg);
    int4* dst = reinterpret_cast<int4*>(&M);
    for (int i = threadIdx.x; i < (int)(sizeof(MixDev) / sizeof(int4)); i += kThreads) dst[i] = src[i];
  }
  __syncthreads();

  auto member = [&](int kind, int item) {  // members of one kind are queued in member order
    int m = -1;
    for (int k = 0; k < M.n_members; ++k)
      if (M.team[k] == kind && M.first[k] <= item) m = k;
    return m;
  };

  // phase 1: one trajectory per work group
  if (kP1 && M.items[0] > 0) {
    double2* X = reinterpret_cast<double2*>(pool);
    TB* xs = reinterpret_cast<TB*>(X + RB * GB);
    for (;;) {
      __syncthreads();
      if (threadIdx.x == 0) shItem = atomicAdd(M.queue, 1);
      __syncthreads();
      const int item = shItem;
      if (item >= M.items[0]) break;
      const int m = member(0, item);
      EngB eng(M.S[m], xs, X, red, (int)threadIdx.x, TDC ? phB : nullptr);
      if constexpr (kDictB) eng.loadDict(dictS, dbaseS, didxS, (int)kIdxB, doffS);
      advance_member_lds<GSLC>(M.S[m], M.E[m], eng, item - M.first[m], rtsB, cntB);
    }
    __syncthreads();
  }

  unsigned char* qbase = pool;  // the quarter team's region (the whole pool of a PH = 14 work group)

  // phase 2: one trajectory per quarter team
  if (kP2 && M.items[2] > 0) {
    const int qlane = threadIdx.x % GQ;
    double2* X = reinterpret_cast<double2*>(qbase);
    TB* xs = reinterpret_cast<TB*>(X + RB * GQ);
    double* qred = reinterpret_cast<double*>(reinterpret_cast<unsigned char*>(xs) + RB * GQ * sizeof(TB));
    double* rts = qred + NWQ * kMaxJumps;
    long long* cnt = reinterpret_cast<long long*>(rts + kMaxJumps);
    TB* qph = TDC ? align16(cnt + kCounters) : nullptr;
    int* slot = qslot;
    for (;;) {
      __syncthreads();
      if (qlane == 0) *slot = atomicAdd(M.queue + 2, 1);
      __syncthreads();
      const int item = *slot;
      if (item >= M.items[2]) break;
      const int m = member(2, item);
      EngQ eng(M.S[m], xs, X, qred, qlane, qph);
      advance_member_lds<GSLC>(M.S[m], M.E[m], eng, item - M.first[m], rts, cnt);
    }
    __syncthreads();  // the team's LDS is free only when all of its wavefronts are done with it
  }

  // phase 3: one trajectory per wavefront (inside its quarter's region)
  if constexpr (kP3) {
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    unsigned char* mine = qbase + kWaveEach * (w % NWQ);
    double2* ps = reinterpret_cast<double2*>(mine);
    double2* pc = ps + RW * 64;
    double* rts = reinterpret_cast<double*>(pc + RW * 64);
    long long* cnt = reinterpret_cast<long long*>(rts + kMaxJumps);
    TW* wph = TDC ? reinterpret_cast<TW*>(align16(cnt + kCounters)) : nullptr;
    for (;;) {
      int item = 0;
      if (lane == 0) item = atomicAdd(M.queue + 1, 1);
      item = __builtin_amdgcn_readfirstlane(item);
      if (item >= M.items[1]) break;
      const int m = member(1, item);
      if (GSLC && M.E[m].gsl) run_wave_traj<EngW, true>(&M.S[m], &M.E[m], ps, pc, rts, cnt, lane, item - M.first[m], wph);
      else run_wave_traj<EngW, false>(&M.S[m], &M.E[m], ps, pc, rts, cnt, lane, item - M.first[m], wph);
      wave_sync();
    }
  }
}

// ---------------------------------------------------------------------------
// compiled variants
// ---------------------------------------------------------------------------

// Every (rows, waves, ELL width) combination compiled into the library.
#define QT_LAYOUTS(X) \
  X(1, 1, 4) X(1, 1, 8) X(1, 1, 16) \
  X(2, 1, 4) X(2, 1, 8) X(2, 1, 16) \
  X(4, 1, 8) X(4, 1, 16) \
  X(1, 4, 8) X(1, 4, 16) X(2, 4, 8) X(2, 4, 16) X(4, 4, 16) \
  X(4, 8, 16) X(2, 16, 16) X(4, 16, 16)

// Batched-kernel variants: (rows, waves, ELL width, minimum waves per SIMD requested from the
// register allocator).  The occupancy hint trades registers (and spills) for latency hiding.
#define QT_BATCH(X) \
  X(1, 1, 4, 1) X(1, 1, 4, 4) X(1, 1, 4, 5) X(1, 1, 4, 6) X(1, 1, 4, 8) \
  X(1, 1, 8, 1) X(1, 1, 8, 4) X(1, 1, 8, 5) \
  X(1, 1, 16, 1) \
  X(2, 1, 4, 1) X(2, 1, 8, 1) X(2, 1, 16, 1) \
  X(4, 1, 8, 1) X(4, 1, 16, 1) \
  X(1, 4, 8, 1) X(1, 4, 16, 1) X(2, 4, 8, 1) X(2, 4, 16, 1) X(4, 4, 16, 1) \
  X(4, 8, 16, 1) X(2, 16, 16, 1) X(4, 16, 16, 1)

// Interaction-picture batched variants (rows, waves, ELL width): one per dimension class
#define QT_BATCH_TD(X) \
  X(1, 1, 4, 1) X(1, 1, 8, 1) X(1, 1, 16, 1) X(2, 1, 8, 1) X(2, 1, 16, 1) X(4, 1, 16, 1) \
  X(1, 4, 16, 1) X(2, 4, 16, 1) X(4, 4, 16, 1) X(4, 8, 16, 1) X(2, 16, 16, 1) X(4, 16, 16, 1)

// GSL-controller variants (ODE_GSL.h), Schroedinger and interaction picture: one per dimension
// class, the ELL width rounded up to a compiled one
#define QT_BATCH_GSL(X) \
  X(1, 1, 4, 1) X(1, 1, 8, 1) X(1, 1, 16, 1) X(2, 1, 8, 1) X(2, 1, 16, 1) X(4, 1, 16, 1) \
  X(1, 4, 16, 1) X(2, 4, 16, 1) X(4, 4, 16, 1) X(4, 8, 16, 1) X(2, 16, 16, 1) X(4, 16, 16, 1)

// One launch table over a list of batched variants for one precision: NAME##_blocks (resident work
// groups per compute unit, registers and LDS permitting) and NAME##_launch.  The translation unit
// that invokes it defines QT_TBL_F32, QT_TBL_TD and QT_TBL_GSL (the kernels' F32, TD and GSL
// arguments); each (picture, controller, precision) table has a unit of its own, so that the
// heavy instantiations compile in parallel.
#define QT_BATCH_TABLE(NAME, LIST)                                                                                 \
  int NAME##_blocks(int rows, int nw, int wh, int minw) {                                                          \
    int n = 0;                                                                                                     \
    LIST(QT_TABLE_B)                                                                                               \
    return 0;                                                                                                      \
  }                                                                                                                \
  hipError_t NAME##_launch(int rows, int nw, int wh, int minw, int grid, const SysDev& S, const EnsDev& E,         \
                           hipStream_t st) {                                                                       \
    LIST(QT_TABLE_L)                                                                                               \
    return hipErrorInvalidValue;                                                                                   \
  }
#define QT_TABLE_B(R, N, W, M)                                                                                     \
  if (rows == R && nw == N && wh == W && minw == M) {                                                              \
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(                                                            \
        &n, mcwf_batch_kernel<R, N, W, QT_TBL_F32, M, QT_TBL_TD, QT_TBL_GSL>, 64 * N, 0);                          \
    return n;                                                                                                      \
  }
#define QT_TABLE_L(R, N, W, M)                                                                                     \
  if (rows == R && nw == N && wh == W && minw == M) {                                                              \
    hipLaunchKernelGGL((mcwf_batch_kernel<R, N, W, QT_TBL_F32, M, QT_TBL_TD, QT_TBL_GSL>), dim3(grid), dim3(64 * N), \
                       0, st, S, E);                                                                               \
    return hipGetLastError();                                                                                      \
  }

#define QT_TABLE_DECL(NAME)                                                                                        \
  int NAME##_blocks(int rows, int nw, int wh, int minw);                                                           \
  hipError_t NAME##_launch(int rows, int nw, int wh, int minw, int grid, const SysDev& S, const EnsDev& E,         \
                           hipStream_t st);
#define QT_TABLE_DECL2(NAME) QT_TABLE_DECL(NAME##_f32) QT_TABLE_DECL(NAME##_f64)
QT_TABLE_DECL2(batch_sch)
QT_TABLE_DECL2(batch_td)
QT_TABLE_DECL2(batch_gsl)
QT_TABLE_DECL2(batch_gsl_td)

}  // namespace qt
