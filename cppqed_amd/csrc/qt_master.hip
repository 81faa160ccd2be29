// qt_master.hip -- batched Master-equation sweep on FP64 MFMA (gfx950).
//
// One work group of eight wavefronts per sweep point (the staged matrix fills most of a
// compute unit's LDS, so a point's work group runs alone on its CU; eight wavefronts keep
// more HBM loads in flight than four: 468 k -> 651 k tries/s on 512 points at d = 80).  The
// density operator rho (d x d complex FP64, d padded to DP = 16 NB) is cut into NB x NB tiles
// of 16 x 16; tile t = r NB + c belongs to wavefront t % 8, which accumulates its block of
// M = (H_nH/i) S with v_mfma_f64_16x16x4_f64 (four real MFMAs per complex 16x16x4 block):
// the A operand is the H_nH/i row block r (read from the point's fragment-ordered copy,
// L2-resident), the B operand the RK stage matrix S staged in LDS as real / imaginary
// planes.  Then, as Master::step's derivative (Master.h:157-184),
//     drho = M + M^dagger + sum_m J_m S J_m^dagger,
// formed as N = M + J S J^dagger / 2 (the jump term is Hermitian) and N + N^dagger through an
// LDS transpose.  Step-size control is ode_engine::Base::step with Boost.Odeint's RKCK54
// controller (ODE.h:234-262); after each step rho is made Hermitian (2 Re, CMatrix.cc:14-30)
// and renormalised by its trace (Master.h:196-206).
//
// MFMA lane layout (measured on gfx950): A[i][k] in lane i + 16k, B[k][j] in lane j + 16k,
// D[i][j] in lane 16 (i % 4) + j, component i / 4.  Lane l of the tile's wavefront owns the
// tile elements (4v + l/16, l%16), v < 4; matrices in HBM are stored tile by tile in that
// order: index (4 t + v) 64 + l (rho), or (4 u + v) 64 + l over the lower tiles u = r (r + 1) / 2
// + c, r >= c, only (the Hermitian RK slots).

#include "qt_device.hpp"
#include "qt_kernels.hpp"
#include "qt_master_dev.hpp"
#include "qt_master_common.hpp"

#include <algorithm>

namespace qt {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int kMasterWaves = 8;

// one entry of the remainder as kept in LDS (the point's copy, master_kernel)
struct RemEntry { double2 v; int c, f; };

// K = M + M^dagger + sum_m J_m S J_m^dagger for the stage matrix S in LDS (Sr, Si planes); kept out
// of line so that its three call sites (a step's first derivative, the stage loop, the stream
// point) share one register allocation (inlined, the 512-point d = 80 sweep ran 632 k instead of
// 654 k tries/s).
// TD: M = E (H_nH/i) E^-1 S with ph[k] = exp(-exz_k tau) (E^-1, applied to the B operand as it
// is read) and phi[i] = exp(exz_i tau) (E, applied to the product's rows); the tables are made by
// the caller.
template <int NB, bool TD>
__device__ __noinline__ void master_apply(const MasterDev& M, int p, double* Sr, double* Si, double2* Kout,
                                          const double2* ph, const double2* phi, const double2* rph,
                                          const RemEntry* remL, const RemEntry* jmpL) {
  constexpr int DP = 16 * NB, KC = DP / 4, LD = DP + 1, NT = NB * NB;
  constexpr int TW = (NT + kMasterWaves - 1) / kMasterWaves;
  const int tid = threadIdx.x, lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);  // the wavefront's index (uniform)
  const int d = M.dim;
  const double* const hf = M.hfrag + (M.shared_h ? (size_t)0 : (size_t)p * NB * KC * 64 * 2);
  // the k-steps whose A fragment is non-zero, per tile (a zero fragment adds exact zeros: skipping
  // its load and MFMAs leaves the product bit for bit the same)
  const unsigned* const hm = M.hmask + (M.shared_h ? (size_t)0 : (size_t)p * NB);
  unsigned mk[TW];
#pragma unroll
  for (int tw = 0; tw < TW; ++tw) {
    const int t = w + kMasterWaves * tw;
    mk[tw] = t < NT ? hm[t / NB] : 0u;
  }
  double* const SrL = Sr + (lane >> 4) * LD + (lane & 15);
  double* const SiL = Si + (lane >> 4) * LD + (lane & 15);
  const double* const SrT = Sr + (lane & 15) * LD + (lane >> 4);
  const double* const SiT = Si + (lane & 15) * LD + (lane >> 4);
  d4 cr[TW], ci[TW];
#pragma unroll
  for (int tw = 0; tw < TW; ++tw) { cr[tw] = d4{0., 0., 0., 0.}; ci[tw] = d4{0., 0., 0., 0.}; }
  // M = (H_nH/i) S on FP64 MFMA.  The A fragments come from global memory (L2 / MALL): each
  // k-step's are loaded one step ahead (two named buffers, KC is even), so their latency hides
  // behind the MFMAs of the step before instead of stalling every step.
  auto loadA = [&](int kk, double2 (&a)[TW]) {
#pragma unroll
    for (int tw = 0; tw < TW; ++tw) {
      const int t = w + kMasterWaves * tw;
      if (t >= NT || kk >= KC || !((mk[tw] >> kk) & 1u)) { a[tw] = make_double2(0., 0.); continue; }
      const double* q = hf + ((size_t)((t / NB) * KC + kk) * 64 + lane) * 2;
      a[tw] = make_double2(q[0], q[1]);
    }
  };
  auto mfmaStep = [&](int kk, const double2 (&a)[TW]) {
#pragma unroll
    for (int tw = 0; tw < TW; ++tw) {
      const int t = w + kMasterWaves * tw;
      if (t >= NT || !((mk[tw] >> kk) & 1u)) continue;
      const int c = t % NB;
      const double ar = a[tw].x, ai = a[tw].y;
      double br = SrL[4 * kk * LD + 16 * c], bi = SiL[4 * kk * LD + 16 * c];
      if constexpr (TD) {  // B[k][j] = exp(-exz_k tau) S[k][j], k = 4 kk + lane / 16
        const double2 e = ph[4 * kk + (lane >> 4)];
        const double r0 = br;
        br = e.x * r0 - e.y * bi;
        bi = e.x * bi + e.y * r0;
      }
      cr[tw] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, br, cr[tw], 0, 0, 0);
      cr[tw] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, -bi, cr[tw], 0, 0, 0);
      ci[tw] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar, bi, ci[tw], 0, 0, 0);
      ci[tw] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai, br, ci[tw], 0, 0, 0);
    }
  };
  static_assert(KC % 2 == 0, "k-steps come in pairs");
  double2 a0[TW], a1[TW];
  loadA(0, a0);
#pragma unroll 1
  for (int kk = 0; kk < KC; kk += 2) {
    loadA(kk + 1, a1);
    mfmaStep(kk, a0);
    loadA(kk + 2, a0);
    mfmaStep(kk + 1, a1);
  }
  if constexpr (TD) {  // M = E M': row i times exp(exz_i tau)
#pragma unroll
    for (int tw = 0; tw < TW; ++tw) {
      const int t = w + kMasterWaves * tw;
      if (t >= NT) continue;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const double2 e = phi[16 * (t / NB) + 4 * v + (lane >> 4)];
        const double xr = cr[tw][v], xi = ci[tw][v];
        cr[tw][v] = e.x * xr - e.y * xi;
        ci[tw][v] = e.x * xi + e.y * xr;
      }
    }
  } else {
    (void)ph; (void)phi;
  }
  // the remainder: M_ij += sum_w rval_iw exp(rfreq tau) S_{rcol_iw, j} (frequencies in the
  // interaction picture only)
  if (M.wr > 0) {
    const size_t rb = (size_t)p * M.wr * d;
#pragma unroll
    for (int tw = 0; tw < TW; ++tw) {
      const int t = w + kMasterWaves * tw;
      if (t >= NT) continue;
      const int j = 16 * (t % NB) + (lane & 15);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = 16 * (t / NB) + 4 * v + (lane >> 4);
        if (i >= d || j >= d) continue;
        double ax = 0., ay = 0.;
        for (int a = 0; a < M.wr; ++a) {
          double2 hv;
          int c, f;
          if (remL) {
            const RemEntry& e = remL[a * d + i];
            hv = e.v; c = e.c; f = e.f;
          } else {
            const size_t o = rb + (size_t)a * d + i;
            hv = M.rval[o]; c = M.rcol[o]; f = M.rfi[o];
          }
          if constexpr (TD) hv = cmul(hv, rph[f]);
          else (void)f;
          const double sx = Sr[c * LD + j], sy = Si[c * LD + j];
          ax += hv.x * sx - hv.y * sy;
          ay += hv.x * sy + hv.y * sx;
        }
        cr[tw][v] += ax;
        ci[tw][v] += ay;
      }
    }
  }
  if constexpr (!TD) (void)rph;
  // N = M + (sum_m J_m S J_m^dagger) / 2, S still in LDS
  if (M.nj > 0) {
#pragma unroll
    for (int tw = 0; tw < TW; ++tw) {
      const int t = w + kMasterWaves * tw;
      if (t >= NT) continue;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = 16 * (t / NB) + 4 * v + (lane >> 4), j = 16 * (t % NB) + (lane & 15);
        if (i >= d || j >= d) continue;
        double ax = 0., ay = 0.;
        for (int m = 0; m < M.nj; ++m) {
          const size_t base = ((size_t)p * M.nj + m) * M.wj * d;
          const int lbase = m * M.wj * d;
          for (int aa = 0; aa < M.wj; ++aa) {
            double2 ja;
            int ca;
            if (jmpL) { const RemEntry& e = jmpL[lbase + aa * d + i]; ja = e.v; ca = e.c; }
            else { ja = M.jval[base + (size_t)aa * d + i]; ca = M.jcol[base + (size_t)aa * d + i]; }
            if (ja.x == 0. && ja.y == 0.) continue;
            for (int b = 0; b < M.wj; ++b) {
              double2 jb;
              int cb;
              if (jmpL) { const RemEntry& e = jmpL[lbase + b * d + j]; jb = e.v; cb = e.c; }
              else { jb = M.jval[base + (size_t)b * d + j]; cb = M.jcol[base + (size_t)b * d + j]; }
              if (jb.x == 0. && jb.y == 0.) continue;
              const double sx = Sr[ca * LD + cb], sy = Si[ca * LD + cb];
              const double ux = ja.x * sx - ja.y * sy, uy = ja.x * sy + ja.y * sx;  // J_ia S_{ca,cb}
              ax += ux * jb.x + uy * jb.y;                                          // ... conj(J_jb)
              ay += uy * jb.x - ux * jb.y;
            }
          }
        }
        cr[tw][v] += 0.5 * ax;
        ci[tw][v] += 0.5 * ay;
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int tw = 0; tw < TW; ++tw) {
    const int t = w + kMasterWaves * tw;
    if (t >= NT) continue;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int o = (16 * (t / NB) + 4 * v) * LD + 16 * (t % NB);
      SrL[o] = cr[tw][v]; SiL[o] = ci[tw][v];
    }
  }
  __syncthreads();
  // drho_ij = N_ij + conj(N_ji), kept for the lower tiles r >= c only (drho is Hermitian: the
  // mirrored entry is the exact conjugate) at lower-tile index u = r (r + 1) / 2 + c
#pragma unroll
  for (int tw = 0; tw < TW; ++tw) {
    const int t = w + kMasterWaves * tw;
    if (t >= NT) continue;
    const int r = t / NB, c = t % NB;
    if (r < c) continue;
    const int u = r * (r + 1) / 2 + c;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int o = 16 * c * LD + 16 * r + 4 * v;
      Kout[(4 * u + v) * 64 + lane] = make_double2(cr[tw][v] + SrT[o], ci[tw][v] - SiT[o]);
    }
  }
}


template <int NB, bool TD>
__global__ void __launch_bounds__(64 * kMasterWaves) master_kernel(MasterDev M) {
  constexpr int DP = 16 * NB, KC = DP / 4, LD = DP + 1, NT = NB * NB, NN = DP * DP;
  constexpr int TW = (NT + kMasterWaves - 1) / kMasterWaves;  // tiles per wavefront (at most)
  __shared__ double Sr[DP * LD];
  __shared__ double Si[DP * LD];
  __shared__ double red[kMasterWaves];
  __shared__ long long cnt[kCounters];
  // interaction picture: exp(-exz tau), exp(exz tau) of the current stage; exp(-exz (t - t0))
  __shared__ double2 ph[TD ? DP : 1], phi[TD ? DP : 1], rph[TD ? kMaxFreqs : 1];
  // the point's remainder when it fits (read by every derivative)
  constexpr int kRemBytes = NB <= 5 ? 8192 : 4096;
  __shared__ RemEntry remS[kRemBytes / sizeof(RemEntry)];
  // the point's jump operators (ELL rows) when they fit, read by every derivative's J S J^dagger
  // (from LDS instead of a chain of L2 round trips per entry)
  constexpr int kJmpBytes = NB <= 5 ? 16384 : 0;
  __shared__ RemEntry jmpS[kJmpBytes ? kJmpBytes / sizeof(RemEntry) : 1];

  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int p = blockIdx.x;
  if (p >= M.n_points) return;
  const int d = M.dim;
  const RemEntry* remL = nullptr;
  if (M.wr > 0 && (size_t)M.wr * d <= kRemBytes / sizeof(RemEntry)) {
    for (int e = tid; e < M.wr * d; e += 64 * kMasterWaves) {
      const size_t o = (size_t)p * M.wr * d + e;
      remS[e] = RemEntry{M.rval[o], M.rcol[o], M.rfi[o]};
    }
    remL = remS;  // ordered before its first use by the derivative's barrier
  }
  const RemEntry* jmpL = nullptr;
  if (kJmpBytes > 0 && M.nj > 0 && (size_t)M.nj * M.wj * d <= kJmpBytes / sizeof(RemEntry)) {
    const int nje = M.nj * M.wj * d;
    for (int e = tid; e < nje; e += 64 * kMasterWaves) {
      const size_t o = (size_t)p * nje + e;
      jmpS[e] = RemEntry{M.jval[o], M.jcol[o], 0};
    }
    jmpL = jmpS;  // ordered before its first use by the derivative's barrier
  }

  // rho (X) is kept whole, tile by tile; the RK slots K1..K6 and the step's result XN hold only
  // the NL lower tiles (r >= c), in lower-tile order u = r (r + 1) / 2 + c: they are exactly
  // Hermitian (drho = N + N^dagger entry by entry, and the stage and step sums act entry by entry
  // with real weights), so the upper tiles are the mirrored conjugates and need not be stored or
  // read -- 40% less slot traffic per try.  rho must be Hermitian (the caller's rho0 and every
  // state the kernel writes are).
  constexpr int NL = NB * (NB + 1) / 2;
  constexpr int UW = (NL + kMasterWaves - 1) / kMasterWaves;  // lower tiles per wavefront (at most)
  double2* const X = M.x + (size_t)p * NN;
  double2* const K = M.k + (size_t)p * 7 * NN;
  double2* const K1 = K;
  double2* const K2 = K + NN;
  double2* const K3 = K + 2 * NN;
  double2* const K4 = K + 3 * NN;
  double2* const K5 = K + 4 * NN;
  double2* const K6 = K + 5 * NN;
  double2* const XN = K + 6 * NN;

  // element e = (tw, v) of this lane: tile t = w + 8 tw, row 16 r + 4 v + lane/16, column 16 c + lane%16
  auto tileOf = [&](int tw) { return w + kMasterWaves * tw; };
  auto has = [&](int tw) { return tileOf(tw) < NT; };
  auto gidx = [&](int tw, int v) { return (4 * tileOf(tw) + v) * 64 + lane; };
  auto rowOf = [&](int tw, int v) { return 16 * (tileOf(tw) / NB) + 4 * v + (lane >> 4); };
  auto colOf = [&](int tw) { return 16 * (tileOf(tw) % NB) + (lane & 15); };
  double* const SrL = Sr + (lane >> 4) * LD + (lane & 15);   // + (16 r + 4 v) LD + 16 c
  double* const SiL = Si + (lane >> 4) * LD + (lane & 15);
  const double* const SrT = Sr + (lane & 15) * LD + (lane >> 4);  // + 16 c LD + 16 r + 4 v
  const double* const SiT = Si + (lane & 15) * LD + (lane >> 4);
  auto lofs = [&](int tw, int v) { const int t = tileOf(tw); return (16 * (t / NB) + 4 * v) * LD + 16 * (t % NB); };
  auto tofs = [&](int tw, int v) { const int t = tileOf(tw); return 16 * (t % NB) * LD + 16 * (t / NB) + 4 * v; };
  // lower tile u = w + 8 uw of this wavefront, its row and column block (lr, lc)
  int lr[UW], lc[UW];
#pragma unroll
  for (int uw = 0; uw < UW; ++uw) {
    const int u = w + kMasterWaves * uw;
    int r = 0;
    while ((r + 1) * (r + 2) / 2 <= u) ++r;
    lr[uw] = r;
    lc[uw] = u - r * (r + 1) / 2;
  }
  auto hasL = [&](int uw) { return w + kMasterWaves * uw < NL; };
  auto lidx = [&](int uw, int v) { return (4 * (w + kMasterWaves * uw) + v) * 64 + lane; };       // slots
  auto xidx = [&](int uw, int v) { return (4 * (lr[uw] * NB + lc[uw]) + v) * 64 + lane; };     // rho
  // LDS offsets of a lower element (i, j) (with SrL) and of its mirror (j, i) (with SrM)
  double* const SrM = Sr + (lane & 15) * LD + (lane >> 4);
  double* const SiM = Si + (lane & 15) * LD + (lane >> 4);
  auto llofs = [&](int uw, int v) { return (16 * lr[uw] + 4 * v) * LD + 16 * lc[uw]; };
  auto lmofs = [&](int uw, int v) { return 16 * lc[uw] * LD + 16 * lr[uw] + 4 * v; };

  auto bsum = [&](double v) {
    v = wave_sum(v);
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    double s = red[0];
#pragma unroll
    for (int i = 1; i < kMasterWaves; ++i) s += red[i];
    return s;
  };
  auto bmax = [&](double v) {
    v = wave_max(v);
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    double s = red[0];
#pragma unroll
    for (int i = 1; i < kMasterWaves; ++i) s = fmax(s, red[i]);
    return s;
  };

  const double2* const exzP = TD ? M.exz + (size_t)p * d : nullptr;
  // the phase tables of picture time tau (interaction picture); the caller's barrier orders them
  auto setPhases = [&](double tau) {
    if constexpr (TD) {
      for (int i = tid; i < DP; i += 64 * kMasterWaves) {
        double2 a = make_double2(1., 0.), b = make_double2(1., 0.);
        if (i < d) {
          const double2 z = exzP[i];
          double sn, cs;
          sincos(-z.y * tau, &sn, &cs);
          const double e = exp(-z.x * tau);
          a = make_double2(e * cs, e * sn);                 // exp(-z tau)
          const double f = exp(z.x * tau);
          b = make_double2(f * cs, -f * sn);                // exp(z tau)
        }
        ph[i] = a; phi[i] = b;
      }
      if (M.wr > 0 && tid < M.nrf) {  // exp(f tau) of the remainder's frequencies
        const double2 f = M.rfreq[(size_t)p * M.nrf + tid];
        double sn, cs;
        sincos(f.y * tau, &sn, &cs);
        const double e = exp(f.x * tau);
        rph[tid] = make_double2(e * cs, e * sn);
      }
    } else {
      (void)tau;
    }
  };
  // Kout = F(S) at picture time tau; S given on the lower tiles by stage(uw, sv), the four entries
  // sv[v] of this lane in lower tile uw, and mirrored (conjugated) into the upper tiles
  auto derivative = [&](auto&& stage, double2* Kout, double tau) {
    __syncthreads();
    setPhases(tau);
    for (int uw = 0; uw < UW; ++uw) {
      if (!hasL(uw)) break;
      const bool mirror = lr[uw] != lc[uw];
      double2 sv[4];
      stage(uw, sv);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        SrL[llofs(uw, v)] = sv[v].x; SiL[llofs(uw, v)] = sv[v].y;
        if (mirror) { SrM[lmofs(uw, v)] = sv[v].x; SiM[lmofs(uw, v)] = -sv[v].y; }
      }
    }
    __syncthreads();
    master_apply<NB, TD>(M, p, Sr, Si, Kout, ph, phi, rph, remL, jmpL);
  };
  auto ld = [&](const double2* A, int tw, int v) { return A[gidx(tw, v)]; };
  auto ldX = [&](int uw, int v) { return X[xidx(uw, v)]; };
  auto stageX = [&](int uw, double2 (&sv)[4]) {
#pragma unroll
    for (int v = 0; v < 4; ++v) sv[v] = ldX(uw, v);
  };

  double t = M.t[p], dtDid = M.dtDid[p], dtTry = M.dtTry[p];
  if (tid < kCounters) cnt[tid] = M.counters[(size_t)p * kCounters + tid];
  __syncthreads();
  int done = M.done[p];
  __shared__ int asStop;

  for (int kI = 0; kI < M.n_intervals; ++kI) {
    // Dt mode: advance(intervals[kI]); dc mode: up to dc steps step(t_end - t) per stream point
    const double endTime = M.dc > 0 ? M.t_end : t + M.intervals[kI];
    const int maxSteps = M.dc > 0 ? M.dc : -1;
    if (!done) {
      for (int steps = 0;; ++steps) {
        const double Dt = endTime - t;
        if (Dt == 0. || steps == maxSteps) break;
        // Master::step -> ode_engine::Base::step(Dt); the pictures coincide at the step's start (t0 = t)
        const double timeBefore = t;
        const double t0 = t;
        const double nextDtTry = (fabs(Dt) < fabs(dtTry / 10.) ? dtTry : 0.);
        if ((Dt > 0.) != (dtTry > 0.)) dtTry = -dtDid;
        if (fabs(dtTry) > fabs(Dt)) dtTry = Dt;
        derivative(stageX, K1, t - t0);  // k1 at (x, t); unchanged on retries
        long long tries = 0, failed = 0;
        for (;;) {
          ++tries;
          const double dt = dtTry;
          // stages 2..6: S = X + dt sum_{i<s} a_si K_i, accumulated in the tableau's order
#pragma unroll 1
          for (int st = 1; st < 6; ++st) {
            // each slot's four entries are loaded together, then added in the tableau's order
            derivative([&](int uw, double2 (&sv)[4]) {
              stageX(uw, sv);
              for (int i = 0; i < st; ++i) {
                const double c = ck_a(st, i) * dt;
                const double2* const Ki = K + (size_t)i * NN;
                double2 kv[4];
#pragma unroll
                for (int v = 0; v < 4; ++v) kv[v] = Ki[lidx(uw, v)];
#pragma unroll
                for (int v = 0; v < 4; ++v) sv[v] = axpy(c, kv[v], sv[v]);
              }
            }, K + (size_t)st * NN, (t + ck_c(st) * dt) - t0);
          }
          // the error's maximum over the lower tiles is the whole matrix's (mirrored entries have
          // the same moduli).  The barrier orders K6's tiles, written by the wavefronts that own
          // the full tiles, before these reads by the owners of the lower tiles.
          __syncthreads();
          double emax = 0.;
#pragma unroll 1
          for (int uw = 0; uw < UW; ++uw) {
            if (!hasL(uw)) break;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
              const int li = lidx(uw, v);
              const double2 x = ldX(uw, v), k1 = K1[li], k3 = K3[li], k4 = K4[li], k5 = K5[li], k6 = K6[li];
              XN[li] = axpy(CK::b6 * dt, k6, axpy(CK::b4 * dt, k4, axpy(CK::b3 * dt, k3, axpy(CK::b1 * dt, k1, x))));
              double2 xe = make_double2(CK::e1 * dt * k1.x, CK::e1 * dt * k1.y);
              xe = axpy(CK::e6 * dt, k6, axpy(CK::e5 * dt, k5, axpy(CK::e4 * dt, k4, axpy(CK::e3 * dt, k3, xe))));
              const double D = M.epsAbs + M.epsRel * (sqrt(cabs2(x)) + dt * sqrt(cabs2(k1)));
              emax = fmax(emax, cabs2(xe) / (D * D));
            }
          }
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
        __syncthreads();
        setPhases(t - t0);  // ph = u
        __syncthreads();
        for (int uw = 0; uw < UW; ++uw) {
          if (!hasL(uw)) break;
          const bool mirror = lr[uw] != lc[uw];
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            // entry (i, j) of XN and its mirror (j, i) = conj, each transformed as its own entry
            const double2 xn = XN[lidx(uw, v)];
            double2 y = xn, ym = make_double2(xn.x, -xn.y);
            if constexpr (TD) {
              const int i = 16 * lr[uw] + 4 * v + (lane >> 4), j = 16 * lc[uw] + (lane & 15);
              if (i < d && j < d) {
                y = cmul(cmul(ph[i], y), make_double2(ph[j].x, -ph[j].y));
                ym = cmul(cmul(ph[j], ym), make_double2(ph[i].x, -ph[i].y));
              }
            }
            SrL[llofs(uw, v)] = y.x; SiL[llofs(uw, v)] = y.y;
            if (mirror) { SrM[lmofs(uw, v)] = ym.x; SiM[lmofs(uw, v)] = ym.y; }
          }
        }
        __syncthreads();
        double tr = 0.;
        for (int tw = 0; tw < TW; ++tw) {
          if (!has(tw)) break;
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int i = rowOf(tw, v), j = colOf(tw);
            if (i == j && i < d) tr += 2. * SrL[lofs(tw, v)];
          }
        }
        const double inv = 1.0 / bsum(tr);
        for (int tw = 0; tw < TW; ++tw) {
          if (!has(tw)) break;
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int i = rowOf(tw, v), j = colOf(tw);
            const double xr = SrL[lofs(tw, v)], xi = SiL[lofs(tw, v)], yr = SrT[tofs(tw, v)], yi = SiT[tofs(tw, v)];
            double2 y;
            if (i == j) y = make_double2(2. * xr, 0.);
            else if (i < j) y = make_double2(xr + yr, xi - yi);
            else y = make_double2(yr + xr, -(yi - xi));
            X[gidx(tw, v)] = make_double2(y.x * inv, y.y * inv);
          }
        }
        if (tid == 0) {
          cnt[kMcwfSteps] += 1; cnt[kOdeSteps] += 1; cnt[kTries] += tries; cnt[kFailed] += failed;
          cnt[kDerivs] += 6 * tries;
        }
      }
    }
    // stream point: averages Tr(rho O_q), ||drho/dt||_F, steady-state check
    __syncthreads();
    for (int tw = 0; tw < TW; ++tw) {
      if (!has(tw)) break;
#pragma unroll
      for (int v = 0; v < 4; ++v) { const double2 y = ld(X, tw, v); SrL[lofs(tw, v)] = y.x; SiL[lofs(tw, v)] = y.y; }
    }
    __syncthreads();
    for (int q = 0; q < M.no; ++q) {
      double part = 0.;
      const bool im = M.oimag[q] != 0;
      for (int i = tid; i < d; i += 64 * kMasterWaves) {
        for (int a = 0; a < M.owid[q]; ++a) {
          const double2 o = M.oval[((size_t)q * M.wo + a) * d + i];
          const int cidx = M.ocol[((size_t)q * M.wo + a) * d + i];
          // Tr(rho O) = sum_i sum_a O_{i,c} rho_{c,i}
          const double rr = Sr[cidx * LD + i], ri = Si[cidx * LD + i];
          part += im ? (o.x * ri + o.y * rr) : (o.x * rr - o.y * ri);
        }
      }
      const double v = bsum(part);
      if (tid == 0) M.obs_out[((size_t)p * M.n_intervals + kI) * M.no + q] = v;
    }
    derivative(stageX, K1, 0.);  // t0 = t at a stream point
    __syncthreads();  // K1's tiles come from other wavefronts (as K6's before the error estimate)
    double n2 = 0.;
    for (int uw = 0; uw < UW; ++uw) {
      if (!hasL(uw)) break;
      const double f = lr[uw] == lc[uw] ? 1. : 2.;  // an off-diagonal tile stands for its mirror too
#pragma unroll
      for (int v = 0; v < 4; ++v) n2 += f * cabs2(K1[lidx(uw, v)]);
    }
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

bool master_supported(int nb) { return nb >= 1 && nb <= kMasterMaxNB; }

template <bool TD>
static hipError_t launch_master_t(int nb, const MasterDev& M, hipStream_t st) {
  const dim3 block(64 * kMasterWaves);
  switch (nb) {
    case 1: hipLaunchKernelGGL((master_kernel<1, TD>), dim3(M.n_points), block, 0, st, M); break;
    case 2: hipLaunchKernelGGL((master_kernel<2, TD>), dim3(M.n_points), block, 0, st, M); break;
    case 3: hipLaunchKernelGGL((master_kernel<3, TD>), dim3(M.n_points), block, 0, st, M); break;
    case 4: hipLaunchKernelGGL((master_kernel<4, TD>), dim3(M.n_points), block, 0, st, M); break;
    case 5: hipLaunchKernelGGL((master_kernel<5, TD>), dim3(M.n_points), block, 0, st, M); break;
    case 6: hipLaunchKernelGGL((master_kernel<6, TD>), dim3(M.n_points), block, 0, st, M); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_master(int nb, const MasterDev& M, hipStream_t st) {
  return M.exz ? launch_master_t<true>(nb, M, st) : launch_master_t<false>(nb, M, st);
}

}  // namespace qt
