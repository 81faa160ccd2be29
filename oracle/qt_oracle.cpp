// qt_oracle.cpp -- CPU restatement of C++QED's quantum-trajectory hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
// only as the parity reference (or the timed CPU baseline), never as the product path.
//
// It restates, in plain FP64 C++ written in the reference's own per-slice
// style (not in the product's flattened-table style), the following reference
// code (paths relative to the reference repository):
//   Tridiagonal algebra ........ CPPQEDcore/quantumoperator/Tridiagonal.{h,tcc,cc}
//   TridiagonalHamiltonian ..... CPPQEDcore/quantumoperator/TridiagonalHamiltonian.h:42-47
//   slice embedding ............ CPPQEDutils/SliceIterator.h:277-354, MultiIndexIterator.h:90-98
//   BinarySystem / Composite ... CPPQEDcore/composites/BinarySystem.cc:141-273, Composite.tcc:156-333
//   Mode / Qbit / PTLA / JC .... CPPQEDelements/frees/{Mode.cc,Mode_.h,Mode.tcc,Qbit.cc,Qbit_.h,
//                                PumpedTwoLevelAtom.cc}, interactions/JaynesCummings_.h:54-66
//   RKCK54 + odeint controller . CPPQEDutils/ODE.h:189-262 (Boost.Odeint controlled_runge_kutta,
//                                runge_kutta_cash_karp54, default_error_checker, default_step_adjuster)
//   GSL controller ............. CPPQEDutils/ODE_GSL.h:44-84 (GSL rkck, control_standard, hadjust);
//                                pinned only by the reference's GSL tolerance tests
//   Simulated trajectory ....... CPPQEDutils/Simulated_.h:26-79 (the stepper tests' classical ODEs)
//   MCWF step .................. CPPQEDcore/quantumtrajectory/MCWF_Trajectory.h:103-118, 250-374
//   ensemble ................... MCWF_Trajectory.h:213-234, 406-420; CPPQEDutils/StochasticTrajectory.h:220
//   Master step ................ CPPQEDcore/quantumtrajectory/Master.h:152-208; CPPQEDutils/CMatrix.cc:14-30
//   run loop ................... CPPQEDutils/Trajectory.h:43-50, 462-610 (incl. continuing from a state file)
//   RNG ........................ CPPQEDutils/Random.h:117-161 (pcg64) + libstdc++ generate_canonical;
//                                Philox4x64-10 (Salmon et al., SC'11) as the counter-based generator.
//
// Each function below cites the reference lines it follows.

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace orc {

using cd = std::complex<double>;
using CV = std::vector<cd>;

static thread_local std::string g_last_error;

// ---------------------------------------------------------------------------
// Random numbers
// ---------------------------------------------------------------------------

// pcg64 = pcg_engines::setseq_xsl_rr_128_64 (O'Neill 2014), constructed as
// pcg64{seed, stream}: inc = 2*stream+1, state = bump(seed+inc), and every
// call advances the LCG first and outputs the XSL-RR of the new state.
// Reference: CPPQEDutils/Random.h:117-126 (EngineWithParameters{seed,stream}).
struct Pcg64 {
  unsigned __int128 state, inc;
  static unsigned __int128 mult() {
    return (((unsigned __int128)2549297995355413924ULL) << 64) | 4865540595714422341ULL;
  }
  Pcg64(uint64_t seed, uint64_t stream) {
    inc = (((unsigned __int128)stream) << 1) | 1u;
    state = ((unsigned __int128)seed + inc) * mult() + inc;
  }
  uint64_t next() {
    state = state * mult() + inc;
    uint64_t xored = (uint64_t)(state >> 64) ^ (uint64_t)state;
    unsigned rot = (unsigned)(state >> 122);
    return (xored >> rot) | (xored << ((64u - rot) & 63u));
  }
  // std::uniform_real_distribution<double>{0,1} with libstdc++'s
  // generate_canonical<double,53>: one 64-bit draw, double(x)/2^64, clamped
  // below 1 (MCWF_Trajectory.h:161,191).
  double uniform() {
    double r = (double)next() / 18446744073709551616.0;
    if (r >= 1.0) r = std::nextafter(1.0, 0.0);
    return r;
  }
};

// Philox4x64-10 (Random123).  The counter-based generator shared with the
// device code: key = (seed, trajectory), counter = (MCWF step ordinal,0,0,0).
static inline void mulhilo64(uint64_t a, uint64_t b, uint64_t& hi, uint64_t& lo) {
  unsigned __int128 p = (unsigned __int128)a * b;
  hi = (uint64_t)(p >> 64);
  lo = (uint64_t)p;
}

void philox4x64_10(const uint64_t ctr_in[4], const uint64_t key_in[2], uint64_t out[4]) {
  uint64_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
  uint64_t k0 = key_in[0], k1 = key_in[1];
  for (int r = 0; r < 10; ++r) {
    if (r) { k0 += 0x9E3779B97F4A7C15ULL; k1 += 0xBB67AE8584CAA73BULL; }
    uint64_t hi0, lo0, hi1, lo1;
    mulhilo64(0xD2E7470EE14C6C93ULL, c0, hi0, lo0);
    mulhilo64(0xCA5A826395121157ULL, c2, hi1, lo1);
    uint64_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// u in [0,1) with 53 random bits from the first Philox word.
double counter_uniform(uint64_t seed, uint64_t traj, uint64_t step) {
  uint64_t ctr[4] = {step, 0, 0, 0}, key[2] = {seed, traj}, out[4];
  philox4x64_10(ctr, key, out);
  return (double)(out[0] >> 11) * (1.0 / 9007199254740992.0);
}

// ---------------------------------------------------------------------------
// Tridiagonal<RANK>   (Tridiagonal.h:71-287, Tridiagonal.tcc, Tridiagonal.cc)
// ---------------------------------------------------------------------------

static int pow3(int r) { int p = 1; while (r--) p *= 3; return p; }

struct Tridiag {
  std::vector<int> dims, K;   // per-axis dimension and offset K
  std::vector<CV> diag, freq; // 3^rank diagonals (row-major, axis 0 slowest), freqs
  double tCurrent = 0.;
  int rank() const { return (int)dims.size(); }
  int length() const { return pow3(rank()); }
  // per-axis extent of diagonal `ind` (digit 0: N, digit 1/2: N-K)
  std::vector<int> shape(int ind) const {
    std::vector<int> s(rank());
    for (int a = rank() - 1; a >= 0; --a) { int dgt = ind % 3; ind /= 3; s[a] = dgt ? dims[a] - K[a] : dims[a]; }
    return s;
  }
};

// Unary ctor, Tridiagonal.cc:39-60
Tridiag unary(const CV& zero, int k, const CV& minus, const CV& plus, bool toFreqs = false) {
  Tridiag t;
  int n = std::max(std::max((int)zero.size(), (int)minus.size() + k), (int)plus.size() + k);
  t.dims = {n}; t.K = {k};
  t.diag.assign(3, CV()); t.freq.assign(3, CV());
  t.diag[0] = toFreqs ? CV() : zero; t.diag[1] = minus; t.diag[2] = plus;
  if (toFreqs) {
    CV mz(zero.size()); for (size_t i = 0; i < zero.size(); ++i) mz[i] = -zero[i];
    int ds = n - k;
    if (ds > 0) { t.freq[1].resize(ds); for (int i = 0; i < ds; ++i) t.freq[1][i] = mz[i + k] - mz[i];
                  t.freq[2].resize(ds); for (int i = 0; i < ds; ++i) t.freq[2][i] = -t.freq[1][i]; }
  }
  auto compat = [&](const CV& a, int da, const CV& b, int db) { return !a.size() || !b.size() || (int)a.size() - da == (int)b.size() - db; };
  if (!compat(zero, k, minus, 0) || !compat(zero, k, plus, 0) || !compat(minus, 0, plus, 0) || ((minus.size() || plus.size()) && !k))
    throw std::invalid_argument("Tridiagonal consistency error");
  return t;
}

// furnishWithFreqs, Tridiagonal.cc:24-36
void furnishWithFreqs(Tridiag& t, const CV& main) {
  int k = t.K[0], ds = t.dims[0] - k;
  if (ds > 0) {
    t.freq[1].resize(ds); for (int n = 0; n < ds; ++n) t.freq[1][n] = main[n + k] - main[n];
    t.freq[2].resize(ds); for (int n = 0; n < ds; ++n) t.freq[2][n] = -t.freq[1][n];
  }
}

// doDirect (ComplexArrayExtensions.h:80-95): outer product / outer sum, empty if either is empty
static CV direct(const CV& a, const CV& b, bool mult) {
  CV r;
  if (!a.size() || !b.size()) return r;
  r.resize(a.size() * b.size());
  for (size_t i = 0; i < a.size(); ++i) for (size_t j = 0; j < b.size(); ++j) r[i * b.size() + j] = mult ? a[i] * b[j] : a[i] + b[j];
  return r;
}

// Direct product ctor, Tridiagonal.tcc:28-70
Tridiag product(const Tridiag& a, const Tridiag& b) {
  if (a.tCurrent != b.tCurrent) throw std::runtime_error("Tridiagonal time mismatch");
  Tridiag t;
  t.dims = a.dims; t.dims.insert(t.dims.end(), b.dims.begin(), b.dims.end());
  t.K = a.K; t.K.insert(t.K.end(), b.K.begin(), b.K.end());
  t.tCurrent = a.tCurrent;
  int lb = b.length();
  t.diag.assign(a.length() * lb, CV()); t.freq.assign(a.length() * lb, CV());
  for (int i = 0; i < a.length(); ++i) for (int j = 0; j < lb; ++j) {
    t.diag[i * lb + j] = direct(a.diag[i], b.diag[j], true);
    t.freq[i * lb + j] = direct(a.freq[i], b.freq[j], false);
  }
  return t;
}

// hermitianConjugate, Tridiagonal.tcc:84-115: digit i -> (3-i)%3, conjugate; freqs untouched
Tridiag hc(const Tridiag& t) {
  Tridiag r = t;
  int R = t.rank();
  for (int ind = 0; ind < t.length(); ++ind) {
    int res = 0, rem = ind;
    for (int k = R - 1; k >= 0; --k) { int p = pow3(k); int ik = rem / p; res += ((3 - ik) % 3) * p; rem -= ik * p; }
    CV c(t.diag[ind].size()); for (size_t i = 0; i < c.size(); ++i) c[i] = std::conj(t.diag[ind][i]);
    r.diag[res] = c;
  }
  return r;
}

Tridiag scaled(const Tridiag& t, cd f) { Tridiag r = t; for (auto& d : r.diag) for (auto& x : d) x *= f; return r; }
Tridiag negated(const Tridiag& t) { Tridiag r = t; for (auto& d : r.diag) for (auto& x : d) x = -x; return r; }

// operator+=, Tridiagonal.tcc:118-149
Tridiag& addTo(Tridiag& a, const Tridiag& b) {
  for (int i = 0; i < a.rank(); ++i) {
    if (!a.K[i]) a.K[i] = b.K[i];
    else if (b.K[i] && a.K[i] != b.K[i]) throw std::invalid_argument("Tridiagonal structure mismatch");
  }
  for (size_t i = 0; i < a.diag.size(); ++i) if (b.diag[i].size()) {
    if (!a.diag[i].size()) a.diag[i] = b.diag[i];
    else { if (a.diag[i].size() != b.diag[i].size()) throw std::invalid_argument("shape mismatch");
           for (size_t j = 0; j < a.diag[i].size(); ++j) a.diag[i][j] += b.diag[i][j]; }
  }
  for (size_t i = 0; i < a.freq.size(); ++i) if (b.freq[i].size()) a.freq[i] = b.freq[i];
  return a;
}
Tridiag sum(Tridiag a, const Tridiag& b) { return addTo(a, b); }
Tridiag minusHC(const Tridiag& t) { return sum(t, negated(hc(t))); } // tridiagMinusHC, Tridiagonal.h:333-336

// propagate, Tridiagonal.tcc:153-163 (cumulative, multiplicative)
void propagate(Tridiag& t, double tt) {
  double dt = tt - t.tCurrent;
  if (dt != 0.)
    for (int i = 1; i < t.length(); ++i)
      if (t.diag[i].size() && t.freq[i].size())
        for (size_t j = 0; j < t.diag[i].size(); ++j) t.diag[i][j] *= std::exp(dt * t.freq[i][j]);
  t.tCurrent = tt;
}

// apply on a strided slice view, Tridiagonal.tcc:193-263, 332-346:
// digit 0: dpsi[n] += a[n] psi[n]; digit 1: dpsi[n+K] += a[n] psi[n]; digit 2: dpsi[n] += a[n] psi[n+K]
void apply(const Tridiag& t, const cd* psi, cd* dpsidt, const std::vector<long>& strides) {
  int R = t.rank();
  std::vector<int> digits(R), sh, a(R);
  for (int ind = 0; ind < t.length(); ++ind) {
    const CV& dg = t.diag[ind];
    if (!dg.size()) continue;
    int rem = ind;
    for (int ax = R - 1; ax >= 0; --ax) { digits[ax] = rem % 3; rem /= 3; }
    sh = t.shape(ind);
    std::fill(a.begin(), a.end(), 0);
    for (size_t lin = 0; lin < dg.size(); ++lin) {
      long oo = 0, ii = 0;
      for (int ax = 0; ax < R; ++ax) {
        int o = a[ax], i = a[ax];
        if (digits[ax] == 1) o += t.K[ax]; else if (digits[ax] == 2) i += t.K[ax];
        oo += o * strides[ax]; ii += i * strides[ax];
      }
      dpsidt[oo] += dg[lin] * psi[ii];
      for (int ax = R - 1; ax >= 0; --ax) { if (++a[ax] < sh[ax]) break; a[ax] = 0; }
    }
  }
}

// ---------------------------------------------------------------------------
// Slice iteration (SliceIterator.h:277-354): retained axes V (in V's order);
// the dummy axes are enumerated odometer-style, last index fastest.
// ---------------------------------------------------------------------------

struct Layout {
  std::vector<int> dims;
  std::vector<long> strides;
  long total = 1;
  explicit Layout(const std::vector<int>& d = {}) : dims(d), strides(d.size()) {
    total = 1;
    for (int i = (int)d.size() - 1; i >= 0; --i) { strides[i] = total; total *= d[i]; }
  }
};

template <typename F>
void forSlices(const Layout& L, const std::vector<int>& V, F&& f) {
  int R = (int)L.dims.size();
  std::vector<int> dummy;
  for (int i = 0; i < R; ++i) if (std::find(V.begin(), V.end(), i) == V.end()) dummy.push_back(i);
  std::vector<long> vs(V.size());
  for (size_t i = 0; i < V.size(); ++i) vs[i] = L.strides[V[i]];
  std::vector<int> idx(dummy.size(), 0);
  while (true) {
    long off = 0;
    for (size_t i = 0; i < dummy.size(); ++i) off += idx[i] * L.strides[dummy[i]];
    f(off, vs);
    int k = (int)dummy.size() - 1;
    for (; k >= 0; --k) { if (++idx[k] < L.dims[dummy[k]]) break; idx[k] = 0; }
    if (k < 0) break;
  }
}

// ---------------------------------------------------------------------------
// Elements
// ---------------------------------------------------------------------------

enum class Picture { Sch, IP, UIP };

struct Jump {
  // acts in place on a strided unary (or RANK-ary) slice
  std::function<void(cd* psi, long stride)> act;
  // rate from the diagonal populations of the slice: sum_n w[n] |psi_n|^2
  CV rateWeights;              // real weights (imag ignored)
  // superoperator on a unary density-operator slice (rows/cols strides); null = not implemented
  std::function<void(const cd* rho, cd* drho, long rs, long cs)> superop;
};

struct Free {
  std::string kind;
  int dim = 0;
  // Hamiltonian: list of tridiagonals (TridiagonalHamiltonian) or a dense CMatrix
  bool hasHa = false, timeDep = false;
  std::vector<Tridiag> hOverIs;
  CV cmatrix;                  // dim*dim, row-major (PTLA Sch)
  // Exact: diagonal U = exp(-exd[n] (t-t0)), exd = zI n for modes (Mode.cc:42-45, FreeExact.h:31),
  // i omrec k^2 for the particle (Particle.cc:29-48); the same diagonal furnishes the frequencies
  bool hasEx = false; cd zI = 0.; CV exd;
  // pumped particle
  double vClass = 0.; std::string mf; int kPart = 0;
  // Liouvillean
  std::vector<Jump> jumps;
  // Averaged: "mode" (n, VAR n, Re a, Im a) or "qbit" (rho00, rho11, Re rho10, Im rho10)
  std::string avKind;
  std::vector<std::pair<double, double>> realFreqs;   // (value, multiplier)
  std::vector<std::pair<cd, double>> complexFreqs;
  double highestFrequency() const {
    double m = 0;
    for (auto& f : realFreqs) m = std::max(m, std::fabs(f.first * f.second));
    for (auto& f : complexFreqs) m = std::max(m, std::abs(f.first * f.second));
    return m;
  }
  int nAvr() const { return avKind.empty() ? 0 : 4; }
};

struct Interaction {
  std::string kind;
  std::vector<int> axes;       // the V of Act<...> (BinarySystem: <0,1>)
  bool hasHa = false, timeDep = false;
  std::vector<Tridiag> hOverIs;
  // chained terms (first, second): dpsidt += (1 (x) second) (first psi), second acting on the
  // interaction's last axis (ParticleAlongCavity::addContribution_v, ParticleCavity.cc:113-140)
  std::vector<std::pair<Tridiag, Tridiag>> chained;
  std::vector<std::pair<cd, double>> complexFreqs;
  std::vector<std::pair<double, double>> realFreqs;
  double highestFrequency() const {
    double m = 0;
    for (auto& f : realFreqs) m = std::max(m, std::fabs(f.first * f.second));
    for (auto& f : complexFreqs) m = std::max(m, std::abs(f.first * f.second));
    return m;
  }
};

// mode::mainDiagonal, Mode.cc:54-63:  z*n + omegaKerr*(n-1)*n/i
CV mainDiagonal(cd z, double omegaKerr, int dim) {
  CV r(dim);
  for (int n = 0; n < dim; ++n) r[n] = z * (double)n + cd((n - 1.) * n, 0) * (omegaKerr / cd(0, 1));
  return r;
}

// mode::aop, Mode.cc:277-282
Tridiag aop(int dim) {
  CV plus(dim - 1);
  for (int n = 0; n < dim - 1; ++n) plus[n] = std::sqrt(n + 1.);
  return unary(CV(), 1, CV(), plus);
}

// mode::pumping, Mode.cc:66-69:  eta a^dagger - h.c.
Tridiag pumping(cd eta, int dim) { return minusHC(scaled(hc(aop(dim)), eta)); }

static bool isNonZero(cd z) { return z.real() != 0. || z.imag() != 0.; }

// hOverI, Mode.cc:75-86
Tridiag hOverI(cd z, cd eta, double omegaKerr, int dim, bool& empty) {
  bool isPumped = isNonZero(eta);
  empty = false;
  if (isNonZero(z)) {
    Tridiag res = unary(mainDiagonal(-z, omegaKerr, dim), 0, CV(), CV());
    if (isPumped) return sum(res, pumping(eta, dim));
    return res;
  } else if (isPumped) return pumping(eta, dim);
  empty = true;
  Tridiag t = unary(CV(), 0, CV(), CV()); t.dims = {dim};
  return t;
}

// mode::Hamiltonian<IS_TD> ctors, Mode.cc:95-107; Kerr-alter (all zero here) pushed second
void modeHamiltonian(Free& f, bool isTD, cd zSch, cd zI, cd eta, int dim) {
  bool empty;
  Tridiag h = hOverI(zSch, eta, 0., dim, empty);
  if (isTD) { furnishWithFreqs(h, mainDiagonal(zI, 0., dim)); f.hasEx = true; f.zI = zI; f.exd = mainDiagonal(zI, 0., dim); }
  f.hOverIs.push_back(h);
  f.hOverIs.push_back(unary(mainDiagonal(0., 0., dim), 0, CV(), CV()));  // (n-1)n*0/i == 0
  for (auto& x : f.hOverIs.back().diag[0]) x = cd(0, 0) * x;  // exact zeros
  f.hasHa = true; f.timeDep = isTD;
}

// mode::details::aJump, Mode.cc:118-125; aSuperoperator, Mode.cc:135-140
Jump aJumpOf(double kappa, int dim) {
  Jump j;
  j.act = [kappa, dim](cd* psi, long s) {
    double fact = std::sqrt(2. * kappa);
    int ub = dim - 1;
    for (int n = 0; n < ub; ++n) psi[n * s] = fact * std::sqrt(n + 1.) * psi[(n + 1) * s];
    psi[ub * s] = 0.;
  };
  j.rateWeights.resize(dim);
  for (int n = 0; n < dim; ++n) j.rateWeights[n] = 2 * kappa * n;  // 2 kappa <n> (Mode_.h:180-184)
  j.superop = [kappa, dim](const cd* rho, cd* drho, long rs, long cs) {
    for (int m = 0; m < dim - 1; ++m)
      for (int n = 0; n < dim - 1; ++n)
        drho[m * rs + n * cs] += 2 * kappa * std::sqrt((double)((m + 1) * (n + 1))) * rho[(m + 1) * rs + (n + 1) * cs];
  };
  return j;
}

std::map<std::string, std::string> parseKV(std::istringstream& is) {
  std::map<std::string, std::string> kv; std::string tok;
  while (is >> tok) { auto p = tok.find('='); if (p == std::string::npos) throw std::invalid_argument("bad token " + tok); kv[tok.substr(0, p)] = tok.substr(p + 1); }
  return kv;
}
double num(const std::map<std::string, std::string>& kv, const std::string& k, double def) {
  auto it = kv.find(k); return it == kv.end() ? def : std::stod(it->second);
}
Picture picture(const std::map<std::string, std::string>& kv) {
  auto it = kv.find("picture"); if (it == kv.end() || it->second == "Sch") return Picture::Sch;
  if (it->second == "IP") return Picture::IP;
  if (it->second == "UIP") return Picture::UIP;
  throw std::invalid_argument("bad picture");
}

// mode::make(ParsPumpedLossy, qmp), Mode.tcc:40-60 (nTh = 0, Kerr = 0)
Free makeMode(const std::map<std::string, std::string>& kv) {
  Free f; f.kind = "mode"; f.avKind = "mode";
  int cutoff = (int)num(kv, "cutoff", 10);
  double delta = num(kv, "delta", -10), kappa = num(kv, "kappa", 0);
  cd eta(num(kv, "eta", 0), num(kv, "etaI", 0));
  Picture p = picture(kv);
  f.dim = cutoff;
  cd zSch(kappa, -delta);
  if (p == Picture::Sch) {
    modeHamiltonian(f, false, zSch, 0., eta, cutoff);
    // frequencies (Mode.tcc:90-93, 127, 160, 193): PumpedLossyModeSch {kappadelta,eta},
    // LossyModeSch {kappa,delta}, PumpedModeSch {delta,eta}, ModeSch {delta}
    if (kappa && isNonZero(eta)) f.complexFreqs.push_back({std::conj(zSch), (double)cutoff});
    else if (kappa) { f.realFreqs.push_back({kappa, (double)cutoff}); f.realFreqs.push_back({delta, (double)cutoff}); }
    else f.realFreqs.push_back({delta, (double)cutoff});
    if (isNonZero(eta)) f.complexFreqs.push_back({eta, std::sqrt((double)cutoff)});
  } else {
    if (!kappa || !isNonZero(eta)) throw std::invalid_argument("oracle: IP/UIP mode needs kappa and eta");
    if (p == Picture::IP) {
      modeHamiltonian(f, true, 0., zSch, eta, cutoff);  // PumpedLossyMode
      f.complexFreqs.push_back({std::conj(zSch), 1.});
    } else {
      modeHamiltonian(f, true, cd(kappa, 0), cd(0, -delta), eta, cutoff);  // PumpedLossyModeUIP
      f.realFreqs.push_back({kappa, (double)cutoff}); f.realFreqs.push_back({delta, 1.});
    }
    f.complexFreqs.push_back({eta, std::sqrt((double)cutoff)});
  }
  if (kappa) f.jumps.push_back(aJumpOf(kappa, cutoff));
  return f;
}

// qbit::make(ParsPumpedLossy, qmp), Qbit.cc:217-246; qubit = two-level mode with -eta (Qbit_.h:99-109)
Free makeQbit(const std::map<std::string, std::string>& kv) {
  Free f; f.kind = "qbit"; f.avKind = "qbit"; f.dim = 2;
  double delta = num(kv, "deltaA", -10), gamma = num(kv, "gamma", 0);
  cd eta(num(kv, "eta", 0), num(kv, "etaI", 0));
  Picture p = picture(kv);
  cd zSch(gamma, -delta);
  if (p == Picture::Sch) {
    modeHamiltonian(f, false, zSch, 0., -eta, 2);
    if (gamma && isNonZero(eta)) { f.complexFreqs.push_back({std::conj(zSch), 1.}); f.complexFreqs.push_back({eta, 1.}); }
    else if (gamma) { f.realFreqs.push_back({gamma, 1.}); f.realFreqs.push_back({delta, 1.}); }
    else { f.realFreqs.push_back({delta, 1.}); if (isNonZero(eta)) f.complexFreqs.push_back({eta, 1.}); }
  } else {
    if (!gamma || !isNonZero(eta)) throw std::invalid_argument("oracle: IP/UIP qbit needs gamma and eta");
    if (p == Picture::IP) { modeHamiltonian(f, true, 0., zSch, -eta, 2); f.complexFreqs.push_back({std::conj(zSch), 1.}); }
    else { modeHamiltonian(f, true, cd(gamma, 0), cd(0, -delta), -eta, 2); f.realFreqs.push_back({gamma, 1.}); f.realFreqs.push_back({delta, 1.}); }
    f.complexFreqs.push_back({eta, 1.});
  }
  if (gamma) f.jumps.push_back(aJumpOf(gamma, 2));
  return f;
}

// PumpedTwoLevelAtomSch, PumpedTwoLevelAtom.cc:13-49
Free makePTLA(const std::map<std::string, std::string>& kv) {
  Free f; f.kind = "ptla"; f.avKind = "qbit"; f.dim = 2;
  double gamma = num(kv, "gamma", 10), delta = num(kv, "deltaA", -10);
  cd eta(num(kv, "eta", 0), num(kv, "etaI", 0));
  cd za(-gamma, delta);
  f.cmatrix = {0., std::conj(eta), -eta, za};
  f.hasHa = true;
  f.complexFreqs.push_back({cd(gamma, delta), 1.}); f.complexFreqs.push_back({eta, 1.});
  Jump j;
  j.act = [za](cd* psi, long s) { psi[0] = std::sqrt(-2. * za.real()) * psi[s]; psi[s] = 0.; };
  j.rateWeights = {0., -2. * za.real()};
  f.jumps.push_back(j);  // no superoperator: Master falls back to J (J rho)^dagger
  return f;
}

// particle::Spatial with deltaK = 1 (Particle.cc:276-290): k_n = n - N/2, x_j = j dx - pi, dx = 2 pi / N
double particleK(int n, int N) { return (double)n - N / 2; }
double particleX(int j, int N) { const double pi = 3.14159265358979323846; return j * (2. * pi / N) - pi; }

// expINKX(dim, nK), Particle.cc:60-72: e^{i nK x} as a Tridiagonal with offset |nK|
Tridiag expINKX(int dim, int nK) {
  int K = std::abs(nK), D = dim - K;
  if (D < 0) return unary(CV(), 0, CV(), CV());
  CV ones(D, 1.);
  if (nK > 0) return unary(CV(), K, ones, CV());
  if (nK == 0) return unary(ones, 0, CV(), CV());
  return unary(CV(), K, CV(), ones);
}
// cosNKX / sinNKX, Particle.cc:76-80, Particle_.h:64-65
Tridiag cosNKXof(int dim, int nK) { return scaled(sum(expINKX(dim, nK), expINKX(dim, -nK)), 0.5); }
Tridiag sinNKXof(int dim, int nK) { return scaled(sum(expINKX(dim, nK), negated(expINKX(dim, -nK))), 1. / cd(0., 2.)); }

// (Pumped)Particle(Sch), Particle.cc:79-110, 210-252: Sch: Tridiagonal(mainDiagonal(space, -omrec))
// (+ hOverI of the pump), mainDiagonal = i omrec k^2 (Particle.cc:92-95); IP: Exact with
// factorExponents -i omrec k^2 and hOverI furnished with the main diagonal.  hOverI =
// vClass (sin: -1) cosNKX(dim, 2 kPart) / (2i) for a real mode function (Particle.cc:84-87).
// Frequencies: RF{omrec, omrec, (1<<fin)^2 (Sch) or 1<<fin (IP)}, RF{vClass, vClass, 1}.
Free makeParticle(const std::map<std::string, std::string>& kv) {
  Free f; f.kind = "particle"; f.avKind = "particle";
  int fin = (int)num(kv, "fin", 6);
  double omrec = num(kv, "omrec", 1.);
  f.vClass = num(kv, "vclass", 0.);
  f.kPart = (int)num(kv, "kpart", 1);
  auto it = kv.find("mf"); f.mf = it == kv.end() ? "sin" : it->second;
  Picture p = picture(kv);
  if (p == Picture::UIP) throw std::invalid_argument("oracle: particle has no UIP");
  f.dim = 1 << fin;
  const bool pumped = f.vClass != 0.;
  CV md(f.dim);  // i omrec k^2
  for (int n = 0; n < f.dim; ++n) md[n] = cd(0, 1) * omrec * (particleK(n, f.dim) * particleK(n, f.dim));
  bool hasPump = pumped && (f.mf == "sin" || f.mf == "cos");
  Tridiag pump = hasPump ? scaled(cosNKXof(f.dim, 2 * f.kPart), cd(f.vClass * (f.mf == "sin" ? -1. : 1.), 0.) / cd(0., 2.))
                         : unary(CV(), 0, CV(), CV());
  if (p == Picture::Sch) {
    CV mdm(f.dim); for (int n = 0; n < f.dim; ++n) mdm[n] = -md[n];
    Tridiag h = unary(mdm, 0, CV(), CV());
    if (hasPump) h = sum(h, pump);
    f.hOverIs.push_back(h);
    f.hasHa = true;
    f.realFreqs.push_back({omrec, (double)f.dim * f.dim});
  } else {
    f.hasEx = true; f.exd = md;
    if (hasPump) { furnishWithFreqs(pump, md); f.hOverIs.push_back(pump); f.hasHa = true; f.timeDep = true; }
    f.realFreqs.push_back({omrec, (double)f.dim});
  }
  if (pumped) f.realFreqs.push_back({f.vClass, 1.});
  return f;
}

// mfNKX(particle, (mf, nK)), Particle.cc:274-289, furnished with the particle's main diagonal
// when it is Exact (expINKX(Ptr, nK), Particle.cc:263-269)
Tridiag mfNKXof(const Free& fp) {
  auto fur = [&](Tridiag t) { if (fp.hasEx) furnishWithFreqs(t, fp.exd); return t; };
  int nK = fp.kPart;
  if (fp.mf == "sin") return scaled(sum(fur(expINKX(fp.dim, nK)), negated(fur(expINKX(fp.dim, -nK)))), 1. / cd(0., 2.));
  if (fp.mf == "cos") return scaled(sum(fur(expINKX(fp.dim, nK)), fur(expINKX(fp.dim, -nK))), 0.5);
  if (fp.mf == "plus") return fur(expINKX(fp.dim, nK));
  return fur(expINKX(fp.dim, -nK));
}

// ffTransformCV with DIR_KX (LazyDensityOperatorFFT.cc:16-40): GSL backward transform
// y_m = sum_n psi_n exp(+2 pi i m n / N), halves swapped with factor N^-1/2, odd entries negated
CV ffTransformKX(const CV& psi) {
  const int N = (int)psi.size();
  const double pi = 3.14159265358979323846;
  CV y(N);
  for (int m = 0; m < N; ++m) {
    cd acc = 0.;
    for (int n = 0; n < N; ++n) acc += psi[n] * std::polar(1.0, 2. * pi * (double)(((long)m * n) % N) / N);
    y[m] = acc;
  }
  CV out(N);
  const int h = N / 2;
  const double nrm = std::pow((double)N, -.5);
  for (int j = 0; j < h; ++j) { out[j] = nrm * y[j + h]; out[j + h] = nrm * y[j]; }
  for (int j = 1; j < N; j += 2) out[j] *= -1.;
  return out;
}

// ---------------------------------------------------------------------------
// System (BinarySystem / Composite) built from a spec string
// ---------------------------------------------------------------------------

struct System {
  std::vector<Free> frees;
  std::vector<Interaction> ias;
  Layout L;
  bool isBinary = false;
  int nJumps() const { int n = 0; for (auto& f : frees) n += (int)f.jumps.size(); return n; }
  int nAvr() const { int n = 0; for (auto& f : frees) n += f.nAvr(); return n; }
  bool hasHa() const { for (auto& f : frees) if (f.hasHa) return true; for (auto& i : ias) if (i.hasHa) return true; return false; }
  bool hasEx() const { for (auto& f : frees) if (f.hasEx) return true; return false; }
  bool hasLi() const { return nJumps() > 0; }
  double highestFrequency() const {
    double m = 0;
    for (auto& f : frees) m = std::max(m, f.highestFrequency());
    if (isBinary) for (auto& i : ias) m = std::max(m, i.highestFrequency());  // Q11: composite ignores acts
    return m;
  }
};

// JaynesCummings Base<true>, JaynesCummings_.h:54-66: tridiagMinusHC(conj(g) sigma * a^dagger)
Interaction makeJC(const System& s, int q, int m, const std::map<std::string, std::string>& kv) {
  Interaction ia; ia.kind = "jc"; ia.axes = {q, m};
  cd g(num(kv, "g", 0), num(kv, "gI", 0));
  const Free& fq = s.frees[q]; const Free& fm = s.frees[m];
  Tridiag sig = aop(fq.dim), a = aop(fm.dim);
  // aop(Ptr) furnishes with freqs when the free is mode::Exact (Mode.cc:288-294)
  if (fq.hasEx) furnishWithFreqs(sig, fq.exd);
  if (fm.hasEx) furnishWithFreqs(a, fm.exd);
  if (isNonZero(g)) {
    ia.hOverIs.push_back(minusHC(product(scaled(sig, std::conj(g)), hc(a))));
    ia.hasHa = true; ia.timeDep = true;
  }
  ia.complexFreqs.push_back({g, std::sqrt((double)fq.dim * fm.dim)});
  return ia;
}

// CoupledModes<CM_XX_RWA> (CoupledModes.cc:36-41): tridiagMinusHC(conj(u) aop(m1) aop(m2)^dagger),
// frequency CF{"u", u, dim1 sqrt(dim2)} (CoupledModes.cc:13-17)
Interaction makeBS(const System& s, int m1, int m2, const std::map<std::string, std::string>& kv) {
  Interaction ia; ia.kind = "bs"; ia.axes = {m1, m2};
  cd u(num(kv, "g", 0), num(kv, "gI", 0));
  const Free& f1 = s.frees[m1]; const Free& f2 = s.frees[m2];
  Tridiag a1 = aop(f1.dim), a2 = aop(f2.dim);
  if (f1.hasEx) furnishWithFreqs(a1, f1.exd);
  if (f2.hasEx) furnishWithFreqs(a2, f2.exd);
  ia.hOverIs.push_back(minusHC(product(scaled(a1, std::conj(u)), hc(a2))));
  ia.hasHa = true; ia.timeDep = true;
  ia.complexFreqs.push_back({u, f1.dim * std::sqrt((double)f2.dim)});
  return ia;
}

// sigma_z (x) sigma_z with sigma_z = diag(-1, +1): H/i = -i J sz sz.  Not a reference element
// (qbit::sigmazop is declared but never defined, Qbit_.h:68-70); frequency J.
Interaction makeZZ(const System& s, int a, int b, const std::map<std::string, std::string>& kv) {
  Interaction ia; ia.kind = "zz"; ia.axes = {a, b};
  if (s.frees[a].dim != 2 || s.frees[b].dim != 2) throw std::invalid_argument("oracle: zz couples qubits");
  double J = num(kv, "J", 0);
  Tridiag sz = unary({-1., 1.}, 0, CV(), CV());
  ia.hOverIs.push_back(scaled(product(sz, sz), cd(0., -J)));
  ia.hasHa = true;
  ia.realFreqs.push_back({J, 1.});
  return ia;
}

// expINKX / cosNKX (Particle.cc:60-82): e^{iKx} = Tridiagonal(Diag(), K, ones) (minus diagonal),
// e^{-iKx} = Tridiagonal(Diag(), K, Diag(), ones) (plus diagonal)
Tridiag cosNKX(int dim, int K) {
  int D = dim - K;
  CV ones(std::max(D, 0), 1.);
  Tridiag r = sum(unary(CV(), K, ones, CV()), unary(CV(), K, CV(), ones));
  return scaled(r, 0.5);
}

Tridiag mfComposition(const Free& fp, const std::string& a, const std::string& b, int nK);

// mfNKX(particle, (mf, nK)) for a given mode function (Particle.cc:278-289), furnished when the
// particle is Exact (expINKX(Ptr, nK), Particle.cc:263-269)
Tridiag mfNKXgiven(const Free& fp, const std::string& mf, int nK) {
  Free g = fp; g.mf = mf; g.kPart = nK;
  return mfNKXof(g);
}

// ParticleAlongCavity (ParticleCavity.cc:26-53, 87-110):
//   dispersive: u0 nop(mode) (+-) cosNKX(particle, 2 kcav) / (2i) for a real cavity mode function
//     (nop = Tridiagonal(mainDiagonal(1,0,dim)) carries no frequencies, so neither does the term);
//   particle without its own pump, or given vclass (1p1mconf 2, 4): etaeff = etaEff(u0, vclass),
//     interfericOneModeFN = etaeff tridiagPlusHC_overI(aop(mode) mfNKX(particle, (mf, kcav)));
//   pumped particle (1p1mconf 3): etaeff = etaEff(u0, particle's vClass); with kcav = |kPart|
//     (isSpecialH_) interfericTwoModeFN = etaeff tridiagPlusHC_overI(aop(mode)^dagger
//     mfComposition(particle, (mf, kcav), (mfPart, kPart)));
//   pumped particle of another wavenumber (ParticleCavity.cc:102-140): firstH = etaeff aop(mode)
//     mfNKX(particle, (mf, kcav)) / i, firstHT = -firstH^dagger, secondH = mfNKX(particle,
//     particle's mf)^dagger, secondHT = secondH^dagger, applied as secondH (firstH psi) +
//     secondHT (firstHT psi), the second factor on the particle axis.
// Frequencies RF{"Unot", u0, mode dim}, RF{"etaeff", etaeff, sqrt(mode dim)} (ParticleCavity.cc:62-66)
Interaction makePcav(const System& s, int m, int p, const std::map<std::string, std::string>& kv) {
  Interaction ia; ia.kind = "pcav"; ia.axes = {m, p};
  const Free& fm = s.frees[m]; const Free& fp = s.frees[p];
  if (fm.kind != "mode" || fp.kind != "particle") throw std::invalid_argument("oracle: pcav couples a mode and a particle");
  double u0 = num(kv, "u0", 0);
  int kcav = (int)num(kv, "kcav", 1);
  auto it = kv.find("mf");
  const std::string mfs = it == kv.end() ? "cos" : it->second;
  const bool givenV = kv.count("vclass") > 0;
  const bool pumpedCtor = fp.vClass != 0. && !givenV;
  const double v = pumpedCtor ? fp.vClass : num(kv, "vclass", 0.);
  if (u0 * v < 0.) throw std::domain_error("Unot VClass sign discrepancy");
  const double etaeff = (u0 > 0 ? 1. : u0 < 0 ? -1. : 0.) * std::sqrt(u0 * v);
  const bool special = !pumpedCtor || kcav == std::abs(fp.kPart);
  double sgn = mfs == "sin" ? -1. : 1.;
  // dispersive(): only for a real mode function (ParticleCavity.cc:26-32)
  if (u0 != 0. && 2 * kcav < fp.dim && (mfs == "cos" || mfs == "sin")) {
    CV nd(fm.dim); for (int n = 0; n < fm.dim; ++n) nd[n] = (double)n;
    Tridiag t = product(scaled(unary(nd, 0, CV(), CV()), u0 * sgn), cosNKX(fp.dim, 2 * kcav));
    ia.hOverIs.push_back(scaled(t, 1. / cd(0., 2.)));
    ia.hasHa = true; ia.timeDep = true;
  }
  if (etaeff != 0. && special) {
    Tridiag a = aop(fm.dim);
    if (fm.hasEx) furnishWithFreqs(a, fm.exd);
    Tridiag t = pumpedCtor ? product(hc(a), mfComposition(fp, mfs, fp.mf, kcav)) : product(a, mfNKXgiven(fp, mfs, kcav));
    ia.hOverIs.push_back(scaled(sum(t, hc(t)), cd(etaeff, 0.) / cd(0., 1.)));
    ia.hasHa = true; ia.timeDep = true;
  }
  if (etaeff != 0. && !special) {
    Tridiag a = aop(fm.dim);
    if (fm.hasEx) furnishWithFreqs(a, fm.exd);
    Tridiag first = scaled(product(a, mfNKXgiven(fp, mfs, kcav)), cd(etaeff, 0.) / cd(0., 1.));
    Tridiag second = hc(mfNKXof(fp));
    ia.chained.push_back({first, second});
    ia.chained.push_back({negated(hc(first)), hc(second)});
    ia.hasHa = true; ia.timeDep = true;
  }
  ia.realFreqs.push_back({u0, (double)fm.dim});
  ia.realFreqs.push_back({etaeff, std::sqrt((double)fm.dim)});
  return ia;
}

// ParticleOrthogonalToCavity (ParticleCavity.cc:47-57, 79-85):
// etaeff tridiagPlusHC_overI(aop(mode) mfNKX(particle, mf)), etaeff = sign(u0) sqrt(u0 vClass)
// (ParticleCavity.cc:20-24; Tridiagonal.h: tridiagPlusHC_overI(t) = (t + t^dagger)/i)
Interaction makePcavOrth(const System& s, int m, int p, const std::map<std::string, std::string>& kv) {
  Interaction ia; ia.kind = "pcav_orth"; ia.axes = {m, p};
  const Free& fm = s.frees[m]; const Free& fp = s.frees[p];
  if (fm.kind != "mode" || fp.kind != "particle" || fp.vClass == 0.) throw std::invalid_argument("oracle: pcav_orth");
  double u0 = num(kv, "u0", -1.);
  if (u0 * fp.vClass < 0.) throw std::domain_error("Unot VClass sign discrepancy");
  double etaeff = (u0 > 0 ? 1. : u0 < 0 ? -1. : 0.) * std::sqrt(u0 * fp.vClass);
  Tridiag a = aop(fm.dim);
  if (fm.hasEx) furnishWithFreqs(a, fm.exd);
  Tridiag t = product(a, mfNKXof(fp));
  ia.hOverIs.push_back(scaled(sum(t, hc(t)), cd(etaeff, 0.) / cd(0., 1.)));
  ia.hasHa = true; ia.timeDep = true;
  ia.realFreqs.push_back({u0, (double)fm.dim});
  ia.realFreqs.push_back({etaeff, std::sqrt((double)fm.dim)});
  return ia;
}

// mfComposition, Particle.cc:290-345 (positive wavenumbers), with the particle's factors
// furnished when it is Exact; identity(dim) is a bare main diagonal (Tridiagonal.cc:64-69)
Tridiag mfComposition(const Free& fp, const std::string& a, const std::string& b, int nK) {
  auto fur = [&](Tridiag t) { if (fp.hasEx) furnishWithFreqs(t, fp.exd); return t; };
  CV ones(fp.dim, 1.);
  Tridiag id = unary(ones, 0, CV(), CV());
  auto e = [&](int k) { return fur(expINKX(fp.dim, k)); };
  auto cos2 = [&]() { return scaled(sum(e(2 * nK), e(-2 * nK)), 0.5); };
  auto sin2 = [&]() { return scaled(sum(e(2 * nK), negated(e(-2 * nK))), 1. / cd(0., 2.)); };
  if (a == "sin" && b == "sin") return scaled(sum(id, negated(cos2())), 0.5);
  if (a == "cos" && b == "cos") return scaled(sum(id, cos2()), 0.5);
  if ((a == "sin" && b == "cos") || (a == "cos" && b == "sin")) return scaled(sin2(), 0.5);
  if (a == b) return id;
  if (a == "plus" && b == "minus") return e(-2 * nK);
  if (a == "minus" && b == "plus") return e(2 * nK);
  if ((a == "plus" && b == "sin") || (a == "sin" && b == "minus")) return scaled(sum(e(-2 * nK), negated(id)), cd(0., 0.5));
  if ((a == "plus" && b == "cos") || (a == "cos" && b == "minus")) return scaled(sum(id, e(-2 * nK)), 0.5);
  if ((a == "minus" && b == "cos") || (a == "cos" && b == "plus")) return scaled(sum(id, e(2 * nK)), 0.5);
  return scaled(sum(e(2 * nK), negated(id)), cd(0., -0.5));
}

// ParticleTwoModes, special Hamiltonian for equal wavenumbers (ParticleTwoModes.cc:19-46):
// tridiagPlusHC_overI(factor aop(mode0)^dagger aop(mode1) mfComposition(part, mf0, mf1)),
// factor = sign(u0) sqrt(u0 u1) e^{i phi}; frequencies RF{Unot0, u0, sqrt(dim0)}, RF{Unot1, u1, sqrt(dim1)}
Interaction makePTM(const System& s, int m0, int m1, int p, const std::map<std::string, std::string>& kv) {
  Interaction ia; ia.kind = "ptm"; ia.axes = {m0, m1, p};
  const Free& f0 = s.frees[m0]; const Free& f1 = s.frees[m1]; const Free& fp = s.frees[p];
  double u0 = num(kv, "u0", -1.), u1 = num(kv, "u1", -1.), phi = num(kv, "phi", 0.);
  int k0 = (int)num(kv, "k0", 1), k1 = (int)num(kv, "k1", 1);
  if (k0 != k1 || u0 * u1 < 0) throw std::invalid_argument("oracle: ptm");
  auto g = [&](const char* key) { auto it = kv.find(key); return it == kv.end() ? std::string("plus") : it->second; };
  cd fac = (u0 > 0 ? 1. : u0 < 0 ? -1. : 0.) * std::sqrt(u0 * u1) * std::exp(cd(0., phi));
  Tridiag a0 = aop(f0.dim), a1 = aop(f1.dim);
  if (f0.hasEx) furnishWithFreqs(a0, f0.exd);
  if (f1.hasEx) furnishWithFreqs(a1, f1.exd);
  Tridiag t = product(product(scaled(hc(a0), fac), a1), mfComposition(fp, g("mf0"), g("mf1"), k0));
  ia.hOverIs.push_back(scaled(sum(t, hc(t)), 1. / cd(0., 1.)));
  ia.hasHa = true; ia.timeDep = true;
  ia.realFreqs.push_back({u0, std::sqrt((double)f0.dim)});
  ia.realFreqs.push_back({u1, std::sqrt((double)f1.dim)});
  return ia;
}

System* buildSystem(const std::string& spec) {
  std::unique_ptr<System> s(new System);
  std::string line;
  std::vector<std::string> lines;
  int forced = -1;
  { std::string cur; for (char c : spec) { if (c == ';' || c == '\n') { lines.push_back(cur); cur.clear(); } else cur += c; } lines.push_back(cur); }
  for (auto& ln : lines) {
    std::istringstream is(ln); std::string what, kind;
    if (!(is >> what)) continue;
    is >> kind;
    if (what == "free") {
      auto kv = parseKV(is);
      if (kind == "mode") s->frees.push_back(makeMode(kv));
      else if (kind == "qbit") s->frees.push_back(makeQbit(kv));
      else if (kind == "ptla") s->frees.push_back(makePTLA(kv));
      else if (kind == "particle") s->frees.push_back(makeParticle(kv));
      else throw std::invalid_argument("oracle: unknown free " + kind);
    } else if (what == "ia") {
      std::vector<int> axes; std::string tok; std::map<std::string, std::string> kv;
      while (is >> tok) { auto p = tok.find('='); if (p == std::string::npos) axes.push_back(std::stoi(tok)); else kv[tok.substr(0, p)] = tok.substr(p + 1); }
      if (kind == "jc") s->ias.push_back(makeJC(*s, axes.at(0), axes.at(1), kv));
      else if (kind == "bs") s->ias.push_back(makeBS(*s, axes.at(0), axes.at(1), kv));
      else if (kind == "zz") s->ias.push_back(makeZZ(*s, axes.at(0), axes.at(1), kv));
      else if (kind == "pcav") s->ias.push_back(makePcav(*s, axes.at(0), axes.at(1), kv));
      else if (kind == "pcav_orth") s->ias.push_back(makePcavOrth(*s, axes.at(0), axes.at(1), kv));
      else if (kind == "ptm") s->ias.push_back(makePTM(*s, axes.at(0), axes.at(1), axes.at(2), kv));
      else throw std::invalid_argument("oracle: unknown interaction " + kind);
    } else if (what == "system") {
      forced = kind == "binary" ? 1 : 0;  // binary::make vs composite::make chosen explicitly
    } else throw std::invalid_argument("oracle: bad statement " + what);
  }
  std::vector<int> dims; for (auto& f : s->frees) dims.push_back(f.dim);
  s->L = Layout(dims);
  s->isBinary = forced >= 0 ? forced == 1 : (s->frees.size() == 2 && s->ias.size() <= 1);
  return s.release();
}

// Hamiltonian::addContribution of the composite: frees in index order over
// fullRange<V_i>, then interactions over fullRange<V_act>  (BinarySystem.cc:194-207,
// Composite.tcc:178-197).  Time-dependent tridiagonals are propagated to t-t0 first
// (TridiagonalHamiltonian.h:42-47).
void addContribution(System& s, double t, const cd* psi, cd* dpsidt, double t0) {
  for (size_t i = 0; i < s.frees.size(); ++i) {
    Free& f = s.frees[i];
    if (!f.hasHa) continue;
    forSlices(s.L, {(int)i}, [&](long off, const std::vector<long>& vs) {
      if (!f.cmatrix.empty()) {  // linalg::apply, CMatrix.cc:7-11
        int d = f.dim;
        for (int r = 0; r < d; ++r) { cd acc = 0.; for (int c = 0; c < d; ++c) acc += f.cmatrix[r * d + c] * psi[off + c * vs[0]]; dpsidt[off + r * vs[0]] += acc; }
      } else {
        for (auto& h : f.hOverIs) { if (f.timeDep) propagate(h, t - t0); apply(h, psi + off, dpsidt + off, vs); }
      }
    });
  }
  for (auto& ia : s.ias) {
    if (!ia.hasHa) continue;
    forSlices(s.L, ia.axes, [&](long off, const std::vector<long>& vs) {
      for (auto& h : ia.hOverIs) { if (ia.timeDep) propagate(h, t - t0); apply(h, psi + off, dpsidt + off, vs); }
      for (auto& c : ia.chained) {  // rank-2 interactions: the second factor on axis 1
        if (ia.timeDep) { propagate(c.first, t - t0); propagate(c.second, t - t0); }
        const int d0 = s.frees[ia.axes[0]].dim, d1 = s.frees[ia.axes[1]].dim;
        std::vector<cd> tmp((size_t)d0 * d1, cd(0.));
        apply(c.first, psi + off, tmp.data(), {(long)d1, 1L});
        for (int r = 0; r < d0; ++r) apply(c.second, tmp.data() + (size_t)r * d1, dpsidt + off + r * vs[0], {vs[1]});
      }
    });
  }
}

// Exact::actWithU (BinarySystem.cc:177-184, FreeExact.h:31, Mode.cc:42-45)
void actWithU(System& s, double t, cd* psi, double t0) {
  for (size_t i = 0; i < s.frees.size(); ++i) {
    Free& f = s.frees[i];
    if (!f.hasEx) continue;
    CV u(f.dim);
    for (int n = 0; n < f.dim; ++n) u[n] = std::exp(-f.exd[n] * (t - t0));
    forSlices(s.L, {(int)i}, [&](long off, const std::vector<long>& vs) { for (int n = 0; n < f.dim; ++n) psi[off + n * vs[0]] *= u[n]; });
  }
}

// rates (binary::average<LA_Li>, BinarySystem.cc:141-156): frees first, summed over slices (partialTrace)
std::vector<double> rates(System& s, const cd* psi) {
  std::vector<double> r;
  for (size_t i = 0; i < s.frees.size(); ++i) {
    for (auto& j : s.frees[i].jumps) {
      double acc = 0;
      bool first = true;
      forSlices(s.L, {(int)i}, [&](long off, const std::vector<long>& vs) {
        double part = 0;
        for (int n = 0; n < s.frees[i].dim; ++n) part += j.rateWeights[n].real() * std::norm(psi[off + n * vs[0]]);
        acc = first ? part : acc + part; first = false;
      });
      r.push_back(acc);
    }
  }
  // Liouvillean::rates passes LiouvilleanAveragedCommon::average, which refuses non-finite values
  // (LiouvilleanAveragedCommon.h:82-93)
  for (double x : r) if (!std::isfinite(x)) throw std::runtime_error("Infinite detected in averages");
  return r;
}

void actWithJ(System& s, cd* psi, int m) {
  for (size_t i = 0; i < s.frees.size(); ++i) {
    int n = (int)s.frees[i].jumps.size();
    if (m < n) { auto& j = s.frees[i].jumps[m]; forSlices(s.L, {(int)i}, [&](long off, const std::vector<long>& vs) { j.act(psi + off, vs[0]); }); return; }
    m -= n;
  }
}

// Averaged::average of the frees (Mode.cc:177-195, Qbit.cc:27-32) summed over slices,
// on the (unnormalised) pure state; process_v applied separately.
void averagesRaw(System& s, const cd* psi, double* out) {
  int k = 0;
  for (size_t i = 0; i < s.frees.size(); ++i) {
    Free& f = s.frees[i];
    if (f.avKind.empty()) continue;
    double a[4] = {0, 0, 0, 0};
    forSlices(s.L, {(int)i}, [&](long off, const std::vector<long>& vs) {
      auto el = [&](int n) { return psi[off + n * vs[0]]; };
      if (f.avKind == "mode") {
        for (int n = 1; n < f.dim; ++n) {
          double dg = std::norm(el(n));
          a[0] += n * dg; a[1] += n * n * dg;
          cd od = std::sqrt((double)n) * (el(n) * std::conj(el(n - 1)));
          a[2] += od.real(); a[3] += od.imag();
        }
      } else if (f.avKind == "particle") {  // particle::Averaged::average_v, Particle.cc:138-160
        CV sl(f.dim);
        for (int n = 0; n < f.dim; ++n) {
          sl[n] = el(n);
          double dg = std::norm(sl[n]), k = particleK(n, f.dim);
          a[0] += k * dg; a[1] += k * k * dg;
        }
        CV px = ffTransformKX(sl);
        for (int j = 0; j < f.dim; ++j) { double dg = std::norm(px[j]), x = particleX(j, f.dim); a[2] += x * dg; a[3] += x * x * dg; }
      } else {
        cd r10 = el(1) * std::conj(el(0));
        a[0] += std::norm(el(0)); a[1] += std::norm(el(1)); a[2] += r10.real(); a[3] += r10.imag();
      }
    });
    for (int q = 0; q < 4; ++q) out[k + q] = a[q];
    k += 4;
  }
}

// density-operator version (Master): ldo(i,j) row-major d x d
void averagesRawRho(System& s, const cd* rho, double* out) {
  int k = 0; long d = s.L.total;
  for (size_t i = 0; i < s.frees.size(); ++i) {
    Free& f = s.frees[i];
    if (f.avKind.empty()) continue;
    double a[4] = {0, 0, 0, 0};
    forSlices(s.L, {(int)i}, [&](long off, const std::vector<long>& vs) {
      auto el = [&](int m, int n) { return rho[(off + m * vs[0]) * d + off + n * vs[0]]; };
      if (f.avKind == "mode") {
        for (int n = 1; n < f.dim; ++n) {
          double dg = el(n, n).real();
          a[0] += n * dg; a[1] += n * n * dg;
          cd od = std::sqrt((double)n) * el(n, n - 1);
          a[2] += od.real(); a[3] += od.imag();
        }
      } else if (f.avKind == "particle") {
        // x populations of F rho F^dagger, F the transform of ffTransformKX (the density-operator
        // ffTransform of LazyDensityOperatorFFT.cc:53-58 transforms rows and columns)
        const int N = f.dim;
        for (int n = 0; n < N; ++n) { double dg = el(n, n).real(), k = particleK(n, N); a[0] += k * dg; a[1] += k * k * dg; }
        std::vector<CV> F(N);  // F[n] = transform of the unit vector |n>: column n of F
        for (int n = 0; n < N; ++n) { CV u(N, 0.); u[n] = 1.; F[n] = ffTransformKX(u); }
        for (int j = 0; j < N; ++j) {
          cd acc = 0.;
          for (int n = 0; n < N; ++n) for (int m = 0; m < N; ++m) acc += F[n][j] * el(n, m) * std::conj(F[m][j]);
          const double x = particleX(j, N);
          a[2] += x * acc.real(); a[3] += x * x * acc.real();
        }
      } else {
        a[0] += el(0, 0).real(); a[1] += el(1, 1).real(); a[2] += el(1, 0).real(); a[3] += el(1, 0).imag();
      }
    });
    for (int q = 0; q < 4; ++q) out[k + q] = a[q];
    k += 4;
  }
}

void processAverages(System& s, double* av) {  // Mode.cc:198-201
  int k = 0;
  for (auto& f : s.frees) {
    if (f.avKind.empty()) continue;
    if (f.avKind == "mode") av[k + 1] -= av[k] * av[k];
    if (f.avKind == "particle") { av[k + 1] -= av[k] * av[k]; av[k + 3] = std::sqrt(av[k + 3] - av[k + 2] * av[k + 2]); }  // Particle.cc:163-167
    k += 4;
  }
}

// ---------------------------------------------------------------------------
// ODE engine: ode_engine::Base<controlled_runge_kutta<runge_kutta_cash_karp54>>
// (ODE.h:125-262) with Boost.Odeint's default_error_checker / default_step_adjuster.
// ---------------------------------------------------------------------------

struct RKCK {
  // Cash & Karp (1990) tableau, as in odeint's runge_kutta_cash_karp54
  static constexpr double c2 = 1. / 5, c3 = 3. / 10, c4 = 3. / 5, c5 = 1., c6 = 7. / 8;
  static constexpr double a21 = 1. / 5;
  static constexpr double a31 = 3. / 40, a32 = 9. / 40;
  static constexpr double a41 = 3. / 10, a42 = -9. / 10, a43 = 6. / 5;
  static constexpr double a51 = -11. / 54, a52 = 5. / 2, a53 = -70. / 27, a54 = 35. / 27;
  static constexpr double a61 = 1631. / 55296, a62 = 175. / 512, a63 = 575. / 13824, a64 = 44275. / 110592, a65 = 253. / 4096;
  static constexpr double b1 = 37. / 378, b2 = 0., b3 = 250. / 621, b4 = 125. / 594, b5 = 0., b6 = 512. / 1771;
  static constexpr double e1 = 37. / 378 - 2825. / 27648, e2 = 0., e3 = 250. / 621 - 18575. / 48384,
                          e4 = 125. / 594 - 13525. / 55296, e5 = -277. / 14336, e6 = 512. / 1771 - 1. / 4;
};

enum class Controller { Odeint = 0, Gsl = 1 };

struct OdeStats { long nDerivs = 0, nSteps = 0, nFailed = 0, nTries = 0; };

struct Ode {
  double dtDid = 0., dtTry;
  double epsAbsOdeint, epsRelOdeint;  // after the Q1 swap: abs = user epsRel, rel = user epsAbs
  Controller controller = Controller::Odeint;
  OdeStats st;
  // scratch
  CV k1, k2, k3, k4, k5, k6, xt, xn, xe;
  Ode(double dtInit, double epsRel, double epsAbs) : dtTry(dtInit), epsAbsOdeint(epsRel), epsRelOdeint(epsAbs) {}

  // controlled_runge_kutta::try_step(sys,x,t,dt) (explicit_error_stepper_tag):
  // sys(x,dxdt,t); do_step with error; error = max_i |xerr_i|/(epsAbs+epsRel*(|x_i|+dt*|dxdt_i|));
  // reject if >1 (dt *= max(0.9*err^(-1/3),0.2)), else accept and grow if err<0.5.
  template <typename F>
  bool tryStep(F&& f, CV& x, double& t, double& dt) {
    size_t n = x.size();
    for (CV* v : {&k1, &k2, &k3, &k4, &k5, &k6, &xt, &xn, &xe}) v->assign(n, 0.);
    st.nTries++;
    auto sys = [&](const CV& y, CV& dy, double tt) { st.nDerivs++; f(y, dy, tt); };
    sys(x, k1, t);
    using R = RKCK;
    double s1 = R::a21 * dt;
    for (size_t i = 0; i < n; ++i) xt[i] = 1.0 * x[i] + s1 * k1[i];
    sys(xt, k2, t + R::c2 * dt);
    { double q1 = R::a31 * dt, q2 = R::a32 * dt;
      for (size_t i = 0; i < n; ++i) xt[i] = 1.0 * x[i] + q1 * k1[i] + q2 * k2[i]; }
    sys(xt, k3, t + R::c3 * dt);
    { double q1 = R::a41 * dt, q2 = R::a42 * dt, q3 = R::a43 * dt;
      for (size_t i = 0; i < n; ++i) xt[i] = 1.0 * x[i] + q1 * k1[i] + q2 * k2[i] + q3 * k3[i]; }
    sys(xt, k4, t + R::c4 * dt);
    { double q1 = R::a51 * dt, q2 = R::a52 * dt, q3 = R::a53 * dt, q4 = R::a54 * dt;
      for (size_t i = 0; i < n; ++i) xt[i] = 1.0 * x[i] + q1 * k1[i] + q2 * k2[i] + q3 * k3[i] + q4 * k4[i]; }
    sys(xt, k5, t + R::c5 * dt);
    { double q1 = R::a61 * dt, q2 = R::a62 * dt, q3 = R::a63 * dt, q4 = R::a64 * dt, q5 = R::a65 * dt;
      for (size_t i = 0; i < n; ++i) xt[i] = 1.0 * x[i] + q1 * k1[i] + q2 * k2[i] + q3 * k3[i] + q4 * k4[i] + q5 * k5[i]; }
    sys(xt, k6, t + R::c6 * dt);
    { double q1 = R::b1 * dt, q2 = R::b2 * dt, q3 = R::b3 * dt, q4 = R::b4 * dt, q5 = R::b5 * dt, q6 = R::b6 * dt;
      for (size_t i = 0; i < n; ++i) xn[i] = 1.0 * x[i] + q1 * k1[i] + q2 * k2[i] + q3 * k3[i] + q4 * k4[i] + q5 * k5[i] + q6 * k6[i]; }
    { double q1 = R::e1 * dt, q2 = R::e2 * dt, q3 = R::e3 * dt, q4 = R::e4 * dt, q5 = R::e5 * dt, q6 = R::e6 * dt;
      for (size_t i = 0; i < n; ++i) xe[i] = q1 * k1[i] + q2 * k2[i] + q3 * k3[i] + q4 * k4[i] + q5 * k5[i] + q6 * k6[i]; }
    if (controller == Controller::Gsl) { sys(xn, k1, t + dt); return gslControl(t, dt, x); }
    double err = 0.;
    for (size_t i = 0; i < n; ++i) {
      double e = std::abs(xe[i]) / (epsAbsOdeint + epsRelOdeint * (1.0 * std::abs(x[i]) + (1.0 * dt) * std::abs(k1[i])));
      err = std::max(err, e);
    }
    if (err > 1.0) {
      dt *= std::max(9. / 10. * std::pow(err, -1. / (4 - 1)), 1. / 5.);
      return false;
    }
    t += dt;
    if (err < 0.5) {
      err = std::max(std::pow(5.0, -5.0), err);
      dt *= 9. / 10. * std::pow(err, -1. / 5);
    }
    x.swap(xn);
    return true;
  }

  // GSL rkck + control_standard(eps_abs=user epsRel, eps_rel=user epsAbs, a_y=1, a_dydt=1)
  // + hadjust (ODE_GSL.h:44-84): the step is always accepted; hadjust then sees the new
  // state and the derivative at the new point (dydt_out, here in k1) as 2d real components
  // (BlitzArray.h:141-163).  hadjust's order is the stepper's reported order, 5 for GSL's rkck:
  // the reference's GSL tolerance tests (EvolvedTestGSL.cc, HarmonicOscillatorComplexGSL.cc) pass
  // with it and fail with 4.
  bool gslControl(double& t, double& dt, CV& x) {
    size_t n = x.size();
    double rmax = 2.2250738585072014e-308;  // DBL_MIN
    for (size_t i = 0; i < n; ++i) {
      double yv[2] = {xn[i].real(), xn[i].imag()}, ye[2] = {xe[i].real(), xe[i].imag()}, dy[2] = {k1[i].real(), k1[i].imag()};
      for (int c = 0; c < 2; ++c) {
        double D0 = epsRelOdeint * (1.0 * std::fabs(yv[c]) + 1.0 * std::fabs(dt * dy[c])) + epsAbsOdeint;
        double r = std::fabs(ye[c]) / std::fabs(D0);
        rmax = std::max(r, rmax);
      }
    }
    const double h_old = dt, S = 0.9;
    constexpr unsigned kOrd = 5;
    t += dt;
    x.swap(xn);
    if (rmax > 1.1) {
      double r = S / std::pow(rmax, 1.0 / kOrd);
      if (r < 0.2) r = 0.2;
      dt = r * h_old;
    } else if (rmax < 0.5) {
      double r = S / std::pow(rmax, 1.0 / (kOrd + 1.0));
      if (r > 5.0) r = 5.0;
      if (r < 1.0) r = 1.0;
      dt = r * h_old;
    }
    return true;
  }

  // ode_engine::Base::step, ODE.h:234-262
  template <typename F>
  void step(double deltaT, F&& f, double& time, CV& state) {
    const double timeBefore = time;
    double nextDtTry = (std::fabs(deltaT) < std::fabs(dtTry / 10.) ? dtTry : 0.);
    if ((deltaT > 0) != (dtTry > 0)) dtTry = -dtDid;
    if (std::fabs(dtTry) > std::fabs(deltaT)) dtTry = deltaT;
    long nf = 0;
    while (!tryStep(f, state, time, dtTry)) ++nf;
    dtDid = time - timeBefore;
    if (nextDtTry) dtTry = nextDtTry;
    st.nSteps++; st.nFailed += nf;
  }
};

// ---------------------------------------------------------------------------
// MCWF trajectory (MCWF_Trajectory.h:86-374)
// ---------------------------------------------------------------------------

enum class RngKind { Pcg64 = 0, Philox = 1 };

struct JumpRec { double t; int ch; };

struct MCWF {
  System* sys;
  CV psi;
  double t = 0., t0 = 0.;
  Ode ode;
  double dpLimit, overshootTol;
  RngKind rk; Pcg64 pcg; uint64_t seed, traj, stepOrdinal = 0;
  long nMCWFSteps = 0, nStepBacks = 0, nOvershot = 0;
  double maxNormDev = 0., dpMaxOvershoot = 0., dpTolMaxOvershoot = 0.;  // MCWF_TrajectoryLogger.cc:45-67
  std::vector<JumpRec> jumps;

  MCWF(System* s, const CV& psi0, double dtInit, double epsRel, double epsAbs, double dpL, double ovT,
       RngKind k, uint64_t sd, uint64_t stream, Controller ctl)
      : sys(s), psi(psi0), ode(dtInit, epsRel, epsAbs), dpLimit(dpL), overshootTol(ovT), rk(k), pcg(sd, stream), seed(sd), traj(stream) {
    ode.controller = ctl;
    if (sys->hasLi() && !t) { auto r = rates(*sys, psi.data()); manageTimeStep(r, 0, 0, false); }  // :114-117
  }

  double sample() { return rk == RngKind::Pcg64 ? pcg.uniform() : counter_uniform(seed, traj, stepOrdinal++); }

  double renorm() {  // StateVector::renorm, StateVector.h:148-153
    double s = 0; for (auto& x : psi) s += std::norm(x);
    double nrm = std::sqrt(s);
    for (auto& x : psi) x /= nrm;
    return nrm;
  }

  void coherentTimeDevelopment(double Dt) {  // :250-276
    if (sys->hasHa()) {
      ode.step(Dt, [this](const CV& y, CV& dy, double tt) {
        std::fill(dy.begin(), dy.end(), cd(0, 0));
        addContribution(*sys, tt, y.data(), dy.data(), t0);
      }, t, psi);
    } else {
      double stepToDo = sys->hasLi() ? std::min(ode.dtTry, Dt) : Dt;
      t += stepToDo; ode.dtDid = stepToDo;
    }
    if (sys->hasEx()) { actWithU(*sys, t, psi.data(), t0); t0 = t; }
    double nrm = renorm();
    maxNormDev = std::max(maxNormDev, std::fabs(1 - nrm));
  }

  bool manageTimeStep(const std::vector<double>& r, double tCache, double dtDidCache, bool logControl = true) {  // :294-317
    double totalRate = 0; for (double x : r) totalRate += x;
    const double dtDid = ode.dtDid, dtTry = ode.dtTry;
    const double sugg = dpLimit / totalRate;
    if (totalRate * dtDid > overshootTol * dpLimit) {
      t = tCache; ode.dtDid = dtDidCache; ode.dtTry = sugg;
      nStepBacks++;
      if (logControl) dpTolMaxOvershoot = std::max(dpTolMaxOvershoot, totalRate * dtDid);
      return true;
    } else if (totalRate * dtTry > dpLimit) {
      nOvershot++;
      if (logControl) dpMaxOvershoot = std::max(dpMaxOvershoot, totalRate * dtTry);
    }
    ode.dtTry = sys->hasHa() ? std::min(ode.dtTry, sugg) : sugg;
    return false;
  }

  void performJump(const std::vector<double>& r) {  // :320-344
    double random = sample() / ode.dtDid;
    int no = 0;
    for (; random > 0 && no != (int)r.size(); random -= r[no++]) ;
    if (random < 0) {
      --no;
      actWithJ(*sys, psi.data(), no);
      double nf = std::sqrt(r[no]);
      if (!std::isfinite(nf)) throw std::runtime_error("Infinite detected in MCWF_Trajectory::performJump");
      for (auto& x : psi) x /= nf;
      jumps.push_back({t, no});
    }
  }

  void step(double Dt) {  // :347-374
    const CV psiCache(psi);
    double tCache = t, dtDidCache = ode.dtDid;
    coherentTimeDevelopment(Dt);
    if (sys->hasLi()) {
      auto r = rates(*sys, psi.data());
      while (manageTimeStep(r, tCache, dtDidCache)) {
        psi = psiCache;
        coherentTimeDevelopment(Dt);
        r = rates(*sys, psi.data());
      }
      performJump(r);
    }
    nMCWFSteps++;
  }

  void advance(double deltaT) {  // Trajectory.h:43-50
    double endTime = t + deltaT;
    while (double dt = endTime - t) step(dt);
  }
};

// Master (Master.h:152-208)
struct Master {
  System* sys;
  CV rho;  // d x d row-major
  double t = 0., t0 = 0.;
  Ode ode;
  long d;
  Master(System* s, const CV& rho0, double dtInit, double epsRel, double epsAbs)
      : sys(s), rho(rho0), ode(dtInit, epsRel, epsAbs), d(s->L.total) {
    for (auto& f : sys->frees) if (f.hasEx) for (auto& z : f.exd) if (z.real() != 0.) throw std::runtime_error("SystemNotApplicable");
  }

  // hermitianConjugateSelf (ComplexArrayExtensions.h:40-46)
  static void hcSelf(CV& m, long d) {
    for (long i = 0; i < d; ++i) for (long j = i; j < d; ++j) {
      cd a = std::conj(m[i * d + j]), b = std::conj(m[j * d + i]);
      m[i * d + j] = b; m[j * d + i] = a;
    }
  }
  // calculateTwoTimesRealPartOfSelf, CMatrix.cc:14-30
  static void twoRe(CV& m, long d) {
    for (long i = 0; i < d; ++i) {
      m[i * d + i] = 2. * m[i * d + i].real();
      for (long j = i + 1; j < d; ++j) { m[i * d + j] = m[i * d + j] + std::conj(m[j * d + i]); m[j * d + i] = std::conj(m[i * d + j]); }
    }
  }

  void derivs(const CV& r, CV& dr, double tt) {
    std::fill(dr.begin(), dr.end(), cd(0, 0));
    // columns of rho: for each column j, the vector rho(:,j) has stride d
    CV col(d), dcol(d);
    for (long j = 0; j < d; ++j) {
      for (long i = 0; i < d; ++i) { col[i] = r[i * d + j]; dcol[i] = 0.; }
      addContribution(*sys, tt, col.data(), dcol.data(), t0);
      for (long i = 0; i < d; ++i) dr[i * d + j] += dcol[i];
    }
    twoRe(dr, d);
    int m = 0;
    for (size_t fi = 0; fi < sys->frees.size(); ++fi) {
      for (auto& jmp : sys->frees[fi].jumps) {
        if (jmp.superop) {
          // fullRange<V0,2> over the density operator: retained (i, i+R), dummies equal
          int R = (int)sys->L.dims.size();
          std::vector<int> others; for (int a = 0; a < R; ++a) if (a != (int)fi) others.push_back(a);
          // iterate over dummy multi-index pairs (rowDummy, colDummy), all combinations
          Layout Lo = sys->L;
          std::vector<int> rd(others.size(), 0);
          long nOther = 1; for (int a : others) nOther *= Lo.dims[a];
          for (long p = 0; p < nOther; ++p) for (long q = 0; q < nOther; ++q) {
            long ro = 0, co = 0, pp = p, qq = q;
            for (int k = (int)others.size() - 1; k >= 0; --k) {
              int a = others[k];
              ro += (pp % Lo.dims[a]) * Lo.strides[a]; pp /= Lo.dims[a];
              co += (qq % Lo.dims[a]) * Lo.strides[a]; qq /= Lo.dims[a];
            }
            jmp.superop(r.data() + ro * d + co, dr.data() + ro * d + co, Lo.strides[fi] * d, Lo.strides[fi]);
          }
        } else {
          // fallback, Master.h:176-182: J applied to columns, h.c., J again
          CV tmp(r);
          CV c(d);
          auto unaryIter = [&]() {
            for (long j = 0; j < d; ++j) {
              for (long i = 0; i < d; ++i) c[i] = tmp[i * d + j];
              actWithJ(*sys, c.data(), m);
              for (long i = 0; i < d; ++i) tmp[i * d + j] = c[i];
            }
          };
          unaryIter(); hcSelf(tmp, d); unaryIter();
          for (size_t i = 0; i < dr.size(); ++i) dr[i] += tmp[i];
        }
        ++m;
      }
    }
  }

  void step(double deltaT) {
    ode.step(deltaT, [this](const CV& y, CV& dy, double tt) { derivs(y, dy, tt); }, t, rho);
    if (sys->hasEx()) {
      // Master.h:187-193: U on the Left slices (columns), hermitianConjugateSelf, U on the Left
      // slices, conj.  hermitianConjugateSelf conjugates the data but transposes only Blitz's
      // view (transposeSelf, ComplexArrayExtensions.h:40-46, 113-116), so the second pass runs
      // over the rows of the stored matrix: rho <- U rho U^dagger.
      CV c(d);
      for (long j = 0; j < d; ++j) {
        for (long i = 0; i < d; ++i) c[i] = rho[i * d + j];
        actWithU(*sys, t, c.data(), t0);
        for (long i = 0; i < d; ++i) rho[i * d + j] = c[i];
      }
      for (auto& x : rho) x = std::conj(x);
      for (long i = 0; i < d; ++i) actWithU(*sys, t, rho.data() + i * d, t0);  // rows: unit stride
      for (auto& x : rho) x = std::conj(x);
    }
    t0 = t;
    twoRe(rho, d);
    double tr = 0; for (long i = 0; i < d; ++i) tr += rho[i * d + i].real();
    for (auto& x : rho) x /= tr;
  }
  void advance(double deltaT) { double endTime = t + deltaT; while (double dt = endTime - t) step(dt); }
};

}  // namespace orc

// AutostopHandlerGeneric (CPPQEDutils/Trajectory.h:179-222): called with every streamed array.
// The first call only sizes the running average (the array itself is not queued); then the
// queue fills to `rep` arrays with their running mean; from there on, a streamed array whose
// max_i |avg_i - v_i| / (|avg_i| + |v_i|) is below eps stops the run, otherwise it replaces the
// oldest array of the moving average.  Entries below epsAbs in modulus are replaced by epsAbs.
struct Autostop {
  double eps, epsAbs; unsigned rep;
  bool started = false;
  std::vector<double> avg;
  std::deque<std::vector<double>> q;
  Autostop(double e, double ea, unsigned r) : eps(e), epsAbs(ea), rep(r) {}
  bool operator()(const double* v, int n) {  // true: StoppingCriterionReachedException
    std::vector<double> buf(v, v + n);
    for (auto& x : buf) if (std::fabs(x) < epsAbs) x = epsAbs;
    if (!rep) return false;
    if (!started) { started = true; avg.assign(n, 0.); return false; }
    if (q.size() == rep) {
      double m = -INFINITY;
      for (int i = 0; i < n; ++i) m = std::max(m, std::fabs(avg[i] - buf[i]) / (std::fabs(avg[i]) + std::fabs(buf[i])));
      if (m < eps) return true;
      for (int i = 0; i < n; ++i) avg[i] = avg[i] + (buf[i] - q.front()[i]) / double(rep);
      q.pop_front();
    } else {
      const double k = (double)q.size();
      for (int i = 0; i < n; ++i) avg[i] = (k * avg[i] + buf[i]) / (k + 1.);
    }
    q.push_back(buf);
    return false;
  }
};

// ===========================================================================
// C ABI for the tests (ctypes)
// ===========================================================================

using namespace orc;

extern "C" {

struct orc_params {
  double epsRel, epsAbs, dpLimit, overshootTol, dtInit;  // dtInit <= 0: 1/(10*highestFrequency)
  int rng;          // 0 = pcg64 (golden replay), 1 = Philox counter
  int controller;   // 0 = odeint, 1 = gsl
  unsigned long long seed, stream;
};

// run modes, Trajectory.h:462-610: mode 0 = dc (dc steps), 1 = Dt up to T, 2 = NDt intervals of Dt
struct orc_run {
  int mode; int dc; double Dt; long NDt; double T;
};

const char* orc_last_error() { return g_last_error.c_str(); }

void* orc_system_create(const char* spec) {
  try { return buildSystem(spec); } catch (std::exception& e) { g_last_error = e.what(); return nullptr; }
}
void orc_system_destroy(void* s) { delete (System*)s; }
long orc_system_dim(void* s) { return ((System*)s)->L.total; }
int orc_system_njumps(void* s) { return ((System*)s)->nJumps(); }
int orc_system_navr(void* s) { return ((System*)s)->nAvr(); }
double orc_system_highest_frequency(void* s) { return ((System*)s)->highestFrequency(); }

// dpsidt += (H/i) psi   (structure::Hamiltonian::addContribution contract)
int orc_apply_heff(void* s, double t, double t0, const double* psi, double* dpsidt) {
  try { addContribution(*(System*)s, t, (const cd*)psi, (cd*)dpsidt, t0); return 0; }
  catch (std::exception& e) { g_last_error = e.what(); return 1; }
}
int orc_rates(void* s, const double* psi, double* out) {
  try { auto r = rates(*(System*)s, (const cd*)psi); std::copy(r.begin(), r.end(), out); return 0; }
  catch (std::exception& e) { g_last_error = e.what(); return 1; }
}
int orc_act_jump(void* s, double* psi, int m) { actWithJ(*(System*)s, (cd*)psi, m); return 0; }
int orc_averages(void* s, const double* psi, double* out) {
  averagesRaw(*(System*)s, (const cd*)psi, out); processAverages(*(System*)s, out); return 0;
}

void orc_pcg64_uniform(unsigned long long seed, unsigned long long stream, long n, double* out, unsigned long long* raw) {
  Pcg64 p(seed, stream);
  for (long i = 0; i < n; ++i) { if (raw) { Pcg64 q = p; raw[i] = q.next(); } out[i] = p.uniform(); }
}
void orc_philox4x64(const unsigned long long* ctr, const unsigned long long* key, unsigned long long* out) {
  uint64_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]}, k[2] = {key[0], key[1]}, o[4];
  philox4x64_10(c, k, o);
  for (int i = 0; i < 4; ++i) out[i] = o[i];
}
double orc_counter_uniform(unsigned long long seed, unsigned long long traj, unsigned long long step) {
  return counter_uniform(seed, traj, step);
}

// Single trajectory run.  rows: (time, dtDid, averages...) per stream; returns #rows (<= cap)
// jumps_t/jumps_ch: jump log (<= jcap); counters[0..7] = MCWF steps, ODE steps, failed, derivs,
// tries, step-backs, overshoots, #jumps.  psi_out: final state.
long orc_mcwf_run(void* sp, const double* psi0, const orc_params* p, const orc_run* run,
                  double* rows, long cap, double* jumps_t, int* jumps_ch, long jcap,
                  long long* counters, double* psi_out) {
  try {
    System* s = (System*)sp;
    CV psi((const cd*)psi0, (const cd*)psi0 + s->L.total);
    double dtInit = p->dtInit > 0 ? p->dtInit : 1. / (10. * s->highestFrequency());
    MCWF tr(s, psi, dtInit, p->epsRel, p->epsAbs, p->dpLimit, p->overshootTol, (RngKind)p->rng, p->seed, p->stream, (Controller)p->controller);
    int nav = s->nAvr(); long nrow = 0;
    auto stream = [&]() {
      if (nrow < cap) {
        double* r = rows + nrow * (2 + nav);
        r[0] = tr.t; r[1] = tr.ode.dtDid;
        averagesRaw(*s, tr.psi.data(), r + 2); processAverages(*s, r + 2);
      }
      ++nrow;
    };
    bool arrayStreamed = false;
    if (run->mode == 0) {
      for (long count = 0; tr.t < run->T; ++count) {
        if (count) { tr.step(run->T - tr.t); arrayStreamed = false; }
        if (!count || !(count % run->dc)) { stream(); arrayStreamed = true; }
      }
      if (!arrayStreamed) stream();
    } else if (run->mode == 1) {
      for (long count = 0; tr.t < run->T; ++count) {
        if (count) tr.advance(std::min(run->Dt, run->T - tr.t));
        stream();
      }
    } else {
      for (long count = 0; count <= run->NDt; ++count) { if (count) tr.advance(run->Dt); stream(); }
    }
    long nj = std::min((long)tr.jumps.size(), jcap);
    for (long i = 0; i < nj; ++i) { if (jumps_t) jumps_t[i] = tr.jumps[i].t; if (jumps_ch) jumps_ch[i] = tr.jumps[i].ch; }
    if (counters) {
      counters[0] = tr.nMCWFSteps; counters[1] = tr.ode.st.nSteps; counters[2] = tr.ode.st.nFailed; counters[3] = tr.ode.st.nDerivs;
      counters[4] = tr.ode.st.nTries; counters[5] = tr.nStepBacks; counters[6] = tr.nOvershot; counters[7] = (long long)tr.jumps.size();
    }
    if (psi_out) std::memcpy(psi_out, tr.psi.data(), sizeof(cd) * tr.psi.size());
    return nrow;
  } catch (std::exception& e) { g_last_error = e.what(); return -1; }
}

// Ensemble (makeEnsemble + Ensemble::advance): trajectory k uses stream p->stream + k
// (pcg64) or trajectory index p->stream + k (Philox). Dt / NDt modes only (Q12).
// rows: (time, mean dtDid, averages of the ensemble density operator, processed)
long orc_ensemble_run(void* sp, const double* psi0, const orc_params* p, const orc_run* run, long nTraj,
                      double* rows, long cap, long long* tries_out) {
  try {
    System* s = (System*)sp;
    CV psi((const cd*)psi0, (const cd*)psi0 + s->L.total);
    double dtInit = p->dtInit > 0 ? p->dtInit : 1. / (10. * s->highestFrequency());
    std::vector<std::unique_ptr<MCWF>> trajs;
    for (long k = 0; k < nTraj; ++k)
      trajs.emplace_back(new MCWF(s, psi, dtInit, p->epsRel, p->epsAbs, p->dpLimit, p->overshootTol, (RngKind)p->rng, p->seed, p->stream + k, (Controller)p->controller));
    int nav = s->nAvr(); long nrow = 0;
    std::vector<double> tmp(nav);
    auto stream = [&]() {
      if (nrow < cap) {
        double* r = rows + nrow * (2 + nav);
        r[0] = trajs.front()->t;
        double dsum = 0; for (auto& tr : trajs) dsum += tr->ode.dtDid;
        r[1] = dsum / (double)nTraj;
        for (int q = 0; q < nav; ++q) r[2 + q] = 0.;
        for (auto& tr : trajs) { averagesRaw(*s, tr->psi.data(), tmp.data()); for (int q = 0; q < nav; ++q) r[2 + q] += tmp[q]; }
        for (int q = 0; q < nav; ++q) r[2 + q] /= (double)nTraj;
        processAverages(*s, r + 2);
      }
      ++nrow;
    };
    auto advanceAll = [&](double dT) { for (auto& tr : trajs) tr->advance(dT); };
    if (run->mode == 2) { for (long count = 0; count <= run->NDt; ++count) { if (count) advanceAll(run->Dt); stream(); } }
    else { for (long count = 0; trajs.front()->t < run->T; ++count) { if (count) advanceAll(std::min(run->Dt, run->T - trajs.front()->t)); stream(); } }
    if (tries_out) { long long tsum = 0; for (auto& tr : trajs) tsum += tr->ode.st.nTries; *tries_out = tsum; }
    return nrow;
  } catch (std::exception& e) { g_last_error = e.what(); return -1; }
}

static long masterRun(void* sp, const double* rho0, const orc_params* p, const orc_run* run,
                      double* rows, long cap, double* rho_out, Autostop* as) {
  try {
    System* s = (System*)sp;
    long d = s->L.total;
    CV rho((const cd*)rho0, (const cd*)rho0 + d * d);
    double dtInit = p->dtInit > 0 ? p->dtInit : 1. / (10. * s->highestFrequency());
    Master m(s, rho, dtInit, p->epsRel, p->epsAbs);
    m.ode.controller = (Controller)p->controller;
    int nav = s->nAvr(); long nrow = 0;
    std::vector<double> av(nav);
    struct Stop {};
    auto stream = [&](bool handler = true) {
      averagesRawRho(*s, m.rho.data(), av.data()); processAverages(*s, av.data());
      if (nrow < cap) { double* r = rows + nrow * (2 + nav); r[0] = m.t; r[1] = m.ode.dtDid; std::copy(av.begin(), av.end(), r + 2); }
      ++nrow;
      if (handler && as && (*as)(av.data(), nav)) throw Stop{};  // Trajectory.h:583-589
    };
    bool arrayStreamed = false;
    try {
      if (run->mode == 0) {
        for (long count = 0; m.t < run->T; ++count) {
          if (count) { m.step(run->T - m.t); arrayStreamed = false; }
          if (!count || !(count % run->dc)) { arrayStreamed = true; stream(); }
        }
      } else if (run->mode == 1) {
        for (long count = 0; m.t < run->T; ++count) { if (count) m.advance(std::min(run->Dt, run->T - m.t)); arrayStreamed = true; stream(); }
      } else {
        for (long count = 0; count <= run->NDt; ++count) { if (count) m.advance(run->Dt); arrayStreamed = true; stream(); }
      }
    } catch (Stop&) {}
    if (!arrayStreamed) stream(false);  // the closing row goes to no handler (Trajectory.h:594)
    if (rho_out) std::memcpy(rho_out, m.rho.data(), sizeof(cd) * m.rho.size());
    return nrow;
  } catch (std::exception& e) { g_last_error = e.what(); return -1; }
}

long orc_master_run(void* sp, const double* rho0, const orc_params* p, const orc_run* run,
                    double* rows, long cap, double* rho_out) {
  return masterRun(sp, rho0, p, run, rows, cap, rho_out, nullptr);
}

// trajectory::run with AutostopHandlerGeneric (Trajectory.h:337): the rows end with the row at
// which the criterion was met
long orc_master_run_autostop(void* sp, const double* rho0, const orc_params* p, const orc_run* run,
                             double eps, double epsAbs, unsigned rep, double* rows, long cap, double* rho_out) {
  Autostop as(eps, epsAbs, rep);
  return masterRun(sp, rho0, p, run, rows, cap, rho_out, &as);
}

void* orc_autostop_create(double eps, double epsAbs, unsigned rep) { return new Autostop(eps, epsAbs, rep); }
void orc_autostop_destroy(void* h) { delete (Autostop*)h; }
int orc_autostop_feed(void* h, const double* v, int n) { return (*(Autostop*)h)(v, n) ? 1 : 0; }

// A classical ODE on the same engine: the reference's stepper tests drive
// Simulated<CArray<1>> (CPPQEDutils/Simulated_.h:26-79; Simulated::step = ode_.step) through
// run() in Dt mode or advanceTo (Trajectory.h:44-69).  which = 0: the Evolved equation of
// EvolvedTest.cc / EvolvedTestGSL.cc, y'' = y'^2 + Z y' + Z^2 exp(2 Z t) (y - y^2), Z = -1+10i,
// y(0) = e, y'(0) = Z e (solution exp(exp(Z t))); which = 1: the driven damped oscillator of
// HarmonicOscillatorComplex.cc, y'' = exp(i omega t) - 2 gamma y' - y, omega = 5, gamma = 0.1,
// y(0) = 10+5i, y'(0) = -1+i.  dtInit and the user tolerances as in those tests.
// times == nullptr: run up to T (Trajectory.h:462-610) in dc mode when dc > 0 (a stream point
// every dc ODE steps: run(S, 70., 1, ...) of HarmonicOscillatorComplex.cc passes an integer
// stream frequency) or else in Dt mode, one row (t, dtDid, Re y, Im y, Re y', Im y') per stream
// point; otherwise advanceTo(times[i]) for i < n (HarmonicOscillatorRandomTimes.cc:43-50), one
// row after each.  stats: derivative calls, steps, failed tries, tries.
long orc_simulated_run(int which, int controller, double dtInit, double epsRel, double epsAbs, double T, double Dt,
                       int dc, const double* times, long n, double* rows, long cap, long long* stats) {
  try {
    const cd Z(-1., 10.), I(0., 1.);
    const double omega = 5., gamma = 0.1;
    CV y(2);
    if (which == 0) { y[0] = std::exp(1.); y[1] = Z * std::exp(1.); }
    else if (which == 1) { y[0] = cd(10., 5.); y[1] = cd(-1., 1.); }
    else throw std::invalid_argument("unknown test equation");
    auto f = [&](const CV& x, CV& dx, double tau) {
      dx[0] = x[1];
      if (which == 0) dx[1] = x[1] * x[1] + Z * x[1] + Z * Z * std::exp(2. * Z * tau) * (x[0] - x[0] * x[0]);
      else dx[1] = std::exp(I * omega * tau) - 2. * gamma * x[1] - x[0];
    };
    Ode ode(dtInit, epsRel, epsAbs);
    ode.controller = (Controller)controller;
    double t = 0.;
    long nrow = 0;
    auto row = [&]() {
      if (nrow < cap) {
        double* r = rows + 6 * nrow;
        r[0] = t; r[1] = ode.dtDid; r[2] = y[0].real(); r[3] = y[0].imag(); r[4] = y[1].real(); r[5] = y[1].imag();
      }
      ++nrow;
    };
    auto advance = [&](double deltaT) { double endTime = t + deltaT; while (double dt = endTime - t) ode.step(dt, f, t, y); };
    if (!times && dc > 0) {
      bool streamed = false;
      for (long count = 0; t < T; ++count) {
        if (count) { ode.step(T - t, f, t, y); streamed = false; }
        if (!count || !(count % dc)) { row(); streamed = true; }
      }
      if (!streamed) row();
    } else if (!times) {
      for (long count = 0; t < T; ++count) {
        if (count) advance(std::min(Dt, T - t));
        row();
      }
    } else {
      for (long i = 0; i < n; ++i) { advance(times[i] - t); row(); }
    }
    if (stats) { stats[0] = ode.st.nDerivs; stats[1] = ode.st.nSteps; stats[2] = ode.st.nFailed; stats[3] = ode.st.nTries; }
    return nrow;
  } catch (std::exception& e) { g_last_error = e.what(); return -1; }
}

// Slice iteration as the element operators see it (forSlices, SliceIterator.h:277-354): for an
// array of extents dims[rank] and retained axes V[nV], each slice's offset in iteration order
// (at most cap) and the retained axes' strides in V's order.  Returns the number of slices.
long orc_slices(const int* dims, int rank, const int* V, int nV, long* offsets, long cap, long* vstrides) {
  try {
    orc::Layout L(std::vector<int>(dims, dims + rank));
    long n = 0;
    orc::forSlices(L, std::vector<int>(V, V + nV), [&](long off, const std::vector<long>& vs) {
      if (n < cap) offsets[n] = off;
      if (n == 0) for (int k = 0; k < nV; ++k) vstrides[k] = vs[k];
      ++n;
    });
    return n;
  } catch (std::exception& e) { g_last_error = e.what(); return -1; }
}

// std::uniform_real_distribution<double>(a, b) over std::mt19937(seed), the sample times of
// HarmonicOscillatorRandomTimes.cc:40-45 (libstdc++'s generate_canonical)
void orc_mt19937_uniform(unsigned seed, long n, double a, double b, double* out) {
  std::mt19937 re{seed};
  for (long i = 0; i < n; ++i) out[i] = std::uniform_real_distribution<double>(a, b)(re);
}

// ---------------------------------------------------------------------------
// Continuing from a saved state (trajectory::run with a state file, Trajectory.h:484-500:
// the trajectory is constructed as usual, then every archived field overwrites it;
// MCWF_Trajectory::stateIO sets t0 = t, MCWF_Trajectory.h:136-143, Master.h:93-99).
// The continuing loop streams no row at count 0 (Trajectory.h:576-585).

struct orc_traj_state {
  double t, dtDid, dtTry;
  double dpMaxOvershoot, dpTolMaxOvershoot, normMaxDev;
  unsigned long long pcg[4];   // pcg64 state (high, low word), increment (high, low word)
  long long ode[3];            // ODE logger: derivative calls, steps, failed steps (ODE.h:117)
  long long mcwf[3];           // MCWF logger: steps, dpLimit overshoots, tolerance overshoots
  unsigned long long ordinal;  // Philox step ordinal
};

static void continueLoop(const orc_run* run, double& t, const std::function<void(double)>& advance,
                         const std::function<void(double)>& step, const std::function<void()>& stream) {
  if (run->mode == 0) {
    bool streamed = false;
    for (long count = 0; t < run->T; ++count) {
      if (count) { step(run->T - t); streamed = false; if (!(count % run->dc)) { stream(); streamed = true; } }
    }
    if (!streamed) stream();
  } else if (run->mode == 1) {
    for (long count = 0; t < run->T; ++count) if (count) { advance(std::min(run->Dt, run->T - t)); stream(); }
  } else {
    for (long count = 1; count <= run->NDt; ++count) { advance(run->Dt); stream(); }
  }
}

// Continues one MCWF trajectory from an archived state; st is read and then
// overwritten with the final state.  jt/jch hold *nj archived jumps on entry and
// the whole jump list (at most jcap) on return.
long orc_mcwf_resume(void* sp, const double* psi0, const orc_params* p, const orc_run* run, orc_traj_state* st,
                     double* jt, int* jch, long* nj, long jcap, double* rows, long cap, double* psi_out) {
  try {
    System* s = (System*)sp;
    CV psi((const cd*)psi0, (const cd*)psi0 + s->L.total);
    double dtInit = p->dtInit > 0 ? p->dtInit : 1. / (10. * s->highestFrequency());
    MCWF tr(s, psi, dtInit, p->epsRel, p->epsAbs, p->dpLimit, p->overshootTol, (RngKind)p->rng, p->seed, p->stream,
            (Controller)p->controller);
    tr.psi = psi;
    tr.t = tr.t0 = st->t; tr.ode.dtDid = st->dtDid; tr.ode.dtTry = st->dtTry;
    tr.dpMaxOvershoot = st->dpMaxOvershoot; tr.dpTolMaxOvershoot = st->dpTolMaxOvershoot; tr.maxNormDev = st->normMaxDev;
    tr.pcg.state = ((unsigned __int128)st->pcg[0] << 64) | st->pcg[1];
    tr.pcg.inc = ((unsigned __int128)st->pcg[2] << 64) | st->pcg[3];
    tr.stepOrdinal = st->ordinal;
    tr.ode.st.nDerivs = st->ode[0]; tr.ode.st.nSteps = st->ode[1]; tr.ode.st.nFailed = st->ode[2];
    tr.nMCWFSteps = st->mcwf[0]; tr.nOvershot = st->mcwf[1]; tr.nStepBacks = st->mcwf[2];
    tr.jumps.clear();
    for (long i = 0; i < *nj; ++i) tr.jumps.push_back({jt[i], jch[i]});
    int nav = s->nAvr(); long nrow = 0;
    auto stream = [&]() {
      if (nrow < cap) {
        double* r = rows + nrow * (2 + nav);
        r[0] = tr.t; r[1] = tr.ode.dtDid;
        averagesRaw(*s, tr.psi.data(), r + 2); processAverages(*s, r + 2);
      }
      ++nrow;
    };
    continueLoop(run, tr.t, [&](double dT) { tr.advance(dT); }, [&](double dT) { tr.step(dT); }, stream);
    st->t = tr.t; st->dtDid = tr.ode.dtDid; st->dtTry = tr.ode.dtTry;
    st->dpMaxOvershoot = tr.dpMaxOvershoot; st->dpTolMaxOvershoot = tr.dpTolMaxOvershoot; st->normMaxDev = tr.maxNormDev;
    st->pcg[0] = (unsigned long long)(tr.pcg.state >> 64); st->pcg[1] = (unsigned long long)tr.pcg.state;
    st->pcg[2] = (unsigned long long)(tr.pcg.inc >> 64); st->pcg[3] = (unsigned long long)tr.pcg.inc;
    st->ordinal = tr.stepOrdinal;
    st->ode[0] = tr.ode.st.nDerivs; st->ode[1] = tr.ode.st.nSteps; st->ode[2] = tr.ode.st.nFailed;
    st->mcwf[0] = tr.nMCWFSteps; st->mcwf[1] = tr.nOvershot; st->mcwf[2] = tr.nStepBacks;
    *nj = std::min((long)tr.jumps.size(), jcap);
    for (long i = 0; i < *nj; ++i) { jt[i] = tr.jumps[i].t; jch[i] = tr.jumps[i].ch; }
    if (psi_out) std::memcpy(psi_out, tr.psi.data(), sizeof(cd) * tr.psi.size());
    return nrow;
  } catch (std::exception& e) { g_last_error = e.what(); return -1; }
}

struct orc_master_state { double t, dtDid, dtTry; long long ode[3]; };

long orc_master_resume(void* sp, const double* rho0, const orc_params* p, const orc_run* run, orc_master_state* st,
                       double* rows, long cap, double* rho_out) {
  try {
    System* s = (System*)sp;
    long d = s->L.total;
    CV rho((const cd*)rho0, (const cd*)rho0 + d * d);
    double dtInit = p->dtInit > 0 ? p->dtInit : 1. / (10. * s->highestFrequency());
    Master m(s, rho, dtInit, p->epsRel, p->epsAbs);
    m.ode.controller = (Controller)p->controller;
    m.t = m.t0 = st->t; m.ode.dtDid = st->dtDid; m.ode.dtTry = st->dtTry;
    m.ode.st.nDerivs = st->ode[0]; m.ode.st.nSteps = st->ode[1]; m.ode.st.nFailed = st->ode[2];
    int nav = s->nAvr(); long nrow = 0;
    auto stream = [&]() {
      if (nrow < cap) { double* r = rows + nrow * (2 + nav); r[0] = m.t; r[1] = m.ode.dtDid; averagesRawRho(*s, m.rho.data(), r + 2); processAverages(*s, r + 2); }
      ++nrow;
    };
    continueLoop(run, m.t, [&](double dT) { m.advance(dT); }, [&](double dT) { m.step(dT); }, stream);
    st->t = m.t; st->dtDid = m.ode.dtDid; st->dtTry = m.ode.dtTry;
    st->ode[0] = m.ode.st.nDerivs; st->ode[1] = m.ode.st.nSteps; st->ode[2] = m.ode.st.nFailed;
    if (rho_out) std::memcpy(rho_out, m.rho.data(), sizeof(cd) * m.rho.size());
    return nrow;
  } catch (std::exception& e) { g_last_error = e.what(); return -1; }
}

}  // extern "C"
