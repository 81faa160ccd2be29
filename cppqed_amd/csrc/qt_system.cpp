// qt_system.cpp -- element construction and lowering to ELL tables (host).
#include "qt_system.hpp"

#include <algorithm>
#include <map>
#include <cmath>
#include <memory>
#include <sstream>
#include <stdexcept>

namespace qt {

// ---------------------------------------------------------------------------
// sparse local-operator algebra
// ---------------------------------------------------------------------------

LocalOp kron(const LocalOp& a, const LocalOp& b) {
  LocalOp r;
  r.dims = a.dims;
  r.dims.insert(r.dims.end(), b.dims.begin(), b.dims.end());
  int nb = b.size();
  for (auto& x : a.entries)
    for (auto& y : b.entries)
      r.entries.push_back({x.row * nb + y.row, x.col * nb + y.col, x.val * y.val, x.freq + y.freq});
  return r;
}

LocalOp adjoint(const LocalOp& a) {
  LocalOp r;
  r.dims = a.dims;
  for (auto& x : a.entries) r.entries.push_back({x.col, x.row, std::conj(x.val), -x.freq});
  return r;
}

LocalOp scale(const LocalOp& a, cd f) {
  LocalOp r = a;
  for (auto& x : r.entries) x.val = x.val * f;
  return r;
}

LocalOp plus(const LocalOp& a, const LocalOp& b) {
  if (a.dims != b.dims) throw std::invalid_argument("operator dimension mismatch");
  LocalOp r = a;
  r.entries.insert(r.entries.end(), b.entries.begin(), b.entries.end());
  return r;
}

// the product a b of two operators on the same axes (entries summed per row and column; frequencies
// of the factors' entries add, as the entries' phases multiply)
static LocalOp matmul(const LocalOp& a, const LocalOp& b) {
  if (a.dims != b.dims) throw std::invalid_argument("operator dimension mismatch");
  std::map<std::pair<int, int>, std::pair<cd, cd>> acc;
  std::vector<std::vector<const LocalOp::Entry*>> bRows(b.size());
  for (auto& y : b.entries) bRows[y.row].push_back(&y);
  for (auto& x : a.entries)
    for (const LocalOp::Entry* y : bRows[x.col]) {
      auto& e = acc[{x.row, y->col}];
      e.first += x.val * y->val;
      e.second = x.freq + y->freq;
    }
  LocalOp r; r.dims = a.dims;
  for (auto& kv : acc) r.entries.push_back({kv.first.first, kv.first.second, kv.second.first, kv.second.second});
  return r;
}

static LocalOp diagonal(int dim, const std::vector<cd>& d) {
  LocalOp r; r.dims = {dim};
  for (int n = 0; n < dim; ++n) r.entries.push_back({n, n, d[n], 0.});
  return r;
}

// annihilation operator a = sum_n sqrt(n+1) |n><n+1|   (Mode.cc:277-282)
static LocalOp destroy(int dim) {
  LocalOp r; r.dims = {dim};
  for (int n = 0; n + 1 < dim; ++n) r.entries.push_back({n, n + 1, cd(std::sqrt(n + 1.), 0.), 0.});
  return r;
}

static bool nonzero(cd z) { return z.real() != 0. || z.imag() != 0.; }

// H/i of a (possibly pumped, damped) mode: -z n  +  eta a^dagger - eta^* a   (Mode.cc:56-86)
static LocalOp modeHOverI(cd z, cd eta, int dim, bool& empty) {
  LocalOp r; r.dims = {dim};
  empty = true;
  if (nonzero(z)) {
    std::vector<cd> d(dim);
    for (int n = 0; n < dim; ++n) d[n] = cd((double)n, 0.) * (-z);
    r = plus(r, diagonal(dim, d)); empty = false;
  }
  if (nonzero(eta)) {
    LocalOp ad = scale(adjoint(destroy(dim)), eta);
    r = plus(r, plus(ad, scale(adjoint(ad), -1.)));
    empty = false;
  }
  return r;
}

// ---------------------------------------------------------------------------
// spec parsing
// ---------------------------------------------------------------------------

using KV = std::map<std::string, std::string>;

static double num(const KV& kv, const std::string& k, double def) {
  auto it = kv.find(k);
  return it == kv.end() ? def : std::stod(it->second);
}

static std::string picture(const KV& kv) {
  auto it = kv.find("picture");
  std::string p = it == kv.end() ? "Sch" : it->second;
  if (p != "Sch" && p != "IP" && p != "UIP") throw std::invalid_argument("unknown picture " + p);
  return p;
}

struct Builder {
  System& s;
  explicit Builder(System& sys) : s(sys) {}
  std::vector<std::vector<Term>> freeJumps;  // per free
  std::vector<std::vector<Observable>> freeObs;
  std::vector<std::vector<int>> freeVar;
  std::vector<std::vector<Term>> iaJumps;

  void avMode(int ax, int dim) {
    std::vector<cd> n(dim), n2(dim);
    for (int k = 0; k < dim; ++k) { n[k] = (double)k; n2[k] = (double)k * k; }
    Observable on, on2, are, aim;
    on.terms.push_back({{ax}, diagonal(dim, n)});
    on2.terms.push_back({{ax}, diagonal(dim, n2)});
    are.terms.push_back({{ax}, destroy(dim)});
    aim = are; aim.imagPart = true;
    freeObs.back() = {on, on2, are, aim};
    freeVar.back() = {-1, 0, -1, -1};  // VAR(n) = <n^2> - <n>^2  (Mode.cc:198-201)
    freeSqrt.back() = {0, 0, 0, 0};
  }
  void avQbit(int ax) {
    Observable p0, p1, sre, sim;
    p0.terms.push_back({{ax}, diagonal(2, {1., 0.})});
    p1.terms.push_back({{ax}, diagonal(2, {0., 1.})});
    LocalOp sig; sig.dims = {2}; sig.entries.push_back({0, 1, 1., 0.});
    sre.terms.push_back({{ax}, sig});
    sim = sre; sim.imagPart = true;
    freeObs.back() = {p0, p1, sre, sim};
    freeVar.back() = {-1, -1, -1, -1};
    freeSqrt.back() = {0, 0, 0, 0};
  }

  // mode::make(ParsPumpedLossy, picture) for nTh = 0 (Mode.tcc:40-60, 160-197)
  void mode(const KV& kv) {
    int ax = (int)s.frees.size();
    int dim = (int)num(kv, "cutoff", 10);
    double delta = num(kv, "delta", -10), kappa = num(kv, "kappa", 0);
    cd eta(num(kv, "eta", 0), num(kv, "etaI", 0));
    std::string pic = picture(kv);
    FreeInfo f; f.kind = "mode"; f.dim = dim; f.avKind = "mode";
    cd z(kappa, -delta);
    bool empty;
    LocalOp h;
    double hf = 0.;
    if (pic == "Sch") {
      h = modeHOverI(z, eta, dim, empty);
      if (kappa && nonzero(eta)) hf = std::abs(std::conj(z) * (double)dim);
      else if (kappa) hf = std::max(std::fabs(kappa * dim), std::fabs(delta * dim));
      else hf = std::fabs(delta * dim);
      if (nonzero(eta)) hf = std::max(hf, std::abs(eta * std::sqrt((double)dim)));
    } else {
      if (!kappa || !nonzero(eta)) throw std::invalid_argument("interaction-picture mode needs kappa and eta");
      f.hasEx = true;
      cd zI;
      if (pic == "IP") { h = modeHOverI(0., eta, dim, empty); zI = z; hf = std::abs(std::conj(z)); }
      else { h = modeHOverI(cd(kappa, 0.), eta, dim, empty); zI = cd(0., -delta); hf = std::max(std::fabs(kappa * dim), std::fabs(delta)); }
      for (int n = 0; n < dim; ++n) f.exd.push_back(zI * (double)n);  // mode::Exact, Mode.cc:42-45
      hf = std::max(hf, std::abs(eta * std::sqrt((double)dim)));
    }
    f.hasHa = true;  // every mode::Hamiltonian carries at least the (zero) Kerr term (Mode.cc:92-107)
    f.highestFrequency = hf;
    s.hTerms.push_back({{ax}, h});
    freeJumps.emplace_back(); freeObs.emplace_back(); freeVar.emplace_back(); freeSqrt.emplace_back();
    if (kappa) {  // sqrt(2 kappa) a  (Mode.cc:118-125)
      LocalOp j; j.dims = {dim};
      double fact = std::sqrt(2. * kappa);
      for (int n = 0; n + 1 < dim; ++n) j.entries.push_back({n, n + 1, cd(fact * std::sqrt(n + 1.), 0.), 0.});
      freeJumps.back().push_back({{ax}, j});
      f.nJumps = 1;
    }
    avMode(ax, dim);
    s.frees.push_back(f);
  }

  // qbit::make(ParsPumpedLossy, picture): a two-level mode pumped with -eta (Qbit.cc:217-246, Qbit_.h:99-118)
  void qbit(const KV& kv) {
    int ax = (int)s.frees.size();
    double delta = num(kv, "deltaA", -10), gamma = num(kv, "gamma", 0);
    cd eta(num(kv, "eta", 0), num(kv, "etaI", 0));
    std::string pic = picture(kv);
    FreeInfo f; f.kind = "qbit"; f.dim = 2; f.avKind = "qbit";
    cd z(gamma, -delta);
    bool empty;
    LocalOp h;
    double hf = 0.;
    if (pic == "Sch") {
      h = modeHOverI(z, -eta, 2, empty);
      if (gamma && nonzero(eta)) hf = std::max(std::abs(std::conj(z)), std::abs(eta));
      else if (gamma) hf = std::max(std::fabs(gamma), std::fabs(delta));
      else hf = std::max(std::fabs(delta), std::abs(eta));
    } else {
      if (!gamma || !nonzero(eta)) throw std::invalid_argument("interaction-picture qbit needs gamma and eta");
      f.hasEx = true;
      cd zI;
      if (pic == "IP") { h = modeHOverI(0., -eta, 2, empty); zI = z; hf = std::abs(std::conj(z)); }
      else { h = modeHOverI(cd(gamma, 0.), -eta, 2, empty); zI = cd(0., -delta); hf = std::max(std::fabs(gamma), std::fabs(delta)); }
      f.exd = {0., zI};
      hf = std::max(hf, std::abs(eta));
    }
    f.hasHa = true;
    f.highestFrequency = hf;
    s.hTerms.push_back({{ax}, h});
    freeJumps.emplace_back(); freeObs.emplace_back(); freeVar.emplace_back(); freeSqrt.emplace_back();
    if (gamma) {
      LocalOp j; j.dims = {2};
      j.entries.push_back({0, 1, cd(std::sqrt(2. * gamma) * std::sqrt(1.), 0.), 0.});
      freeJumps.back().push_back({{ax}, j});
      f.nJumps = 1;
    }
    avQbit(ax);
    s.frees.push_back(f);
  }

  // PumpedTwoLevelAtomSch: dense 2x2 H/i = [[0, eta^*], [-eta, za]], za = (-gamma, deltaA)
  // (PumpedTwoLevelAtom.cc:13-49)
  void ptla(const KV& kv) {
    int ax = (int)s.frees.size();
    double gamma = num(kv, "gamma", 10), delta = num(kv, "deltaA", -10);
    cd eta(num(kv, "eta", 0), num(kv, "etaI", 0));
    cd za(-gamma, delta);
    FreeInfo f; f.kind = "ptla"; f.dim = 2; f.avKind = "qbit"; f.hasHa = true;
    f.highestFrequency = std::max(std::abs(cd(gamma, delta)), std::abs(eta));
    LocalOp h; h.dims = {2};
    h.entries.push_back({0, 1, std::conj(eta), 0.});
    h.entries.push_back({1, 0, -eta, 0.});
    h.entries.push_back({1, 1, za, 0.});
    s.hTerms.push_back({{ax}, h});
    freeJumps.emplace_back(); freeObs.emplace_back(); freeVar.emplace_back(); freeSqrt.emplace_back();
    LocalOp j; j.dims = {2};
    j.entries.push_back({0, 1, cd(std::sqrt(-2. * za.real()), 0.), 0.});
    freeJumps.back().push_back({{ax}, j});
    f.nJumps = 1;
    avQbit(ax);
    s.frees.push_back(f);
  }

  const FreeInfo& freeAt(int i) const {
    if (i < 0 || i >= (int)s.frees.size()) throw std::invalid_argument("interaction refers to an undefined free");
    return s.frees[i];
  }

  // JaynesCummings: conj(g) sigma (x) a^dagger - h.c.   (JaynesCummings_.h:54-66)
  void jc(int q, int m, const KV& kv) {
    const FreeInfo& fq = freeAt(q); const FreeInfo& fm = freeAt(m);
    cd g(num(kv, "g", 0), num(kv, "gI", 0));
    if (nonzero(g)) {
      LocalOp t = kron(scale(destroy(fq.dim), std::conj(g)), adjoint(destroy(fm.dim)));
      s.hTerms.push_back({{q, m}, plus(t, scale(adjoint(t), -1.))});
    } else {
      LocalOp t; t.dims = {fq.dim, fm.dim};
      s.hTerms.push_back({{q, m}, t});
    }
    iaHF = std::max(iaHF, std::abs(g * std::sqrt((double)fq.dim * fm.dim)));
  }

  // CoupledModes<CM_XX_RWA>: conj(u) a_1 a_2^dagger - h.c.   (CoupledModes.cc:36-41)
  void bs(int m1, int m2, const KV& kv) {
    const FreeInfo& f1 = freeAt(m1); const FreeInfo& f2 = freeAt(m2);
    cd u(num(kv, "g", 0), num(kv, "gI", 0));
    LocalOp t = kron(scale(destroy(f1.dim), std::conj(u)), adjoint(destroy(f2.dim)));
    s.hTerms.push_back({{m1, m2}, plus(t, scale(adjoint(t), -1.))});
    iaHF = std::max(iaHF, std::abs(u * (f1.dim * std::sqrt((double)f2.dim))));
  }

  // J sigma_z (x) sigma_z with sigma_z = diag(-1, +1) (|1> excited); H/i = -i J sz sz.
  // The reference declares but never defines qbit::sigmazop (Qbit_.h:68-70): this element is new.
  void zz(int a, int b, const KV& kv) {
    const FreeInfo& fa = freeAt(a); const FreeInfo& fb = freeAt(b);
    if (fa.dim != 2 || fb.dim != 2) throw std::invalid_argument("zz couples two qubits");
    double J = num(kv, "J", 0);
    LocalOp sz = diagonal(2, {-1., 1.});
    s.hTerms.push_back({{a, b}, scale(kron(sz, sz), cd(0., -J))});
    iaHF = std::max(iaHF, std::fabs(J));
  }

  // e^{iKx} on the momentum grid: sum_q |q+K><q| (expINKX, Particle.cc:60-72)
  static LocalOp expINKX(int dim, int K) {
    LocalOp r; r.dims = {dim};
    const int a = std::abs(K);
    for (int q = 0; q + a < dim; ++q) {
      if (K >= 0) r.entries.push_back({q + a, q, 1., 0.});
      else r.entries.push_back({q, q + a, 1., 0.});
    }
    return r;
  }
  // the pump's mode function at wavenumber nK (mfNKX, Particle.cc:274-289)
  static LocalOp modeFunction(int dim, const std::string& mf, int nK) {
    if (mf == "sin") return scale(plus(expINKX(dim, nK), scale(expINKX(dim, -nK), -1.)), 1. / cd(0., 2.));
    if (mf == "cos") return scale(plus(expINKX(dim, nK), expINKX(dim, -nK)), 0.5);
    if (mf == "plus") return expINKX(dim, nK);
    return expINKX(dim, -nK);
  }

  // (Pumped)Particle(Sch) (Particle.cc:79-110, 210-252).  Momentum grid k_n = n - 2^(fin-1)
  // (particle::Spatial with deltaK = 1, Particle.cc:276-290); kinetic term -i omrec k^2 (Sch) or
  // the Exact part exp(-i omrec k^2 t) (IP); classical pump vClass (+-) cos(2 kPart x) / (2i) for
  // a real mode function (hOverI, Particle.cc:84-87); no jumps.
  void particle(const KV& kv) {
    int ax = (int)s.frees.size();
    int fin = (int)num(kv, "fin", 6);
    if (fin < 1 || fin > 12) throw std::invalid_argument("particle finesse out of range");
    const std::string pic = picture(kv);
    if (pic == "UIP") throw std::invalid_argument("the particle has no unitary interaction picture of its own");
    double omrec = num(kv, "omrec", 1.), vclass = num(kv, "vclass", 0.);
    int kpart = (int)num(kv, "kpart", 1);
    auto it = kv.find("mf");
    std::string mf = it == kv.end() ? "sin" : it->second;
    if (mf != "sin" && mf != "cos" && mf != "plus" && mf != "minus") throw std::invalid_argument("unknown particle mode function");
    int dim = 1 << fin;
    FreeInfo f; f.kind = "particle"; f.dim = dim; f.avKind = "particle";
    f.pumped = vclass != 0.; f.mf = mf; f.kpart = kpart;
    LocalOp h; h.dims = {dim};
    if (f.pumped && (mf == "sin" || mf == "cos") && 2 * kpart < dim) {
      LocalOp c = scale(plus(expINKX(dim, 2 * kpart), expINKX(dim, -2 * kpart)), 0.5);
      h = scale(c, cd(vclass * (mf == "sin" ? -1. : 1.), 0.) / cd(0., 2.));
    }
    if (pic == "Sch") {
      std::vector<cd> d(dim);
      for (int n = 0; n < dim; ++n) { double k = n - dim / 2; d[n] = cd(0., -omrec * k * k); }
      h = plus(diagonal(dim, d), h);
      f.hasHa = true;
      f.highestFrequency = std::fabs(omrec) * (double)dim * dim;  // RF{"omrec", omrec, sqr(1<<fin)}
    } else {
      f.hasEx = true;
      for (int n = 0; n < dim; ++n) { double k = n - dim / 2; f.exd.push_back(cd(0., omrec * k * k)); }
      f.hasHa = !h.entries.empty();
      f.highestFrequency = std::fabs(omrec) * (double)dim;        // RF{"omrec", omrec, 1<<fin}
    }
    if (f.pumped) f.highestFrequency = std::max(f.highestFrequency, std::fabs(vclass));  // RF{"vClass", vClass, 1}
    f.vclass = vclass;
    if (f.hasHa) s.hTerms.push_back({{ax}, h});
    freeJumps.emplace_back(); freeObs.emplace_back(); freeVar.emplace_back(); freeSqrt.emplace_back();
    avParticle(ax, dim);
    s.frees.push_back(f);
  }

  // ParticleOrthogonalToCavity (ParticleCavity.cc:47-57, 79-85): the interference term
  // etaeff (a (x) mf(kPart x) + h.c.) / i with etaeff = sign(u0) sqrt(u0 vClass) and the pumped
  // particle's own mode function.
  void pcavOrth(int m, int p, const KV& kv) {
    const FreeInfo& fm = freeAt(m); const FreeInfo& fp = particleAt(p);
    if (fm.kind != "mode") throw std::invalid_argument("pcav_orth needs a mode as its first free");
    if (!fp.pumped) throw std::invalid_argument("pcav_orth needs a pumped particle");
    double u0 = num(kv, "u0", -1.);
    if (u0 * fp.vclass < 0.) throw std::invalid_argument("Unot VClass sign discrepancy");
    const double etaeff = (u0 > 0 ? 1. : u0 < 0 ? -1. : 0.) * std::sqrt(u0 * fp.vclass);
    LocalOp a = kron(destroy(fm.dim), modeFunction(fp.dim, fp.mf, fp.kpart));
    s.hTerms.push_back({{m, p}, scale(plus(a, adjoint(a)), cd(etaeff, 0.) / cd(0., 1.))});
    iaHF = std::max(iaHF, std::max(std::fabs(u0) * fm.dim, std::fabs(etaeff) * std::sqrt((double)fm.dim)));
  }

  // particle::Averaged (Particle.cc:138-167): <P>, VAR(P), <X>, DEV(X).  The position
  // populations are those of the k -> x transform of quantumdata::ffTransform
  // (LazyDensityOperatorFFT.cc:16-40: GSL backward DFT, halves swapped, N^-1/2), so
  // <f(X)> = <psi| F^dagger f(x) F |psi> with F[j][n] = N^-1/2 exp(2 pi i (j + N/2) n / N):
  // a dense (circulant) operator in the momentum basis.
  void avParticle(int ax, int dim) {
    std::vector<cd> k(dim), k2(dim);
    for (int n = 0; n < dim; ++n) { double kk = n - dim / 2; k[n] = kk; k2[n] = kk * kk; }
    const double pi = 3.14159265358979323846, dx = 2. * pi / dim, xMax = pi;
    LocalOp X, X2; X.dims = {dim}; X2.dims = {dim};
    for (int n = 0; n < dim; ++n)
      for (int m = 0; m < dim; ++m) {
        cd a = 0., b = 0.;
        for (int j = 0; j < dim; ++j) {
          const double x = j * dx - xMax;
          const long ph = ((long)(j + dim / 2) * (m - n)) % dim;
          const cd e = std::polar(1.0 / dim, 2. * pi * (double)ph / dim);
          a += x * e; b += x * x * e;
        }
        X.entries.push_back({n, m, a, 0.});
        X2.entries.push_back({n, m, b, 0.});
      }
    Observable op, op2, ox, ox2;
    op.terms.push_back({{ax}, diagonal(dim, k)});
    op2.terms.push_back({{ax}, diagonal(dim, k2)});
    ox.terms.push_back({{ax}, X});
    ox2.terms.push_back({{ax}, X2});
    freeObs.back() = {op, op2, ox, ox2};
    freeVar.back() = {-1, 0, -1, 2};       // VAR(P) = <P^2> - <P>^2, DEV(X) = sqrt(<X^2> - <X>^2)
    freeSqrt.back() = {0, 0, 0, 1};
  }

  const FreeInfo& particleAt(int i) const {
    const FreeInfo& f = freeAt(i);
    if (f.kind != "particle") throw std::invalid_argument("pcav needs a particle as its second free");
    return f;
  }

  // ParticleAlongCavity (ParticleCavity.cc:26-53, 87-110).  The dispersive term u0 n (x) (+-)
  // cos(2 kcav x) / (2i) for a real cavity mode function; cos(K x) = (e^{iKx} + e^{-iKx}) / 2 shifts
  // the momentum index by +-K (Particle.cc:60-80); the constant u0 n / 2 of cos^2 belongs to the
  // mode detuning.  The interference term with etaeff = sign(u0) sqrt(u0 v): for a particle
  // without its own pump, or with vclass= given (v = vclass), etaeff (a (x) mf(kcav x) + h.c.) / i;
  // for a pumped particle (v = its vClass) and kcav = kPart, etaeff (a^dagger (x)
  // mfComposition(mf, mfPart) + h.c.) / i.  A pumped particle of another wavenumber takes the
  // chained product of ParticleCavity.cc:102-140: secondH (firstH psi) + secondHT (firstHT psi) with
  // firstH = etaeff a (x) mf(kcav x) / i, firstHT = -firstH^dagger, secondH = mfPart(kPart x)^dagger
  // on the particle, i.e. etaeff (a (x) mfPart^dagger mf + a^dagger (x) mfPart mf^dagger) / i, each
  // product formed on the (truncated) momentum grid.  Its entries shift the particle by
  // +-kPart +- kcav, never 0, so in the interaction picture each rotates as one entry when both
  // frees are Exact; with only the particle Exact the reference rotates the second factor alone
  // (per intermediate momentum), which one ELL entry cannot hold: refused.
  void pcav(int m, int p, const KV& kv) {
    const FreeInfo& fm = freeAt(m); const FreeInfo& fp = particleAt(p);
    if (fm.kind != "mode") throw std::invalid_argument("pcav needs a mode as its first free");
    double u0 = num(kv, "u0", 0);
    int kcav = (int)num(kv, "kcav", 1);
    auto mf = kv.find("mf");
    std::string mfs = mf == kv.end() ? "cos" : mf->second;
    if (mfs != "cos" && mfs != "sin" && mfs != "plus" && mfs != "minus")
      throw std::invalid_argument("unknown pcav mode function");
    const bool pumpedCtor = fp.pumped && !kv.count("vclass");
    const double v = pumpedCtor ? fp.vclass : num(kv, "vclass", 0.);
    if (u0 * v < 0.) throw std::invalid_argument("Unot VClass sign discrepancy");
    const double etaeff = (u0 > 0 ? 1. : u0 < 0 ? -1. : 0.) * std::sqrt(u0 * v);
    const bool special = !pumpedCtor || kcav == std::abs(fp.kpart);
    if (!special && etaeff != 0. && fp.hasEx && !fm.hasEx)
      throw std::invalid_argument("pcav with a pumped particle of another wavenumber needs the mode in the interaction "
                                  "picture when the particle is");
    const double sgn = mfs == "sin" ? -1. : 1.;
    const int K = 2 * kcav;
    LocalOp t; t.dims = {fm.dim, fp.dim};
    // a complex mode function (plus / minus) has no dispersive term (ParticleCavity.cc:27)
    if (u0 != 0. && K < fp.dim && (mfs == "cos" || mfs == "sin")) {
      LocalOp n; n.dims = {fm.dim};
      for (int a = 0; a < fm.dim; ++a) n.entries.push_back({a, a, (double)a, 0.});
      LocalOp c; c.dims = {fp.dim};
      for (int q = 0; q + K < fp.dim; ++q) {
        c.entries.push_back({q + K, q, 0.5, 0.});
        c.entries.push_back({q, q + K, 0.5, 0.});
      }
      t = scale(kron(n, c), cd(u0 * sgn, 0.) / cd(0., 2.));
    }
    if (etaeff != 0. && special) {
      LocalOp x = pumpedCtor ? kron(adjoint(destroy(fm.dim)), mfComposition(fp.dim, mfs, fp.mf, kcav))
                             : kron(destroy(fm.dim), modeFunction(fp.dim, mfs, kcav));
      t = plus(t, scale(plus(x, adjoint(x)), cd(etaeff, 0.) / cd(0., 1.)));
    }
    if (etaeff != 0. && !special) {
      const LocalOp mc = modeFunction(fp.dim, mfs, kcav), mp = modeFunction(fp.dim, fp.mf, fp.kpart);
      const LocalOp x = plus(kron(destroy(fm.dim), matmul(adjoint(mp), mc)),
                             kron(adjoint(destroy(fm.dim)), matmul(mp, adjoint(mc))));
      t = plus(t, scale(x, cd(etaeff, 0.) / cd(0., 1.)));
    }
    s.hTerms.push_back({{m, p}, t});
    // RF{"Unot", uNot, mode dim}, RF{"etaeff", etaeff, sqrt(mode dim)}
    iaHF = std::max(iaHF, std::max(std::fabs(u0) * fm.dim, std::fabs(etaeff) * std::sqrt((double)fm.dim)));
  }

  // mfComposition(particle, (mf0, K), (mf1, K)), Particle.cc:290-345, for positive wavenumbers
  static LocalOp mfComposition(int dim, const std::string& a, const std::string& b, int nK) {
    LocalOp id; id.dims = {dim};
    for (int q = 0; q < dim; ++q) id.entries.push_back({q, q, 1., 0.});
    auto cos2 = [&]() { return scale(plus(expINKX(dim, 2 * nK), expINKX(dim, -2 * nK)), 0.5); };
    auto sin2 = [&]() { return scale(plus(expINKX(dim, 2 * nK), scale(expINKX(dim, -2 * nK), -1.)), 1. / cd(0., 2.)); };
    const cd half_i = cd(0., 0.5);
    if (a == "sin" && b == "sin") return scale(plus(id, scale(cos2(), -1.)), 0.5);
    if (a == "cos" && b == "cos") return scale(plus(id, cos2()), 0.5);
    if ((a == "sin" && b == "cos") || (a == "cos" && b == "sin")) return scale(sin2(), 0.5);
    if (a == b) return id;  // plus-plus, minus-minus
    if (a == "plus" && b == "minus") return expINKX(dim, -2 * nK);
    if (a == "minus" && b == "plus") return expINKX(dim, 2 * nK);
    if ((a == "plus" && b == "sin") || (a == "sin" && b == "minus")) return scale(plus(expINKX(dim, -2 * nK), scale(id, -1.)), half_i);
    if ((a == "plus" && b == "cos") || (a == "cos" && b == "minus")) return scale(plus(id, expINKX(dim, -2 * nK)), 0.5);
    if ((a == "minus" && b == "cos") || (a == "cos" && b == "plus")) return scale(plus(id, expINKX(dim, 2 * nK)), 0.5);
    return scale(plus(expINKX(dim, 2 * nK), scale(id, -1.)), -half_i);  // minus-sin, sin-plus
  }

  // ParticleTwoModes with equal wavenumbers (ParticleTwoModes.cc:35-46, the special Hamiltonian):
  // tridiagPlusHC_overI(f a0^dagger (x) a1 (x) mfComposition(mf0, mf1)), f = sign(u0) sqrt(u0 u1) e^{i phi}
  void particleTwoModes(int m0, int m1, int p, const KV& kv) {
    const FreeInfo& f0 = freeAt(m0); const FreeInfo& f1 = freeAt(m1); const FreeInfo& fp = particleAt(p);
    if (f0.kind != "mode" || f1.kind != "mode") throw std::invalid_argument("ptm couples two modes and a particle");
    double u0 = num(kv, "u0", -1.), u1 = num(kv, "u1", -1.), phi = num(kv, "phi", 0.);
    if (u0 * u1 < 0) throw std::invalid_argument("Unots sign discrepancy");
    int k0 = (int)num(kv, "k0", 1), k1 = (int)num(kv, "k1", 1);
    if (k0 != k1) throw std::invalid_argument("ptm with unequal wavenumbers is not built");
    auto g = [&](const char* key) { auto it = kv.find(key); return it == kv.end() ? std::string("plus") : it->second; };
    const std::string mf0 = g("mf0"), mf1 = g("mf1");
    const cd fac = (u0 > 0 ? 1. : u0 < 0 ? -1. : 0.) * std::sqrt(u0 * u1) * std::polar(1., phi);
    LocalOp x = scale(kron(kron(adjoint(destroy(f0.dim)), destroy(f1.dim)), mfComposition(fp.dim, mf0, mf1, k0)), fac);
    s.hTerms.push_back({{m0, m1, p}, scale(plus(x, adjoint(x)), 1. / cd(0., 1.))});
    iaHF = std::max(iaHF, std::max(std::fabs(u0) * std::sqrt((double)f0.dim), std::fabs(u1) * std::sqrt((double)f1.dim)));
  }

  std::vector<std::vector<char>> freeSqrt;
  double iaHF = 0.;
};

System* buildFromSpec(const std::string& spec) {
  std::unique_ptr<System> sp(new System);
  System& s = *sp;
  Builder b(s);
  std::vector<std::string> lines;
  {
    std::string cur;
    for (char c : spec) { if (c == ';' || c == '\n') { lines.push_back(cur); cur.clear(); } else cur += c; }
    lines.push_back(cur);
  }
  int forced = -1;
  int nIa = 0;
  for (auto& ln : lines) {
    std::istringstream is(ln);
    std::string what, kind;
    if (!(is >> what)) continue;
    if (!(is >> kind)) throw std::invalid_argument("incomplete statement: " + ln);
    std::vector<int> axes;
    KV kv;
    std::string tok;
    while (is >> tok) {
      auto p = tok.find('=');
      if (p == std::string::npos) axes.push_back(std::stoi(tok));
      else kv[tok.substr(0, p)] = tok.substr(p + 1);
    }
    if (what == "free") {
      if (!axes.empty()) throw std::invalid_argument("free element takes no axes");
      if (kind == "mode") b.mode(kv);
      else if (kind == "qbit") b.qbit(kv);
      else if (kind == "ptla") b.ptla(kv);
      else if (kind == "particle") b.particle(kv);
      else throw std::invalid_argument("unknown free element " + kind);
    } else if (what == "ia") {
      const size_t want = kind == "ptm" ? 3 : 2;
      if (axes.size() != want) throw std::invalid_argument(kind == "ptm" ? "ptm needs three axes" : "interaction needs two axes");
      for (size_t i = 0; i < axes.size(); ++i)
        for (size_t j = i + 1; j < axes.size(); ++j)
          if (axes[i] == axes[j]) throw std::invalid_argument("interaction axes must differ");
      if (kind == "jc") b.jc(axes[0], axes[1], kv);
      else if (kind == "bs") b.bs(axes[0], axes[1], kv);
      else if (kind == "zz") b.zz(axes[0], axes[1], kv);
      else if (kind == "pcav") b.pcav(axes[0], axes[1], kv);
      else if (kind == "pcav_orth") b.pcavOrth(axes[0], axes[1], kv);
      else if (kind == "ptm") b.particleTwoModes(axes[0], axes[1], axes[2], kv);
      else throw std::invalid_argument("unknown interaction " + kind);
      ++nIa;
    } else if (what == "system") {
      if (kind == "binary") forced = 1; else if (kind == "composite") forced = 0;
      else throw std::invalid_argument("unknown system kind " + kind);
    } else throw std::invalid_argument("unknown statement " + what);
  }
  if (s.frees.empty()) throw std::invalid_argument("system without free elements");
  s.isBinary = forced >= 0 ? forced == 1 : (s.frees.size() == 2 && nIa <= 1);
  for (auto& f : s.frees) s.dims.push_back(f.dim);
  s.strides.assign(s.dims.size(), 1);
  s.dim = 1;
  for (int i = (int)s.dims.size() - 1; i >= 0; --i) { s.strides[i] = s.dim; s.dim *= s.dims[i]; }
  // jump channels: frees first (BinarySystem.cc:217-239)
  for (auto& js : b.freeJumps) for (auto& t : js) s.jumps.push_back({t});
  for (size_t i = 0; i < b.freeObs.size(); ++i) {
    int base = (int)s.obs.size();
    for (size_t k = 0; k < b.freeObs[i].size(); ++k) {
      s.obs.push_back(b.freeObs[i][k]);
      s.processVarOf.push_back(b.freeVar[i][k] < 0 ? -1 : base + b.freeVar[i][k]);
      s.processSqrt.push_back(b.freeSqrt[i][k]);
    }
  }
  // highest frequency (DynamicsBase.cc:49-66); binary systems include the interaction's (Q11)
  for (auto& f : s.frees) s.highestFrequency = std::max(s.highestFrequency, f.highestFrequency);
  if (s.isBinary) s.highestFrequency = std::max(s.highestFrequency, b.iaHF);
  for (auto& f : s.frees) s.hasEx = s.hasEx || f.hasEx;
  if (s.hasEx) {
    s.exactZ.assign(s.dim, 0.);
    for (long r = 0; r < s.dim; ++r)
      for (size_t a = 0; a < s.frees.size(); ++a)
        if (s.frees[a].hasEx) s.exactZ[r] += s.frees[a].exd[(r / s.strides[a]) % s.dims[a]];
  }
  s.lower();
  return sp.release();
}

// ---------------------------------------------------------------------------
// embedding: local operators on axes V -> row-gather ELL over the full space
// ---------------------------------------------------------------------------

Ell embed(const System& s, const std::vector<Term>& terms) {
  struct E { int col; cd val; cd freq; };
  std::vector<std::vector<E>> rows(s.dim);
  for (auto& t : terms) {
    const LocalOp& op = t.op;
    int r = (int)t.axes.size();
    if ((int)op.dims.size() != r) throw std::invalid_argument("operator rank does not match its axes");
    for (int i = 0; i < r; ++i)
      if (t.axes[i] < 0 || t.axes[i] >= (int)s.dims.size() || op.dims[i] != s.dims[t.axes[i]])
        throw std::invalid_argument("operator axes do not match the system dimensions");
    std::vector<int> lstr(r, 1);
    for (int i = r - 2; i >= 0; --i) lstr[i] = lstr[i + 1] * op.dims[i + 1];
    // The reference's direct product of Tridiagonals carries frequencies only where every factor
    // does (doDirect of an empty array is empty: Tridiagonal.tcc:61-67, ComplexArrayExtensions.h:80-92);
    // a factor has them on its off-diagonals only, and only when its free is Exact
    // (Tridiagonal.cc:24-36 leaves the main diagonal's empty).  So an entry of a term rotates with
    // the interaction picture only if all of the term's frees are Exact and the entry is off the
    // main diagonal on every one of its axes (a single-axis term needs only the former).
    bool allEx = !t.explicitFreq;
    for (int i = 0; i < r; ++i) allEx = allEx && s.frees[t.axes[i]].hasEx;
    // group entries by local row
    std::vector<std::vector<const LocalOp::Entry*>> byRow(op.size());
    for (auto& e : op.entries) byRow[e.row].push_back(&e);
    for (long row = 0; row < s.dim; ++row) {
      int lr = 0;
      for (int i = 0; i < r; ++i) lr += (int)((row / s.strides[t.axes[i]]) % s.dims[t.axes[i]]) * lstr[i];
      for (const LocalOp::Entry* e : byRow[lr]) {
        long col = row;
        cd freq = 0.;
        bool offEverywhere = true;
        for (int i = 0; i < r; ++i) {
          int ax = t.axes[i];
          int rc = (e->row / lstr[i]) % op.dims[i], cc = (e->col / lstr[i]) % op.dims[i];
          col += (long)(cc - rc) * s.strides[ax];
          offEverywhere = offEverywhere && rc != cc;
          // interaction-picture phase of |r><c| under U = exp(-exd[n] t): exp((exd[r] - exd[c]) (t - t0))
          if (allEx) freq += s.frees[ax].exd[rc] - s.frees[ax].exd[cc];
        }
        if (r > 1 && !offEverywhere) freq = 0.;
        if (t.explicitFreq) freq = e->freq;
        auto& rv = rows[row];
        auto it = std::find_if(rv.begin(), rv.end(), [&](const E& x) { return x.col == (int)col && x.freq == freq; });
        if (it == rv.end()) rv.push_back({(int)col, e->val, freq});
        else it->val += e->val;
      }
    }
  }
  Ell ell;
  size_t w = 1;
  for (auto& rv : rows) w = std::max(w, rv.size());
  ell.width = (int)w;
  ell.cols.assign(s.dim * w, 0);
  ell.vals.assign(s.dim * w, 0.);
  bool td = false;
  for (auto& rv : rows) for (auto& e : rv) if (e.freq != cd(0.)) td = true;
  if (td) ell.freqs.assign(s.dim * w, 0.);
  for (long row = 0; row < s.dim; ++row) {
    auto& rv = rows[row];
    // keep the diagonal first (stable otherwise): a fixed order across rows helps the gather
    std::stable_sort(rv.begin(), rv.end(), [&](const E& a, const E& b) { return (a.col == row) > (b.col == row); });
    for (size_t k = 0; k < w; ++k) {
      if (k < rv.size()) {
        ell.cols[row * w + k] = rv[k].col; ell.vals[row * w + k] = rv[k].val;
        if (td) ell.freqs[row * w + k] = rv[k].freq;
      } else {
        ell.cols[row * w + k] = (int)row;  // padding: zero weight on the row itself
      }
    }
  }
  return ell;
}

void System::lower() {
  H = embed(*this, hTerms);
  J.clear();
  for (auto& j : jumps) J.push_back(embed(*this, j));
  O.clear();
  for (auto& o : obs) O.push_back(embed(*this, o.terms));
}

}  // namespace qt
