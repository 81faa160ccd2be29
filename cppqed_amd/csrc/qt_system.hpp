// qt_system.hpp -- host-side system compiler.
//
// Builds a quantum system from elements (the reference's Free / Interaction /
// BinarySystem / Composite element API, CPPQEDcore/structure and
// CPPQEDcore/composites) and lowers it to the flat, device-ready form the HIP
// kernels consume:
//
//   * H_nH/i as an ELL matrix over the full Hilbert space (row-gather form:
//     every row owns up to W (column, value, frequency) entries; the rows
//     are padded with zero entries pointing at themselves);
//   * one ELL matrix per jump operator J_m (channels ordered frees first,
//     then interactions, as BinarySystem.cc:141-156 / Composite.tcc:238-303);
//   * one ELL matrix per linear observable (the element Averaged classes),
//     plus the post-processing rule (Mode.cc:198-201);
//   * optional per-basis-state exponent sum for the Exact (interaction-picture)
//     diagonal U = exp(-z (t - t0)).
//
// Operators are held as sparse COO matrices on an ordered set of axes (the V of
// the reference's Act<...>/slice iteration) and embedded into the full space by
// stride arithmetic, which replaces the reference's run-time slice iteration
// (CPPQEDutils/SliceIterator.h:277-354) by tables precomputed once per system.
#pragma once

#include <complex>
#include <map>
#include <string>
#include <vector>

namespace qt {

using cd = std::complex<double>;

// A sparse operator on the tensor product of `dims` (local row-major indexing).
struct LocalOp {
  struct Entry { int row, col; cd val; cd freq; };
  std::vector<int> dims;
  std::vector<Entry> entries;
  int size() const { int n = 1; for (int d : dims) n *= d; return n; }
};

LocalOp kron(const LocalOp& a, const LocalOp& b);          // freqs add (direct sum)
LocalOp adjoint(const LocalOp& a);                         // conjugate transpose, freqs negated
LocalOp scale(const LocalOp& a, cd f);
LocalOp plus(const LocalOp& a, const LocalOp& b);          // concatenation (merged on embedding)

// op acting on the listed axes (V order).  The interaction-picture frequency of an entry comes
// from the Exact frees' exponents (the catalogue's elements), or, with explicitFreq, from the
// entry itself (terms handed over as Tridiagonal data, whose frequencies they carry).
struct Term { std::vector<int> axes; LocalOp op; bool explicitFreq = false; };

struct Observable {
  std::vector<Term> terms;  // summed
  bool imagPart = false;    // stream Im<O> instead of Re<O>
};

struct FreeInfo {
  std::string kind;
  int dim = 0;
  bool hasHa = false, hasEx = false;
  std::vector<cd> exd;       // Exact part: U = exp(-exd[n] (t - t0)) on level n (empty without one)
  double highestFrequency = 0.;
  int nJumps = 0;
  std::string avKind;  // "mode" | "qbit" | "particle" | ""
  // pumped particle: classical pump strength, mode function and wavenumber
  bool pumped = false;
  double vclass = 0.;
  std::string mf;
  int kpart = 0;
};

// ELL matrix over the full space
struct Ell {
  int width = 0;
  std::vector<int> cols;    // dim * width, row-major by row
  std::vector<cd> vals;
  std::vector<cd> freqs;    // empty unless some entry carries a frequency
  bool timeDependent() const { return !freqs.empty(); }
};

struct System {
  std::vector<int> dims;
  std::vector<long> strides;
  long dim = 1;
  std::vector<FreeInfo> frees;
  std::vector<Term> hTerms;               // H_nH/i, in application order
  std::vector<std::vector<Term>> jumps;   // per channel
  std::vector<Observable> obs;            // raw linear averages, 4 per free with averages
  std::vector<int> processVarOf;          // for each obs: index of <n> to subtract squared, or -1
  std::vector<char> processSqrt;          // for each obs: take the square root after the subtraction
  double highestFrequency = 0.;
  bool isBinary = false;
  bool hasEx = false;
  std::vector<cd> exactZ;                 // per basis state, sum over frees of exd[n_free] (empty if no Exact)

  // lowered tables
  Ell H;
  std::vector<Ell> J;
  std::vector<Ell> O;
  void lower();
};

// Parse the element spec (one statement per line or ';'):
//   free mode  cutoff=N delta=.. kappa=.. eta=.. [etaI=..] [picture=Sch|IP|UIP]
//   free qbit  deltaA=.. gamma=.. eta=.. [etaI=..] [picture=..]
//   free ptla  gamma=.. deltaA=.. eta=..                   (PumpedTwoLevelAtomSch)
//   free particle fin=.. omrec=.. [vclass=.. kpart=.. mf=sin|cos|plus|minus] [picture=..]
//                                                          ((Pumped)Particle(Sch), momentum basis)
//   ia jc i j g=.. [gI=..]                                 (JaynesCummings)
//   ia bs i j g=..                                         (CoupledModes, RWA beam splitter)
//   ia zz i j J=..                                         (qubit sigma_z sigma_z coupling)
//   ia pcav m p u0=.. kcav=.. mf=.. [vclass=..]            (ParticleAlongCavity: dispersive and
//                                                          interference terms)
//   ia pcav_orth m p u0=..                                 (ParticleOrthogonalToCavity, pumped particle)
//   ia ptm m0 m1 p u0=.. u1=.. phi=.. k0=.. k1=.. mf0=.. mf1=..  (ParticleTwoModes, equal wavenumbers)
//   system binary|composite
System* buildFromSpec(const std::string& spec);

Ell embed(const System& s, const std::vector<Term>& terms);

}  // namespace qt
