"""The C++ of the integration guide compiles and links against the library: the device-backed
ensemble a C++QED maintainer would add (a Trajectory-concept class over qt_mcwf_*), and the
adapter from quantumoperator::Tridiagonal to qt_tridiag, each built with g++ against small
stand-ins for the reference's headers (the reference itself cannot be built here).  The adapter
also runs (qt_sys_build_terms is host code): a driven lossy mode handed over as Tridiagonals
lowers to its textbook tables.  CPU only: the ensemble class is compiled and linked, not run."""
import os
import re
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

MCWF_STANDIN = r"""
#pragma once
#include <complex>
#include <ostream>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>
struct pcg64 {};
namespace structure {
struct Averages {
  std::vector<double> v;
  explicit Averages(int n) : v(n) {}
  double& operator()(int i) { return v[i]; }
};
}
namespace quantumdata {
template<int RANK> struct StateVector {
  std::vector<std::complex<double>> a;
  const std::vector<std::complex<double>>& getArray() const { return a; }
};
}
namespace mcwf {
template<typename RNG> struct Pars { double epsRel, epsAbs, dpLimit, overshootTolerance; unsigned long long seed, prngStream; };
}
struct FormDouble {
  int p;
  explicit FormDouble(int pr) : p(pr) {}
  double operator()(double x) const { return x; }
};
namespace formdouble { inline int positive(int p) { return p; } }
"""

TRIDIAGONAL_STANDIN = r"""
#pragma once
#include <array>
#include <complex>
#include <cstddef>
#include <vector>
namespace quantumoperator {
constexpr int pow3(int r) { return r ? 3 * pow3(r - 1) : 1; }
struct Diagonal {
  std::vector<std::complex<double>> v;
  size_t size() const { return v.size(); }
  const std::complex<double>* data() const { return v.data(); }
};
template<int RANK> struct Tridiagonal {
  static const int LENGTH = pow3(RANK);
  struct Diagonals {
    std::vector<Diagonal> d = std::vector<Diagonal>(LENGTH);
    const Diagonal& operator()(int i) const { return d[i]; }
  };
  Diagonals diagonals, freqs;
  std::array<size_t, RANK> differences{};
  const Diagonals& get() const { return diagonals; }
  const Diagonals& getFreqs() const { return freqs; }
  const std::array<size_t, RANK>& getDifferences() const { return differences; }
};
}
"""

DRIVER = r"""
#include <cmath>
#include <cstdio>
#include <iostream>

// what cppqedutils::run asks of a trajectory (Trajectory.h:310-339)
template<class Traj> void runLike(Traj& tr, double T, double Dt, std::ostream& os) {
  while (tr.getTime() < T) {
    tr.advance(std::min(Dt, T - tr.getTime()), os);
    tr.stream(os, 6);
    os << tr.getDtDid() << '\n';
  }
}

int main(int argc, char** argv) {
  if (argc > 99) {  // linked, never run here (needs a GPU)
    quantumdata::StateVector<2> psi{std::vector<std::complex<double>>(4, 0.5)};
    mcwf::Pars<pcg64> p{1e-6, 1e-12, 0.01, 10., 1001, 1};
    quantumtrajectory::DeviceEnsembleMCWF<2> e(argv[1], psi, p, 16);
    runLike(e, 1.0, 0.1, std::cout);
  }
  // a driven lossy mode, H/i = -z n + eta a^dagger - eta^* a, as one Tridiagonal, through DeviceTerms
  const int N = 5;
  const std::complex<double> z(0.5, 1.0), eta(0.7, 0.2);
  quantumoperator::Tridiagonal<1> h;
  h.differences = {1};
  for (int n = 0; n < N; ++n) h.diagonals.d[0].v.push_back(-z * double(n));
  for (int n = 0; n + 1 < N; ++n) {
    h.diagonals.d[1].v.push_back(eta * std::sqrt(n + 1.));             // |n+1><n|
    h.diagonals.d[2].v.push_back(-std::conj(eta) * std::sqrt(n + 1.));  // |n><n+1|
  }
  quantumtrajectory::DeviceTerms terms;
  terms.add(h, {0});
  quantumoperator::Tridiagonal<1> jump;
  jump.differences = {1};
  for (int n = 0; n + 1 < N; ++n) jump.diagonals.d[2].v.push_back(std::sqrt(2. * z.real()) * std::sqrt(n + 1.));
  quantumtrajectory::DeviceTerms jterms;
  jterms.add(jump, {0});
  const qt_free_desc f{N, nullptr};
  const qt_operator j = jterms.op();
  qt_system_desc d{1, &f, terms.op(), 1, &j, 0, nullptr, nullptr, nullptr, nullptr, 5.0};
  qt_sys* sys = nullptr;
  if (qt_sys_build_terms(&d, &sys)) { std::printf("error %s\n", qt_last_error()); return 1; }
  int w = 0;
  qt_sys_table(sys, 0, &w, nullptr, nullptr, nullptr);
  std::vector<int> cols(N * w);
  std::vector<double> vals(2 * N * w);
  qt_sys_table(sys, 0, nullptr, cols.data(), vals.data(), nullptr);
  for (int r = 0; r < N; ++r)
    for (int k = 0; k < w; ++k) std::printf("%d %d %.17g %.17g\n", r, cols[r * w + k], vals[2 * (r * w + k)], vals[2 * (r * w + k) + 1]);
  qt_sys_destroy(sys);
  return 0;
}
"""


def _blocks():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    return re.findall(r"```cpp\n(.*?)```", text, re.S)


def test_integration_cpp_compiles_links_and_lowers(tmp_path):
    blocks = _blocks()
    ens = next(b for b in blocks if "class DeviceEnsembleMCWF" in b)
    terms = next(b for b in blocks if "struct DeviceTerms" in b)
    (tmp_path / "MCWF_Trajectory.h").write_text(MCWF_STANDIN)
    (tmp_path / "Tridiagonal.h").write_text(TRIDIAGONAL_STANDIN)
    (tmp_path / "binding.cpp").write_text(ens + "\n" + terms + "\n" + DRIVER)
    lib = os.path.join(ROOT, "cppqed_amd")
    exe = tmp_path / "binding"
    cmd = ["g++", "-std=c++17", "-Wall", "-Werror", "-I", str(tmp_path), "-I", os.path.join(ROOT, "include"),
           str(tmp_path / "binding.cpp"), "-o", str(exe), "-L", lib, "-lqtraj", "-Wl,-rpath," + lib,
           "-Wl,--allow-shlib-undefined"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    N, z, eta = 5, 0.5 + 1.0j, 0.7 + 0.2j
    H = np.zeros((N, N), complex)
    for ln in out.stdout.split("\n"):
        if ln.strip():
            r_, c_, re_, im_ = ln.split()
            H[int(r_), int(c_)] += float(re_) + 1j * float(im_)
    a = np.diag(np.sqrt(np.arange(1, N)), 1)
    want = -z * np.diag(np.arange(N)) + eta * a.T - np.conj(eta) * a
    np.testing.assert_allclose(H, want, rtol=0, atol=1e-15)
