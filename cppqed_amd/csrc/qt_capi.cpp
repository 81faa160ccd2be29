// qt_capi.cpp -- the C ABI (include/qtraj.h): system handles, device-resident MCWF
// ensembles, the plain one-trajectory form.
#include "../../include/qtraj.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <deque>
#include <map>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "qt_kernels.hpp"
#include "qt_philox.h"
#include "qt_system.hpp"

namespace qt {
bool layout_supported(int rows, int nw, int wh);
bool batch_supported(int rows, int nw, int wh, int minw, bool td, bool gsl);
int batch_blocks_per_cu(int rows, int nw, int wh, int minw, bool f32, bool td, bool gsl);
bool sub_supported(int g, int rows, int wh);
int sub_blocks_per_cu(int g, int rows, int wh, bool f32);
hipError_t launch_sub(int g, int rows, int wh, bool f32, int grid, const SysDev& S, const EnsDev& E, hipStream_t st);
hipError_t launch_batch(int rows, int nw, int wh, int minw, bool f32, bool td, bool gsl, int grid, const SysDev& S,
                        const EnsDev& E, hipStream_t st);
hipError_t launch_jump_draws(const unsigned long long* ordinal, unsigned long long seed, unsigned long long traj0,
                             int n_traj, int cap, double* draws, hipStream_t st, unsigned long long* base);
bool mixed_supported(int nwb, int rb, int whb, int rw, int whw);
int mixed_blocks_per_cu(int nwb, int rb, int whb, int rw, int whw, int ph, bool f32, bool td, bool gsl);
int mixed_variants(int* table, int cap);
hipError_t launch_mixed(int nwb, int rb, int whb, int rw, int whw, int ph, bool f32, bool td, bool gsl, int grid,
                        const MixDev* M, hipStream_t st);
hipError_t launch_fill_int(int* p, int n, int v, hipStream_t st);
hipError_t launch_ensemble_rho(const double2* psi, int n, int d, int splits, double2* part, double2* rho,
                               hipStream_t st);
hipError_t launch_zero_u64(unsigned long long* p, size_t n, hipStream_t st);
hipError_t launch_stream_rows(const unsigned long long* acc, int n_int, int no, int n_traj, double t_start,
                              const double* intervals, const int* pvar, const int* psqrt, double* rows, hipStream_t st);
hipError_t launch_plain_try(int rows, int nw, int wh, bool td, bool gsl, const SysDev& S, const PlainTry& P,
                            hipStream_t st);
hipError_t launch_plain_post(int rows, int nw, int wh, const SysDev& S, double2* psi, double* out, double dtau,
                             hipStream_t st);
hipError_t launch_plain_jump(int rows, int nw, int wh, const SysDev& S, double2* psi, int m, double nf, hipStream_t st);
hipError_t launch_plain_obs(int rows, int nw, int wh, const SysDev& S, const double2* psi, double* out, hipStream_t st);
hipError_t launch_apply_heff(int rows, int nw, int wh, bool f32, bool td, const SysDev& S, const double2* psi,
                             double2* dpsidt, int n, double tau, hipStream_t st);
struct MasterDev;
bool master_supported(int nb);
hipError_t launch_master(int nb, const MasterDev& M, hipStream_t st);
bool master_sparse_supported(int d);
hipError_t launch_master_sparse(const MasterDev& M, hipStream_t st);
}  // namespace qt

#include "qt_master_dev.hpp"

namespace {

thread_local std::string g_err;

struct Fail : std::runtime_error {
  int code;
  Fail(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void hipCheck(hipError_t e, const char* what) {
  if (e != hipSuccess) throw Fail(QT_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

template <typename F>
int guard(F&& f) {
  try {
    f();
    return QT_OK;
  } catch (Fail& e) {
    g_err = e.what();
    return e.code;
  } catch (std::bad_alloc&) {
    g_err = "out of host memory";
    return QT_ENOMEM;
  } catch (std::exception& e) {
    g_err = e.what();
    return QT_EINVAL;
  }
}

template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { if (p) (void)hipFree(p); }
  void alloc(size_t count) {
    if (count <= n && p) return;
    if (p) { (void)hipFree(p); p = nullptr; n = 0; }
    if (!count) return;
    hipError_t e = hipMalloc(&p, count * sizeof(T));
    if (e != hipSuccess) { p = nullptr; throw Fail(QT_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e)); }
    n = count;
  }
  void upload(const std::vector<T>& v) {
    alloc(v.size());
    if (!v.empty()) hipCheck(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy H2D");
  }
};

struct LayoutChoice { int rows, nw, wh, occ = 1, glanes = 0; };  // glanes > 0: sub-wave groups

// Layout table (rows per lane x wavefronts per group) by dimension; the ELL width is
// rounded up to a compiled width.
LayoutChoice pick_layout(long dim, int width, int rows_req, int nw_req) {
  int wh = width <= 4 ? 4 : width <= 8 ? 8 : width <= 16 ? 16 : -1;
  if (wh < 0) throw Fail(QT_ENOTSUP, "H_nH/i rows wider than 16 entries are not supported");
  auto fits = [&](int r, int n) { return (long)r * 64 * n >= dim; };
  if (rows_req > 0 && nw_req > 0) {
    if (!fits(rows_req, nw_req)) throw Fail(QT_EINVAL, "requested layout does not cover the dimension");
    for (int w : {wh, 8, 16}) if (w >= wh && qt::layout_supported(rows_req, nw_req, w)) return {rows_req, nw_req, w};
    throw Fail(QT_ENOTSUP, "requested layout is not compiled in");
  }
  static const int cand[][2] = {{1, 1}, {2, 1}, {4, 1}, {1, 4}, {2, 4}, {4, 4}, {4, 8}, {2, 16}, {4, 16}};
  for (auto& c : cand) {
    if (!fits(c[0], c[1])) continue;
    for (int w : {wh, 8, 16}) if (w >= wh && qt::layout_supported(c[0], c[1], w)) return {c[0], c[1], w};
  }
  throw Fail(QT_ENOTSUP, "dimension too large for the compiled layouts");
}

// sub-wave group width by dimension (measured on MI355X)
int default_group_lanes(long dim) { return dim <= 64 ? 16 : 64; }

// occupancy hint per layout (measured on MI355X; 1 = no hint)
int default_occupancy(const LayoutChoice& L) {
  // one trajectory per wavefront at dim <= 64: asking for 4 waves/SIMD schedules the
  // kernel for latency hiding (JC-60: 453 vs 413 M tries/s FP32-deriv, 398 vs 387 FP64)
  if (L.rows == 1 && L.nw == 1 && L.wh == 4) return 4;
  return 1;
}

// sub-wave group layout: G lanes per trajectory, rows per lane covering the dimension
LayoutChoice pick_sub(long dim, int width, int g_req, int rows_req) {
  int wh = width <= 4 ? 4 : width <= 8 ? 8 : width <= 16 ? 16 : -1;
  if (wh < 0) throw Fail(QT_ENOTSUP, "H_nH/i rows wider than 16 entries are not supported");
  for (int g : {16, 32, 64}) {
    if (g_req > 0 && g != g_req) continue;
    for (int r : {1, 2, 4}) {
      if (rows_req > 0 && r != rows_req) continue;
      if ((long)r * g < dim) continue;
      for (int w : {wh, 8, 16})
        if (w >= wh && qt::sub_supported(g, r, w)) return LayoutChoice{r, 1, w, 1, g};
    }
  }
  throw Fail(QT_ENOTSUP, "no compiled sub-wave layout covers this system");
}

void check_params(const qt_mcwf_params* p) {
  if (!p) throw Fail(QT_EINVAL, "null parameters");
  if (!(p->epsRel > 0) || !(p->epsAbs >= 0)) throw Fail(QT_EINVAL, "tolerances must be positive");
  if (!(p->dpLimit > 0) || !(p->overshootTol >= 1)) throw Fail(QT_EINVAL, "dpLimit > 0 and overshootTol >= 1 required");
  if (p->precision != QT_FP64 && p->precision != QT_FP32_DERIV) throw Fail(QT_EINVAL, "unknown precision");
  if (p->controller != QT_CONTROLLER_ODEINT && p->controller != QT_CONTROLLER_GSL)
    throw Fail(QT_EINVAL, "unknown controller");
}

}  // namespace

namespace {

// Constant-offset form of H (SysDev::hoff): Kronecker-structured operators shift each axis by
// fixed amounts, so the entries' column offsets (col - row) often take no more distinct values
// than a row has entries.  Then slot k of every row holds the entry at column row + offs[k]
// (slots in order of first appearance, rows in order; zero and a clamped column where the row
// has none).  cols / vals come out row-major [dim][width] like the ELL table.  False (and
// nothing written) for tables with frequencies or with more distinct offsets than the width.
bool offset_form(const qt::Ell& e, int d, std::vector<int>& offs, std::vector<int>& cols, std::vector<qt::cd>& vals) {
  offs.clear();
  if (e.timeDependent()) return false;
  const int w = e.width;
  for (int r = 0; r < d; ++r)
    for (int k = 0; k < w; ++k) {
      if (e.vals[(size_t)r * w + k] == qt::cd(0.)) continue;
      const int o = e.cols[(size_t)r * w + k] - r;
      if (std::find(offs.begin(), offs.end(), o) == offs.end()) {
        if ((int)offs.size() == w) { offs.clear(); return false; }
        offs.push_back(o);
      }
    }
  offs.resize(w, 0);
  cols.assign((size_t)d * w, 0);
  vals.assign((size_t)d * w, qt::cd(0.));
  for (int r = 0; r < d; ++r)
    for (int k = 0; k < w; ++k) cols[(size_t)r * w + k] = std::min(std::max(r + offs[k], 0), d - 1);
  for (int r = 0; r < d; ++r)
    for (int k = 0; k < w; ++k) {
      const qt::cd z = e.vals[(size_t)r * w + k];
      if (z == qt::cd(0.)) continue;
      const int col = e.cols[(size_t)r * w + k];
      const int slot = (int)(std::find(offs.begin(), offs.end(), col - r) - offs.begin());
      cols[(size_t)r * w + slot] = col;
      vals[(size_t)r * w + slot] = z;
    }
  return true;
}

}  // namespace

struct qt_sys {
  std::unique_ptr<qt::System> h;
  bool uploaded = false;
  bool interactionPicture() const { return h->H.timeDependent() || h->hasEx; }
  // the device kernels take an interaction-picture system when its frequencies fit the table
  void checkPictureSupported() {
    upload();
    if (h->H.timeDependent() && nfreq > qt::kMaxFreqs)
      throw Fail(QT_ENOTSUP, "interaction picture with more than 32 distinct frequencies");
  }
  DevBuf<int> hcol, jcol, ocol, oimag, owid, hfidx, hoff, pvar, psqrt;
  DevBuf<double2> hval, jval, oval, freq, exz;
  DevBuf<double> jwts;   // rate weights [nj][dim] (SysDev::jw)
  int nfreq = 0;          // distinct interaction-picture frequencies of H (0: Schroedinger picture)
  DevBuf<float2> hval32;
  // dictionary form of the constant-offset table (SysDev::hdict)
  DevBuf<double2> hdict;
  DevBuf<float2> hdict32;
  DevBuf<int> hdbase;
  DevBuf<unsigned char> hidx;
  qt::SysDev dev{};

  void upload() {
    if (uploaded) return;
    const qt::System& s = *h;
    const int d = (int)s.dim;
    auto colmajor = [d](const qt::Ell& e, std::vector<int>& c, std::vector<double2>& v) {
      for (int w = 0; w < e.width; ++w)
        for (int r = 0; r < d; ++r) {
          c.push_back(e.cols[(size_t)r * e.width + w]);
          const qt::cd z = e.vals[(size_t)r * e.width + w];
          v.push_back(make_double2(z.real(), z.imag()));
        }
    };
    std::vector<int> c; std::vector<double2> v;
    colmajor(s.H, c, v);
    // constant-offset form of H when it has one (offset_form): uploaded in its place
    std::vector<int> offs;
    {
      std::vector<int> oc;
      std::vector<qt::cd> ov;
      if (offset_form(s.H, d, offs, oc, ov)) {
        const int w0 = s.H.width;
        for (int k = 0; k < w0; ++k)
          for (int r = 0; r < d; ++r) {
            c[(size_t)k * d + r] = oc[(size_t)r * w0 + k];
            const qt::cd z = ov[(size_t)r * w0 + k];
            v[(size_t)k * d + r] = make_double2(z.real(), z.imag());
          }
        hoff.upload(offs);
      }
    }
    hcol.upload(c); hval.upload(v);
    std::vector<float2> v32;
    for (auto& z : v) v32.push_back(make_float2((float)z.x, (float)z.y));
    hval32.upload(v32);
    // dictionary form of the constant-offset table: when every slot's values take at most 256
    // distinct bit patterns (kMaxDict in all, at most kMaxDictSlots slots), the distinct values go
    // to a dictionary, slot by slot, and each entry becomes a byte index into its slot's part
    // (slot-major like the table); the team engines copy both into LDS, so that a gather reads
    // its coefficients from LDS instead of global memory
    int ndict = 0;
    if (!offs.empty() && s.H.width <= qt::kMaxDictSlots) {
      const int w0 = s.H.width;
      std::vector<double2> dict;
      std::vector<int> base(qt::kMaxDictSlots, 0);
      std::vector<unsigned char> idx((size_t)w0 * d, 0);
      bool ok = true;
      for (int k = 0; k < w0 && ok; ++k) {
        std::map<std::pair<unsigned long long, unsigned long long>, int> seen;
        base[k] = (int)dict.size();
        for (int r = 0; r < d && ok; ++r) {
          const double2 z = v[(size_t)k * d + r];
          std::pair<unsigned long long, unsigned long long> key;
          std::memcpy(&key.first, &z.x, 8);
          std::memcpy(&key.second, &z.y, 8);
          auto it = seen.find(key);
          if (it == seen.end()) {
            if (seen.size() == 256 || dict.size() == (size_t)qt::kMaxDict) { ok = false; break; }
            it = seen.emplace(key, (int)seen.size()).first;
            dict.push_back(z);
          }
          idx[(size_t)k * d + r] = (unsigned char)it->second;
        }
      }
      if (ok && !dict.empty()) {
        std::vector<float2> d32;
        for (auto& z : dict) d32.push_back(make_float2((float)z.x, (float)z.y));  // as hval32
        hdict.upload(dict); hdict32.upload(d32); hdbase.upload(base); hidx.upload(idx);
        ndict = (int)dict.size();
      }
    }
    int wj = 1;
    for (auto& j : s.J) wj = std::max(wj, j.width);
    c.clear(); v.clear();
    for (auto& j : s.J) {
      qt::Ell p = j;  // pad to the common width
      if (p.width < wj) {
        qt::Ell q; q.width = wj; q.cols.assign((size_t)d * wj, 0); q.vals.assign((size_t)d * wj, 0.);
        for (int r = 0; r < d; ++r)
          for (int w = 0; w < wj; ++w) {
            q.cols[(size_t)r * wj + w] = w < p.width ? p.cols[(size_t)r * p.width + w] : r;
            q.vals[(size_t)r * wj + w] = w < p.width ? p.vals[(size_t)r * p.width + w] : 0.;
          }
        p = q;
      }
      colmajor(p, c, v);
    }
    if (c.empty()) { c.push_back(0); v.push_back(make_double2(0, 0)); }
    jcol.upload(c); jval.upload(v);
    {  // rate weights (SysDev::jw) when every J_m^dagger J_m is diagonal
      bool diag = !s.J.empty();
      std::vector<double> wts;
      for (auto& j : s.J) {
        std::vector<double> wm(d, 0.);
        std::vector<char> used(d, 0);
        for (int r = 0; r < d && diag; ++r) {
          int cnt = 0;
          for (int k = 0; k < j.width; ++k) {
            const qt::cd z = j.vals[(size_t)r * j.width + k];
            if (z == qt::cd(0.)) continue;
            const int col = j.cols[(size_t)r * j.width + k];
            if (++cnt > 1 || used[col]) { diag = false; break; }
            used[col] = 1;
            wm[col] = std::norm(z);
          }
        }
        wts.insert(wts.end(), wm.begin(), wm.end());
      }
      if (diag) jwts.upload(wts);
      dev.jw = diag ? jwts.p : nullptr;
    }
    int wo = 1;
    for (auto& o : s.O) wo = std::max(wo, o.width);
    c.clear(); v.clear();
    std::vector<int> im, ow;
    for (size_t q = 0; q < s.O.size(); ++q) {
      const qt::Ell& o = s.O[q];
      ow.push_back(o.width);
      for (int w = 0; w < wo; ++w)
        for (int r = 0; r < d; ++r) {
          c.push_back(w < o.width ? o.cols[(size_t)r * o.width + w] : r);
          const qt::cd z = w < o.width ? o.vals[(size_t)r * o.width + w] : qt::cd(0.);
          v.push_back(make_double2(z.real(), z.imag()));
        }
      im.push_back(s.obs[q].imagPart ? 1 : 0);
    }
    if (c.empty()) { c.push_back(0); v.push_back(make_double2(0, 0)); im.push_back(0); ow.push_back(0); }
    ocol.upload(c); oval.upload(v); oimag.upload(im); owid.upload(ow);
    {  // Averaged::process rules of the streamed averages
      std::vector<int> pv(std::max<size_t>(1, s.O.size()), -1), ps(std::max<size_t>(1, s.O.size()), 0);
      for (size_t q = 0; q < s.processVarOf.size() && q < s.O.size(); ++q) { pv[q] = s.processVarOf[q]; ps[q] = s.processSqrt[q]; }
      pvar.upload(pv); psqrt.upload(ps);
    }
    // interaction picture: distinct entry frequencies (index 0 = 0) and per-entry indices,
    // column-major like the values; the Exact diagonal's exponents per basis state
    if (s.H.timeDependent()) {
      std::vector<qt::cd> fl{qt::cd(0.)};
      std::vector<int> fi((size_t)d * s.H.width, 0);
      for (int w = 0; w < s.H.width; ++w)
        for (int r = 0; r < d; ++r) {
          const qt::cd f = s.H.freqs[(size_t)r * s.H.width + w];
          size_t k = 0;
          while (k < fl.size() && fl[k] != f) ++k;
          if (k == fl.size()) fl.push_back(f);
          fi[(size_t)w * d + r] = (int)k;
        }
      nfreq = (int)fl.size();
      if (nfreq <= qt::kMaxFreqs) {
        std::vector<double2> fv;
        for (auto& f : fl) fv.push_back(make_double2(f.real(), f.imag()));
        freq.upload(fv);
        hfidx.upload(fi);
      }
    }
    if (s.hasEx) {
      std::vector<double2> z;
      for (auto& c : s.exactZ) z.push_back(make_double2(c.real(), c.imag()));
      exz.upload(z);
    }
    dev.nfreq = nfreq <= qt::kMaxFreqs ? nfreq : 0;
    dev.freq = freq.p; dev.hfidx = hfidx.p; dev.exz = s.hasEx ? exz.p : nullptr;
    dev.dim = d; dev.wh = s.H.width;
    dev.hcol = hcol.p; dev.hval = hval.p; dev.hval32 = hval32.p;
    dev.hoff = offs.empty() ? nullptr : hoff.p;
    dev.ndict = ndict;
    dev.hdict = ndict ? hdict.p : nullptr; dev.hdict32 = ndict ? hdict32.p : nullptr;
    dev.hdbase = ndict ? hdbase.p : nullptr; dev.hidx = ndict ? hidx.p : nullptr;
    dev.nj = (int)s.J.size(); dev.wj = wj; dev.jcol = jcol.p; dev.jval = jval.p;
    dev.no = (int)s.O.size(); dev.wo = wo; dev.ocol = ocol.p; dev.oval = oval.p; dev.oimag = oimag.p; dev.owid = owid.p;
    dev.pvar = pvar.p; dev.psqrt = psqrt.p;
    uploaded = true;
  }
};

struct qt_ens {
  qt_sys* sys = nullptr;
  int n = 0;
  qt_mcwf_params p{};
  LayoutChoice lay{};
  double time = 0.;           // common time of all trajectories (they all stop at each endTime)
  DevBuf<double2> psi, psiInit;
  DevBuf<double> t, t0, dtDid, dtTry, dtTryInit, intervals, jt, draws, lastObs, rowsDev;
  DevBuf<unsigned long long> acc;   // stream sums of the last launch (acc_add)
  // host-supplied jump draws (qt_mcwf_set_draws): [n][hostDrawsPer], indexed from drawBase[traj]
  DevBuf<double> hostDraws;
  DevBuf<unsigned long long> drawBase;
  int hostDrawsPer = 0;
  DevBuf<unsigned long long> ord;
  DevBuf<unsigned long long> ordBase;  // each trajectory's RNG ordinal at launch start (index 0 of its bulk draws)
  DevBuf<long long> cnt;
  DevBuf<int> jch, work, errflag;
  DevBuf<int> prog;                    // finished chunks per trajectory (interleaved scheduling)
  DevBuf<double2> rhoDev, rhoPart;     // the ensemble's density operator on request (qt_mcwf_rho)
  int lastIntervals = 0;
  int grid = 0;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  double kernelMs = 0.;
  int launches = 0;
  double dtMinSeen = 0.02;  // typical accepted step, sizes the jump-draw budget
  // captured launch sequence (jump draws, queue reset, batched kernel), replayed while the
  // launch description is unchanged
  bool useGraph = false;
  hipStream_t capStream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  qt::EnsDev gKey{};
  long long graphLaunches = 0, graphCaptures = 0;
  void dropGraph() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    gexec = nullptr; graph = nullptr;
  }
  ~qt_ens() {
    dropGraph();
    if (capStream) (void)hipStreamDestroy(capStream);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  }

  qt::EnsDev view() {
    qt::EnsDev E{};
    E.n_traj = n; E.psi = psi.p; E.t = t.p; E.t0 = t0.p; E.dtDid = dtDid.p; E.dtTry = dtTry.p; E.ordinal = ord.p;
    E.seed = p.seed; E.traj0 = p.traj0; E.counters = cnt.p; E.jt = jt.p; E.jch = jch.p; E.jcap = p.jump_log_cap;
    E.epsAbs = p.epsRel; E.epsRel = p.epsAbs;  // ODE.h:205 hands (epsRel, epsAbs) to make_controlled(abs, rel, ...)
    E.dpLimit = p.dpLimit; E.overshootTol = p.overshootTol;
    E.work = work.p; E.err_flag = errflag.p;
    E.gsl = p.controller == QT_CONTROLLER_GSL;
    if (hostDrawsPer > 0) {
      E.draws = hostDraws.p; E.draw_cap = hostDrawsPer; E.draw_base = drawBase.p; E.draws_strict = 1;
    }
    return E;
  }
};

struct qt_master {
  int n = 0, d = 0, nb = 0, nj = 0, wj = 1;
  qt_master_params p{};
  qt_sys* sys0 = nullptr;
  double time = 0.;
  DevBuf<double> hfrag, t, dtDid, dtTry, tss, intervals, obs, dtd, dn, tout;
  DevBuf<unsigned> hmask;          // [point][NB]: non-zero A fragments per row block (bit per k-step)
  DevBuf<double2> exz;             // [point][d] (interaction picture only)
  bool td = false;
  int wr = 0, nrf = 0;             // remainder of H_nH/i outside the MFMA product (all of it: sparse)
  bool sharedH = false;            // one copy of the fragments for every point
  bool sparse = false;             // the sparse kernel (every matrix in HBM, H_nH/i as ELL rows)
  DevBuf<int> rcol, rfi;
  DevBuf<double2> rval, rfreq;
  DevBuf<int> jcol, done;
  DevBuf<double2> jval, x, k;
  DevBuf<long long> cnt;
  DevBuf<double> asbuf;            // autostop: running average + queue ring per point
  DevBuf<int> asmeta, pvar, psqrt; // autostop: started / queue size / head per point; processing rules
  hipEvent_t e0 = nullptr, e1 = nullptr;
  double kernelMs = 0.;
  int launches = 0;
  double* hbuf = nullptr;  // pinned staging for the stream outputs (copies stay asynchronous)
  size_t hcap = 0;
  ~qt_master() {
    if (hbuf) (void)hipHostFree(hbuf);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  }
};

namespace {

// The Master kernel takes an interaction-picture system when it is unitary: the reference's Master
// refuses the others (applicableInMaster, Master.h:73; --picture IP becomes UIP, QM_Picture.cc:45-47).
void master_picture_check(const qt::System& s) {
  if (!s.hasEx) {
    if (s.H.timeDependent()) throw Fail(QT_ENOTSUP, "the Master sweep's interaction picture needs Exact frees");
    return;
  }
  for (auto& z : s.exactZ)
    if (z.real() != 0.)
      throw Fail(QT_EINVAL, "SystemNotApplicable: the Master equation needs a unitary interaction picture (UIP)");
}

// Does entry i (row r) of H_nH/i rotate with exz[r] - exz[col] in the interaction picture?  Then the
// MFMA product takes it (H(tau) = E H E^-1); otherwise -- a term that couples an Exact free with
// frequencies the reference leaves out (Tridiagonal.tcc:61-67), e.g. ParticleAlongCavity's
// dispersive term with its frequency-less nop -- it goes to the sparse remainder.
bool master_entry_rotates(const qt::System& s, long r, size_t i) {
  if (!s.hasEx) return true;
  double scale = 1.;
  for (auto& z : s.exactZ) scale = std::max(scale, std::abs(z));
  const qt::cd f = s.H.freqs.empty() ? qt::cd(0.) : s.H.freqs[i];
  return std::abs(f - (s.exactZ[r] - s.exactZ[s.H.cols[i]])) <= 1e-12 * scale;
}

// what qt_master_create refuses about its parameters and the sweep's systems
void check_master(qt_sys* const* systems, int n_points, const qt_master_params* p) {
  if (!systems || !p || n_points <= 0) throw Fail(QT_EINVAL, "bad arguments");
  for (int k = 0; k < n_points; ++k) if (!systems[k]) throw Fail(QT_EINVAL, "null system");
  if (!(p->epsRel > 0) || !(p->epsAbs >= 0)) throw Fail(QT_EINVAL, "tolerances must be positive");
  if (p->autostop_rep < 0 || (p->autostop_rep > 0 && !(p->autostop_eps > 0.)))
    throw Fail(QT_EINVAL, "autostop needs a positive epsilon and a non-negative repetition");
  const qt::System& s0 = *systems[0]->h;
  if (p->autostop_rep > 0 && (int)s0.O.size() > qt::kMaxAutostopAvr)
    throw Fail(QT_ENOTSUP, "autostop watches at most 64 averages");
  if (p->kernel < QT_MASTER_AUTO || p->kernel > QT_MASTER_SPARSE) throw Fail(QT_EINVAL, "unknown Master kernel");
  if (p->kernel == QT_MASTER_MFMA && !qt::master_supported((int)(s0.dim + 15) / 16))
    throw Fail(QT_ENOTSUP, "the MFMA Master kernel supports dimensions up to 96");
  if (!qt::master_sparse_supported((int)s0.dim)) throw Fail(QT_ENOTSUP, "Master sweep supports dimensions up to 4096");
  for (int k = 0; k < n_points; ++k) {
    const qt::System& s = *systems[k]->h;
    if (s.dim != s0.dim || s.J.size() != s0.J.size())
      throw Fail(QT_EINVAL, "sweep points must share dimension and jump channels");
    master_picture_check(s);
  }
}


// The reference throws at the first non-finite rate (LiouvilleanAveragedCommon.h:82-93, via
// Liouvillean::rates) or jump norm (MCWF_Trajectory.h:338), and at non-finite streamed averages;
// the device flags them per launch (bits 0-1), and a streamed average too large for the
// fixed-point stream sums (bit 2).
void ens_check_flag(int flag) {
  if (flag & 2) throw Fail(QT_ENUMERIC, "Infinite detected in averages");
  if (flag & 1) throw Fail(QT_ENUMERIC, "Infinite detected in MCWF_Trajectory::performJump");
  if (flag & 4) throw Fail(QT_ENUMERIC, "a streamed average exceeds the range of the ensemble sums (2^43 / n_traj)");
  if (flag & 8) throw Fail(QT_EINVAL, "a trajectory needed more jump draws than qt_mcwf_set_draws supplied");
  if (flag & 16) throw Fail(QT_EDEVICE, "a chunk of the interleaved schedule waited too long for its trajectory's previous chunk");
}

// rates of a host state through the host ELL tables: ||J_m psi||^2
std::vector<double> host_rates(const qt::System& s, const std::vector<qt::cd>& psi) {
  std::vector<double> r;
  for (auto& j : s.J) {
    double acc = 0.;
    for (long row = 0; row < s.dim; ++row) {
      qt::cd y = 0.;
      for (int w = 0; w < j.width; ++w) y += j.vals[row * j.width + w] * psi[j.cols[row * j.width + w]];
      acc += std::norm(y);
    }
    r.push_back(acc);
  }
  return r;
}

int resident_groups(const LayoutChoice& L, bool f32, bool td, bool gsl = false) {
  int dev = 0;
  hipCheck(hipGetDevice(&dev), "hipGetDevice");
  hipDeviceProp_t prop;
  hipCheck(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
  int per_cu = L.glanes > 0 ? qt::sub_blocks_per_cu(L.glanes, L.rows, L.wh, f32) : qt::batch_blocks_per_cu(L.rows, L.nw, L.wh, L.occ, f32, td, gsl);
  if (per_cu < 1) per_cu = 1;
  return prop.multiProcessorCount * per_cu;
}

void process_rows(const qt::System& s, double* av) {
  for (size_t q = 0; q < s.processVarOf.size(); ++q)
    if (s.processVarOf[q] >= 0) {
      av[q] -= av[s.processVarOf[q]] * av[s.processVarOf[q]];
      if (s.processSqrt[q]) av[q] = std::sqrt(av[q]);
    }
}

}  // namespace

extern "C" {

const char* qt_last_error(void) { return g_err.c_str(); }

int qt_sys_build(const char* spec, qt_sys** out) {
  return guard([&] {
    if (!spec || !out) throw Fail(QT_EINVAL, "null argument");
    std::unique_ptr<qt_sys> s(new qt_sys);
    s->h.reset(qt::buildFromSpec(spec));
    *out = s.release();
  });
}

}  // extern "C"

namespace {

// the entries of one Tridiagonal term (qt_tridiag): diagonal i's digits select per axis the main
// (|n><n|), lower (|n+K><n|) or upper (|n><n+K|) diagonal; its values run row-major over the
// axes with extent dim (main) or dim - K
qt::Term tridiag_term(const qt::System& s, const qt_tridiag& t) {
  const int R = t.rank;
  if (R < 1 || R > (int)s.dims.size() || !t.axes || !t.k || !t.diags) throw Fail(QT_EINVAL, "bad Tridiagonal term");
  qt::Term term;
  term.explicitFreq = true;
  std::vector<int> lstr(R, 1);
  for (int a = 0; a < R; ++a) {
    const int ax = t.axes[a];
    if (ax < 0 || ax >= (int)s.dims.size()) throw Fail(QT_EINVAL, "Tridiagonal axis out of range");
    for (int b = 0; b < a; ++b)
      if (t.axes[b] == ax) throw Fail(QT_EINVAL, "Tridiagonal axes must differ");
    if (t.k[a] < 0) throw Fail(QT_EINVAL, "negative Tridiagonal offset");
    term.axes.push_back(ax);
    term.op.dims.push_back(s.dims[ax]);
  }
  for (int a = R - 2; a >= 0; --a) lstr[a] = lstr[a + 1] * term.op.dims[a + 1];
  int L = 1;
  for (int a = 0; a < R; ++a) L *= 3;
  std::vector<int> code(R), ext(R);
  for (int i = 0; i < L; ++i) {
    if (!t.diags[i]) continue;
    long n = 1;
    for (int a = R - 1, x = i; a >= 0; --a, x /= 3) {
      code[a] = x % 3;
      // Tridiagonal consistency (Tridiagonal.cc:47-70): off-diagonals need an offset
      if (code[a] && t.k[a] == 0) throw Fail(QT_EINVAL, "Tridiagonal consistency error");
      ext[a] = code[a] ? std::max(0, term.op.dims[a] - t.k[a]) : term.op.dims[a];
      n *= ext[a];
    }
    const double* dv = t.diags[i];
    const double* fv = t.freqs && i > 0 ? t.freqs[i] : nullptr;  // the main diagonal never rotates
    for (long e = 0; e < n; ++e) {
      int row = 0, col = 0;
      long rem = e;
      for (int a = R - 1; a >= 0; --a) {
        const int m = (int)(rem % ext[a]);
        rem /= ext[a];
        row += (m + (code[a] == 1 ? t.k[a] : 0)) * lstr[a];
        col += (m + (code[a] == 2 ? t.k[a] : 0)) * lstr[a];
      }
      term.op.entries.push_back({row, col, qt::cd(dv[2 * e], dv[2 * e + 1]),
                                 fv ? qt::cd(fv[2 * e], fv[2 * e + 1]) : qt::cd(0.)});
    }
  }
  return term;
}

qt::Term coo_term(const qt::System& s, const qt_coo& c) {
  const int R = c.rank;
  if (R < 1 || R > (int)s.dims.size() || !c.axes || c.nnz < 0 || (c.nnz && (!c.rows || !c.cols || !c.vals)))
    throw Fail(QT_EINVAL, "bad sparse term");
  qt::Term term;
  term.explicitFreq = true;
  long size = 1;
  for (int a = 0; a < R; ++a) {
    const int ax = c.axes[a];
    if (ax < 0 || ax >= (int)s.dims.size()) throw Fail(QT_EINVAL, "sparse term axis out of range");
    for (int b = 0; b < a; ++b)
      if (c.axes[b] == ax) throw Fail(QT_EINVAL, "sparse term axes must differ");
    term.axes.push_back(ax);
    term.op.dims.push_back(s.dims[ax]);
    size *= s.dims[ax];
  }
  for (long long e = 0; e < c.nnz; ++e) {
    if (c.rows[e] < 0 || c.rows[e] >= size || c.cols[e] < 0 || c.cols[e] >= size)
      throw Fail(QT_EINVAL, "sparse term index out of range");
    const qt::cd f = c.freqs ? qt::cd(c.freqs[2 * e], c.freqs[2 * e + 1]) : qt::cd(0.);
    term.op.entries.push_back({(int)c.rows[e], (int)c.cols[e], qt::cd(c.vals[2 * e], c.vals[2 * e + 1]), f});
  }
  return term;
}

std::vector<qt::Term> operator_terms(const qt::System& s, const qt_operator& o) {
  if (o.n_tri < 0 || o.n_coo < 0 || (o.n_tri && !o.tri) || (o.n_coo && !o.coo)) throw Fail(QT_EINVAL, "bad operator");
  std::vector<qt::Term> out;
  for (int i = 0; i < o.n_tri; ++i) out.push_back(tridiag_term(s, o.tri[i]));
  for (int i = 0; i < o.n_coo; ++i) out.push_back(coo_term(s, o.coo[i]));
  return out;
}

}  // namespace

extern "C" {

int qt_sys_build_terms(const qt_system_desc* d, qt_sys** out) {
  return guard([&] {
    if (!d || !out) throw Fail(QT_EINVAL, "null argument");
    if (d->n_frees < 1 || !d->frees) throw Fail(QT_EINVAL, "system without free elements");
    if (d->n_jumps < 0 || d->n_avr < 0 || (d->n_jumps && !d->jumps) ||
        (d->n_avr && (!d->avr || !d->avr_imag || !d->avr_var_of || !d->avr_sqrt)))
      throw Fail(QT_EINVAL, "bad jump or average list");
    if (!(d->highest_frequency >= 0.)) throw Fail(QT_EINVAL, "bad highest frequency");
    std::unique_ptr<qt::System> sp(new qt::System);
    qt::System& s = *sp;
    for (int a = 0; a < d->n_frees; ++a) {
      const qt_free_desc& f = d->frees[a];
      if (f.dim < 1) throw Fail(QT_EINVAL, "free of no dimension");
      qt::FreeInfo fi;
      fi.kind = "terms";
      fi.dim = f.dim;
      if (f.exact) {
        fi.hasEx = true;
        for (int n = 0; n < f.dim; ++n) fi.exd.push_back(qt::cd(f.exact[2 * n], f.exact[2 * n + 1]));
      }
      s.frees.push_back(fi);
      s.dims.push_back(f.dim);
    }
    s.strides.assign(s.dims.size(), 1);
    s.dim = 1;
    for (int i = (int)s.dims.size() - 1; i >= 0; --i) {
      s.strides[i] = s.dim;
      s.dim *= s.dims[i];
      if (s.dim > (1L << 30)) throw Fail(QT_EINVAL, "system dimension too large");
    }
    s.hTerms = operator_terms(s, d->h);
    for (int m = 0; m < d->n_jumps; ++m) s.jumps.push_back(operator_terms(s, d->jumps[m]));
    for (int q = 0; q < d->n_avr; ++q) {
      qt::Observable o;
      o.terms = operator_terms(s, d->avr[q]);
      o.imagPart = d->avr_imag[q] != 0;
      s.obs.push_back(o);
      const int v = d->avr_var_of[q];
      if (v < -1 || v >= d->n_avr) throw Fail(QT_EINVAL, "bad variance rule");
      s.processVarOf.push_back(v);
      s.processSqrt.push_back(d->avr_sqrt[q] ? 1 : 0);
    }
    s.highestFrequency = d->highest_frequency;
    for (auto& f : s.frees) s.hasEx = s.hasEx || f.hasEx;
    if (s.hasEx) {
      s.exactZ.assign(s.dim, 0.);
      for (long r = 0; r < s.dim; ++r)
        for (size_t a = 0; a < s.frees.size(); ++a)
          if (s.frees[a].hasEx) s.exactZ[r] += s.frees[a].exd[(r / s.strides[a]) % s.dims[a]];
    }
    try {
      s.lower();
    } catch (const std::invalid_argument& ex) {
      throw Fail(QT_EINVAL, ex.what());
    }
    std::unique_ptr<qt_sys> h(new qt_sys);
    h->h = std::move(sp);
    *out = h.release();
  });
}

int qt_sys_destroy(qt_sys* sys) { delete sys; return QT_OK; }

int qt_sys_info(const qt_sys* sys, qt_sys_info_t* info) {
  return guard([&] {
    if (!sys || !info) throw Fail(QT_EINVAL, "null argument");
    const qt::System& s = *sys->h;
    std::memset(info, 0, sizeof(*info));
    info->dim = s.dim;
    info->rank = (int)s.dims.size();
    for (size_t i = 0; i < s.dims.size() && i < 16; ++i) info->dims[i] = s.dims[i];
    info->n_jumps = (int)s.J.size();
    info->n_avr = (int)s.O.size();
    info->h_width = s.H.width;
    int wj = 0; for (auto& j : s.J) wj = std::max(wj, j.width);
    int wo = 0; for (auto& o : s.O) wo = std::max(wo, o.width);
    info->j_width = wj; info->o_width = wo;
    info->time_dependent = s.H.timeDependent() || s.hasEx;
    info->highest_frequency = s.highestFrequency;
    long long nnz = 0;
    for (auto& z : s.H.vals) if (z != qt::cd(0.)) ++nnz;
    info->h_nnz = nnz;
    long long mt = 0;
    for (auto& t : s.hTerms) mt += (long long)t.op.entries.size() * (s.dim / std::max(1, t.op.size()));
    info->m_terms = mt;
    std::vector<int> offs, oc;
    std::vector<qt::cd> ov;
    info->h_offsets = offset_form(s.H, (int)s.dim, offs, oc, ov) ? (int)offs.size() : 0;
  });
}

int qt_sys_rules(const qt_sys* sys, double* exact_z, int* avr_imag, int* avr_var_of, int* avr_sqrt) {
  return guard([&] {
    if (!sys) throw Fail(QT_EINVAL, "null argument");
    const qt::System& s = *sys->h;
    if (exact_z)
      for (long r = 0; r < s.dim; ++r) {
        const qt::cd z = s.hasEx ? s.exactZ[r] : qt::cd(0.);
        exact_z[2 * r] = z.real(); exact_z[2 * r + 1] = z.imag();
      }
    for (size_t q = 0; q < s.obs.size(); ++q) {
      if (avr_imag) avr_imag[q] = s.obs[q].imagPart ? 1 : 0;
      if (avr_var_of) avr_var_of[q] = s.processVarOf[q];
      if (avr_sqrt) avr_sqrt[q] = s.processSqrt[q] ? 1 : 0;
    }
  });
}

int qt_sys_table(const qt_sys* sys, int which, int* width, int* cols, double* vals, double* freqs) {
  return guard([&] {
    if (!sys) throw Fail(QT_EINVAL, "null argument");
    const qt::System& s = *sys->h;
    const qt::Ell* e = nullptr;
    qt::Ell form;  // which == -1: H in its constant-offset form
    if (which == -1) {
      std::vector<int> offs;
      if (!offset_form(s.H, (int)s.dim, offs, form.cols, form.vals))
        throw Fail(QT_EINVAL, "H has no constant-offset form");
      form.width = s.H.width;
      e = &form;
    } else if (which == 0) e = &s.H;
    else if (which - 1 < (int)s.J.size()) e = &s.J[which - 1];
    else if (which - 1 - (int)s.J.size() < (int)s.O.size()) e = &s.O[which - 1 - s.J.size()];
    else throw Fail(QT_EINVAL, "table index out of range");
    if (width) *width = e->width;
    size_t n = (size_t)s.dim * e->width;
    for (size_t i = 0; i < n; ++i) {
      if (cols) cols[i] = e->cols[i];
      if (vals) { vals[2 * i] = e->vals[i].real(); vals[2 * i + 1] = e->vals[i].imag(); }
      if (freqs) {
        qt::cd f = e->freqs.empty() ? qt::cd(0.) : e->freqs[i];
        freqs[2 * i] = f.real(); freqs[2 * i + 1] = f.imag();
      }
    }
  });
}

int qt_apply_heff(qt_sys* sys, double t, double t0, const double* psi, double* dpsidt, int n, int precision,
                  qt_stream stream) {
  return guard([&] {
    if (!sys || !psi || !dpsidt || n <= 0) throw Fail(QT_EINVAL, "bad arguments");
    sys->checkPictureSupported();
    LayoutChoice L = pick_layout(sys->h->dim, sys->h->H.width, 0, 0);
    hipCheck(qt::launch_apply_heff(L.rows, L.nw, L.wh, precision == QT_FP32_DERIV, sys->interactionPicture(), sys->dev,
                                   (const double2*)psi,
                                   (double2*)dpsidt, n, t - t0, (hipStream_t)stream), "apply_heff launch");
  });
}

int qt_mcwf_create(qt_sys* sys, int n_traj, const qt_mcwf_params* p, const double* psi0, qt_ens** out) {
  return guard([&] {
    if (!sys || !psi0 || !out || n_traj <= 0) throw Fail(QT_EINVAL, "bad arguments");
    if (n_traj >= (1 << 23)) throw Fail(QT_ENOTSUP, "at most 2^23 - 1 trajectories per ensemble (range of the stream sums)");
    check_params(p);
    const qt::System& s = *sys->h;
    if ((int)s.J.size() > qt::kMaxJumps) throw Fail(QT_ENOTSUP, "too many jump channels");
    const long d = s.dim;
    std::vector<qt::cd> psi(d);
    for (long i = 0; i < d; ++i) psi[i] = qt::cd(psi0[2 * i], psi0[2 * i + 1]);
    double dtInit = p->dtInit > 0 ? p->dtInit : 1. / (10. * s.highestFrequency);
    // constructor check (MCWF_Trajectory.h:114-117): dtTry <= dpLimit / total rate at t = 0; the
    // rates pass LiouvilleanAveragedCommon::average's check (LiouvilleanAveragedCommon.h:82-93).
    // Host tables only: refused before the device is touched.
    double dtTry = dtInit;
    if (!s.J.empty()) {
      auto r = host_rates(s, psi);
      double tot = 0.; for (double x : r) tot += x;
      if (!std::isfinite(tot)) throw Fail(QT_ENUMERIC, "Infinite detected in averages");
      dtTry = std::min(dtTry, p->dpLimit / tot);
    }
    sys->checkPictureSupported();
    if (sys->interactionPicture() && p->layout_group_lanes > 0)
      throw Fail(QT_ENOTSUP, "sub-wave layouts run the Schroedinger picture only");
    if (p->controller == QT_CONTROLLER_GSL && p->layout_group_lanes > 0)
      throw Fail(QT_ENOTSUP, "sub-wave layouts run the odeint controller only");
    std::unique_ptr<qt_ens> e(new qt_ens);
    e->sys = sys; e->n = n_traj; e->p = *p;
    if (e->p.jump_log_cap < 1) e->p.jump_log_cap = 1;
    const bool useSub = p->layout_group_lanes > 0;
    if (useSub) {
      e->lay = pick_sub(s.dim, s.H.width, p->layout_group_lanes > 0 ? p->layout_group_lanes : default_group_lanes(s.dim),
                        p->layout_rows);
    } else {
      e->lay = pick_layout(s.dim, s.H.width, p->layout_rows, p->layout_waves);
      const bool gsl = p->controller == QT_CONTROLLER_GSL;
      if (sys->interactionPicture() || gsl) {
        // interaction-picture and GSL variants: no occupancy hint, the ELL width may round up
        e->lay.occ = 1;
        int w = e->lay.wh;
        while (w <= 16 && !qt::batch_supported(e->lay.rows, e->lay.nw, w, 1, sys->interactionPicture(), gsl)) w *= 2;
        if (w > 16) throw Fail(QT_ENOTSUP, "no compiled variant of this picture and controller for this layout");
        e->lay.wh = w;
      } else {
        e->lay.occ = p->layout_occupancy > 0 ? p->layout_occupancy : default_occupancy(e->lay);
        if (!qt::batch_supported(e->lay.rows, e->lay.nw, e->lay.wh, e->lay.occ, false, false))
          throw Fail(QT_ENOTSUP, "requested occupancy variant is not compiled in");
      }
    }
    e->psi.alloc((size_t)n_traj * d);
    std::vector<double2> rep((size_t)n_traj * d);
    for (int k = 0; k < n_traj; ++k)
      for (long i = 0; i < d; ++i) rep[(size_t)k * d + i] = make_double2(psi[i].real(), psi[i].imag());
    e->psi.upload(rep);
    e->psiInit.upload(rep);
    e->t.upload(std::vector<double>(n_traj, 0.));
    e->t0.upload(std::vector<double>(n_traj, 0.));
    e->dtDid.upload(std::vector<double>(n_traj, 0.));
    e->dtTry.upload(std::vector<double>(n_traj, dtTry));
    e->dtTryInit.upload(std::vector<double>(n_traj, dtTry));
    e->ord.upload(std::vector<unsigned long long>(n_traj, 0ull));
    e->cnt.upload(std::vector<long long>((size_t)n_traj * qt::kCounters, 0));
    e->jt.alloc((size_t)n_traj * e->p.jump_log_cap);
    e->jch.alloc((size_t)n_traj * e->p.jump_log_cap);
    e->work.alloc(1);
    e->prog.alloc(n_traj);
    e->ordBase.alloc(n_traj);
    e->lastObs.alloc((size_t)n_traj * std::max<size_t>(1, s.O.size()));  // also what qt_mcwf_advance_noalloc writes
    e->errflag.upload(std::vector<int>(1, 0));
    const int perGroup = e->lay.glanes > 0 ? 64 / e->lay.glanes : 1;  // trajectories per work group
    e->grid = std::min((n_traj + perGroup - 1) / perGroup, resident_groups(e->lay, p->precision == QT_FP32_DERIV, sys->interactionPicture(),
                                                                        p->controller == QT_CONTROLLER_GSL));
    hipCheck(hipEventCreate(&e->e0), "hipEventCreate");
    hipCheck(hipEventCreate(&e->e1), "hipEventCreate");
    *out = e.release();
  });
}

int qt_mcwf_destroy(qt_ens* ens) { delete ens; return QT_OK; }

// Interleaved scheduling of the one-wavefront engines (EnsDev::chunk): with more trajectories than
// resident wavefronts, a launch is cut into chunks of about n_intervals / 8 stream intervals
// (qt_mcwf_params::sched_chunk overrides: > 0 intervals per chunk, < 0 off) and every
// trajectory's chunk c runs before any chunk c + 1.  Results do not depend on it: draws are
// indexed by RNG ordinal and the stream sums are order-independent.
static void ens_schedule(const qt_ens* e, qt::EnsDev& E) {
  E.chunk = 0; E.prog = nullptr;
  if (e->lay.glanes > 0 || e->lay.nw != 1 || E.dc > 0 || E.n_intervals < 2) return;
  int ch = e->p.sched_chunk;
  if (ch < 0) return;
  if (ch == 0) {
    if (e->n <= e->grid) return;  // every trajectory is resident from the start
    ch = (E.n_intervals + 7) / 8;  // JC-60, 16384 trajectories, 200 intervals: 14.4 ms with 25, 17 ms whole
  }
  const long long items = (long long)((E.n_intervals + ch - 1) / ch) * e->n;
  if (items <= e->n || items > (1ll << 30)) return;  // one chunk, or a queue beyond the 32-bit head
  E.chunk = ch; E.prog = e->prog.p;
}

// Resets the work queue (and the chunk progress counts) before a launch, with fill kernels so
// that the sequence can be captured.
static void ens_reset_queue(qt_ens* e, const qt::EnsDev& E, hipStream_t st) {
  hipCheck(qt::launch_fill_int(e->work.p, 1, 0, st), "reset work queue");
  if (E.chunk > 0) hipCheck(qt::launch_fill_int(E.prog, e->n, 0, st), "reset chunk progress");
}

// Per-launch preparation of one ensemble: uploads the intervals, sizes the stream outputs,
// generates the launch's jump draws (unless launchDraws is false: then only sized); returns the
// kernel's view of the ensemble.
static qt::EnsDev ens_prepare(qt_ens* e, const double* intervals, int n_int, hipStream_t st, bool bulkDraws,
                              bool launchDraws = true) {
  const qt::System& s = *e->sys->h;
  if (n_int <= 0) throw Fail(QT_EINVAL, "no intervals");
  for (int k = 0; k < n_int; ++k) if (!(intervals[k] >= 0) || !std::isfinite(intervals[k])) throw Fail(QT_EINVAL, "intervals must be finite and non-negative");
  e->intervals.alloc(n_int);
  hipCheck(hipMemcpyAsync(e->intervals.p, intervals, n_int * sizeof(double), hipMemcpyHostToDevice, st), "intervals H2D");
  const int no = (int)s.O.size();
  e->acc.alloc((size_t)qt::kAccStripes * n_int * (no + 1) * 2);
  e->lastObs.alloc((size_t)e->n * std::max(1, no));
  qt::EnsDev E = e->view();
  E.intervals = e->intervals.p; E.n_intervals = n_int;
  E.acc = e->acc.p; E.acc_max = std::ldexp(1.0, 43) / e->n; E.last_obs = e->lastObs.p;
  // jump draws for the launch, generated in bulk (one Philox per lane); trajectories that take
  // more steps than the budget evaluate Philox in the kernel
  if (!s.J.empty() && bulkDraws && !e->hostDrawsPer) {
    double span = 0.; for (int k = 0; k < n_int; ++k) span += intervals[k];
    const int cap = (int)std::min(2048.0, std::max(64.0, std::ceil(span / std::max(e->dtMinSeen, 1e-6) * 1.25) + 8));
    e->draws.alloc((size_t)e->n * cap);
    if (launchDraws)
      hipCheck(qt::launch_jump_draws(e->ord.p, e->p.seed, e->p.traj0, e->n, cap, e->draws.p, st, e->ordBase.p), "jump draws");
    // index 0 of a trajectory's draws is its ordinal at launch start (the draw kernel records it):
    // a chunk that resumes the trajectory mid-launch finds its draws from there
    E.draws = e->draws.p; E.draw_cap = cap; E.draw_base = e->ordBase.p;
  }
  ens_schedule(e, E);
  return E;
}

// 64-bit words of a launch's stream sums
static size_t ens_acc_words(const qt_ens* e, int n_int) {
  return (size_t)qt::kAccStripes * n_int * (e->sys->h->O.size() + 1) * 2;
}

static void ens_kernel(qt_ens* e, const qt::EnsDev& E, hipStream_t st) {
  if (e->lay.glanes > 0)
    hipCheck(qt::launch_sub(e->lay.glanes, e->lay.rows, e->lay.wh, e->p.precision == QT_FP32_DERIV, e->grid, e->sys->dev, E, st),
             "mcwf sub-wave launch");
  else
    hipCheck(qt::launch_batch(e->lay.rows, e->lay.nw, e->lay.wh, e->lay.occ, e->p.precision == QT_FP32_DERIV,
                              e->sys->interactionPicture(), e->p.controller == QT_CONTROLLER_GSL, e->grid,
                              e->sys->dev, E, st),
             "mcwf batch launch");
}

static void ens_launch(qt_ens* e, const double* intervals, int n_int, hipStream_t st, bool timed) {
  if (e->useGraph) {
    // the launch description (buffers, interval count, draw budget) decides whether the
    // captured graph still applies; the interval values are re-uploaded before every replay
    qt::EnsDev E = ens_prepare(e, intervals, n_int, st, e->lay.glanes == 0, false);
    if (!e->gexec || std::memcmp(&E, &e->gKey, sizeof(E)) != 0) {
      e->dropGraph();
      if (!e->capStream) hipCheck(hipStreamCreate(&e->capStream), "hipStreamCreate");
      hipCheck(hipStreamBeginCapture(e->capStream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
      if (E.draws && !E.draws_strict)
        hipCheck(qt::launch_jump_draws(e->ord.p, e->p.seed, e->p.traj0, e->n, E.draw_cap, e->draws.p, e->capStream,
                                       e->ordBase.p), "jump draws");
      hipCheck(qt::launch_zero_u64(E.acc, ens_acc_words(e, n_int), e->capStream), "zero stream sums");
      ens_reset_queue(e, E, e->capStream);
      ens_kernel(e, E, e->capStream);
      hipCheck(hipStreamEndCapture(e->capStream, &e->graph), "hipStreamEndCapture");
      hipCheck(hipGraphInstantiate(&e->gexec, e->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
      e->gKey = E;
      ++e->graphCaptures;
    }
    if (timed) hipCheck(hipEventRecord(e->e0, st), "hipEventRecord");
    hipCheck(hipGraphLaunch(e->gexec, st), "hipGraphLaunch");
    if (timed) hipCheck(hipEventRecord(e->e1, st), "hipEventRecord");
    ++e->graphLaunches;
  } else {
    qt::EnsDev E = ens_prepare(e, intervals, n_int, st, e->lay.glanes == 0);
    hipCheck(qt::launch_zero_u64(E.acc, ens_acc_words(e, n_int), st), "zero stream sums");
    ens_reset_queue(e, E, st);
    if (timed) hipCheck(hipEventRecord(e->e0, st), "hipEventRecord");
    ens_kernel(e, E, st);
    if (timed) hipCheck(hipEventRecord(e->e1, st), "hipEventRecord");
  }
  e->lastIntervals = n_int;
  for (int k = 0; k < n_int; ++k) e->time += intervals[k];
}

// The ensemble rows of the last launch (time, mean dtDid, processed averages) from its stream
// sums, made on the device and copied to the host; synchronises the stream.  Throws the
// reference's exception for a non-finite rate, jump norm or streamed average.
static void ens_rows(qt_ens* e, double tStart, const double* intervals, int n_int, double* rows, hipStream_t st) {
  const qt::System& s = *e->sys->h;
  const int no = (int)s.O.size();
  (void)intervals;  // on the device since ens_prepare
  e->rowsDev.alloc((size_t)n_int * (2 + no));
  hipCheck(qt::launch_stream_rows(e->acc.p, n_int, no, e->n, tStart, e->intervals.p, e->sys->dev.pvar, e->sys->dev.psqrt,
                                  e->rowsDev.p, st), "stream rows launch");
  std::vector<double> r((size_t)n_int * (2 + no));
  hipCheck(hipMemcpyAsync(r.data(), e->rowsDev.p, r.size() * sizeof(double), hipMemcpyDeviceToHost, st), "D2H");
  int flag = 0;
  hipCheck(hipMemcpyAsync(&flag, e->errflag.p, sizeof(int), hipMemcpyDeviceToHost, st), "D2H");
  hipCheck(hipStreamSynchronize(st), "hipStreamSynchronize");
  ens_check_flag(flag);
  for (double v : r) if (std::isnan(v)) throw Fail(QT_ENUMERIC, "Infinite detected in averages");
  if (rows) std::memcpy(rows, r.data(), r.size() * sizeof(double));
}

int qt_mcwf_advance(qt_ens* e, const double* intervals, int n_int, double* rows, qt_stream stream) {
  return guard([&] {
    if (!e || !intervals) throw Fail(QT_EINVAL, "null argument");
    hipStream_t st = (hipStream_t)stream;
    const double tStart = e->time;
    ens_launch(e, intervals, n_int, st, true);
    ens_rows(e, tStart, intervals, n_int, rows, st);
    float ms = 0.f;
    hipCheck(hipEventElapsedTime(&ms, e->e0, e->e1), "hipEventElapsedTime");
    e->kernelMs += ms; e->launches += 1;
  });
}

int qt_mcwf_reset(qt_ens* e, qt_stream stream) {
  return guard([&] {
    if (!e) throw Fail(QT_EINVAL, "null argument");
    hipStream_t st = (hipStream_t)stream;
    const size_t n = e->n, d = e->sys->h->dim;
    hipCheck(hipMemcpyAsync(e->psi.p, e->psiInit.p, n * d * sizeof(double2), hipMemcpyDeviceToDevice, st), "reset psi");
    hipCheck(hipMemcpyAsync(e->dtTry.p, e->dtTryInit.p, n * sizeof(double), hipMemcpyDeviceToDevice, st), "reset dtTry");
    hipCheck(hipMemsetAsync(e->t.p, 0, n * sizeof(double), st), "reset t");
    hipCheck(hipMemsetAsync(e->t0.p, 0, n * sizeof(double), st), "reset t0");
    hipCheck(hipMemsetAsync(e->dtDid.p, 0, n * sizeof(double), st), "reset dtDid");
    hipCheck(hipMemsetAsync(e->ord.p, 0, n * sizeof(unsigned long long), st), "reset ordinal");
    hipCheck(hipMemsetAsync(e->cnt.p, 0, n * qt::kCounters * sizeof(long long), st), "reset counters");
    hipCheck(qt::launch_fill_int(e->errflag.p, 1, 0, st), "reset error flag");
    e->time = 0.;
  });
}

}  // extern "C"

// ------------------------------------------------------------- allocation-free advance

// The per-launch buffers of qt_mcwf_advance_noalloc (intervals, stream sums, rows) live in this
// static device array; each GPU the library runs on holds its own copy until the process ends.
// In host code `scratch` names a host-side stand-in, not a device address: its device address is
// taken once per device with hipGetSymbolAddress (scratch_base) and only that address is used.
constexpr size_t kNoallocScratchBytes = (size_t)8 << 20;
static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

namespace {

std::mutex g_scratch_mu;
constexpr int kMaxDevices = 64;
unsigned char* g_scratch_base[kMaxDevices] = {};

// device address of the current device's scratch (looked up on its first use on that device)
unsigned char* scratch_base() {
  int dev = 0;
  hipCheck(hipGetDevice(&dev), "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) throw Fail(QT_EDEVICE, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_base[dev]) {
    void* p = nullptr;
    hipCheck(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (!p) throw Fail(QT_EDEVICE, "hipGetSymbolAddress(scratch) returned null");
    g_scratch_base[dev] = static_cast<unsigned char*>(p);
  }
  return g_scratch_base[dev];
}

}  // namespace

extern "C" {

int qt_mcwf_advance_noalloc(qt_ens* e, const double* intervals, int n_int, double* rows, qt_stream stream) {
  return guard([&] {
    if (!e || !intervals) throw Fail(QT_EINVAL, "null argument");
    if (n_int <= 0) throw Fail(QT_EINVAL, "no intervals");
    for (int k = 0; k < n_int; ++k)
      if (!(intervals[k] >= 0) || !std::isfinite(intervals[k])) throw Fail(QT_EINVAL, "intervals must be finite and non-negative");
    const qt::System& s = *e->sys->h;
    const int no = (int)s.O.size();
    // scratch layout: intervals | stream sums | rows, each on a 256-byte boundary
    size_t off = 0;
    auto take = [&](size_t bytes) { const size_t o = off; off += (bytes + 255) / 256 * 256; return o; };
    const size_t oIv = take((size_t)n_int * sizeof(double));
    const size_t oAcc = take(ens_acc_words(e, n_int) * sizeof(unsigned long long));
    const size_t oRows = take((size_t)n_int * (2 + no) * sizeof(double));
    if (off > kNoallocScratchBytes)
      throw Fail(QT_EINVAL, "n_intervals = " + std::to_string(n_int) + " with " + std::to_string(no) +
                                " averages needs " + std::to_string(off) + " bytes of scratch; the allocation-free advance has " +
                                std::to_string(kNoallocScratchBytes) + " (16 n_intervals (n_avr + 1) 16 bytes of stream sums)");
    unsigned char* base = scratch_base();
    double* ivDev = reinterpret_cast<double*>(base + oIv);
    unsigned long long* acc = reinterpret_cast<unsigned long long*>(base + oAcc);
    double* rowsDev = reinterpret_cast<double*>(base + oRows);
    hipStream_t st = (hipStream_t)stream;
    e->sys->upload();  // made at qt_mcwf_create: no allocation here
    hipCheck(hipMemcpyAsync(ivDev, intervals, (size_t)n_int * sizeof(double), hipMemcpyHostToDevice, st), "intervals H2D");
    qt::EnsDev E = e->view();
    E.intervals = ivDev; E.n_intervals = n_int;
    E.acc = acc; E.acc_max = std::ldexp(1.0, 43) / e->n; E.last_obs = e->lastObs.p;
    if (!e->hostDrawsPer) { E.draws = nullptr; E.draw_cap = 0; }  // jump draws from the counter RNG inside the kernel
    ens_schedule(e, E);
    const double tStart = e->time;
    hipCheck(qt::launch_zero_u64(acc, ens_acc_words(e, n_int), st), "zero stream sums");
    ens_reset_queue(e, E, st);
    ens_kernel(e, E, st);
    hipCheck(qt::launch_stream_rows(acc, n_int, no, e->n, tStart, ivDev, e->sys->dev.pvar, e->sys->dev.psqrt, rowsDev, st),
             "stream rows launch");
    int flag = 0;
    hipCheck(hipMemcpyAsync(&flag, e->errflag.p, sizeof(int), hipMemcpyDeviceToHost, st), "D2H");
    if (rows) hipCheck(hipMemcpyAsync(rows, rowsDev, (size_t)n_int * (2 + no) * sizeof(double), hipMemcpyDeviceToHost, st), "D2H");
    hipCheck(hipStreamSynchronize(st), "hipStreamSynchronize");
    for (int k = 0; k < n_int; ++k) e->time += intervals[k];
    e->lastIntervals = n_int;
    ens_check_flag(flag);
    if (rows)
      for (size_t i = 0; i < (size_t)n_int * (2 + no); ++i)
        if (std::isnan(rows[i])) throw Fail(QT_ENUMERIC, "Infinite detected in averages");
  });
}

int qt_mcwf_advance_dc(qt_ens* e, double t_end, int dc, int max_rows, double* rows, int* n_rows, qt_stream stream) {
  return guard([&] {
    if (!e || !rows || !n_rows || dc < 1 || max_rows < 1 || !std::isfinite(t_end)) throw Fail(QT_EINVAL, "bad arguments");
    if (e->n != 1) throw Fail(QT_EINVAL, "dc mode drives a single trajectory (an Ensemble has no step(), Trajectory.h:317-323)");
    if (e->lay.glanes > 0) throw Fail(QT_ENOTSUP, "dc mode needs a layout of one or more wavefronts per trajectory");
    if (t_end < e->time) throw Fail(QT_EINVAL, "t_end lies before the trajectory's time");
    *n_rows = 0;
    if (e->time >= t_end) return;  // the run loop's `while (t < T)` does not start
    const qt::System& s = *e->sys->h;
    const int no = (int)s.O.size();
    hipStream_t st = (hipStream_t)stream;
    std::vector<double> iv(max_rows, 0.);
    iv[0] = t_end - e->time;  // sizes the launch's draw budget; dc mode ignores the intervals
    qt::EnsDev E = ens_prepare(e, iv.data(), max_rows, st, true);
    E.dc = dc; E.t_end = t_end;
    ens_schedule(e, E);  // (dc mode runs unchunked)
    e->rowsDev.alloc((size_t)max_rows * (2 + no));
    E.rows_out = e->rowsDev.p;
    hipCheck(qt::launch_zero_u64(E.acc, ens_acc_words(e, max_rows), st), "zero stream sums");
    ens_reset_queue(e, E, st);
    ens_kernel(e, E, st);
    std::vector<double> r((size_t)max_rows * (2 + no));
    hipCheck(hipMemcpyAsync(r.data(), e->rowsDev.p, r.size() * sizeof(double), hipMemcpyDeviceToHost, st), "D2H");
    int flag = 0;
    double tNow = 0.;
    hipCheck(hipMemcpyAsync(&flag, e->errflag.p, sizeof(int), hipMemcpyDeviceToHost, st), "D2H");
    hipCheck(hipMemcpyAsync(&tNow, e->t.p, sizeof(double), hipMemcpyDeviceToHost, st), "D2H");
    hipCheck(hipStreamSynchronize(st), "hipStreamSynchronize");
    e->time = tNow;
    e->lastIntervals = max_rows;
    ens_check_flag(flag);
    // the rows up to the one at t_end (the run loop streams the end instant once, Trajectory.h:597)
    int n = 0;
    while (n < max_rows) {
      double* row = r.data() + (size_t)n * (2 + no);
      for (int q = 0; q < no; ++q) if (!std::isfinite(row[2 + q])) throw Fail(QT_ENUMERIC, "Infinite detected in averages");
      process_rows(s, row + 2);
      ++n;
      if (row[0] >= t_end) break;
    }
    std::memcpy(rows, r.data(), (size_t)n * (2 + no) * sizeof(double));
    *n_rows = n;
  });
}

int qt_mcwf_set_draws(qt_ens* e, const double* draws, int n_per) {
  return guard([&] {
    if (!e || n_per < 0 || (n_per > 0 && !draws)) throw Fail(QT_EINVAL, "bad arguments");
    if (e->lay.glanes > 0) throw Fail(QT_ENOTSUP, "host-supplied jump draws need a layout of one or more wavefronts per trajectory");
    hipCheck(hipDeviceSynchronize(), "hipDeviceSynchronize");
    e->dropGraph();
    if (n_per == 0) { e->hostDrawsPer = 0; return; }
    const size_t n = (size_t)e->n * n_per;
    for (size_t i = 0; i < n; ++i)
      if (!(draws[i] >= 0. && draws[i] < 1.)) throw Fail(QT_EINVAL, "jump draws must lie in [0, 1)");
    e->hostDraws.alloc(n);
    hipCheck(hipMemcpy(e->hostDraws.p, draws, n * sizeof(double), hipMemcpyHostToDevice), "H2D");
    e->drawBase.alloc(e->n);
    hipCheck(hipMemcpy(e->drawBase.p, e->ord.p, e->n * sizeof(unsigned long long), hipMemcpyDeviceToDevice), "D2D");
    hipCheck(hipDeviceSynchronize(), "hipDeviceSynchronize");
    e->hostDrawsPer = n_per;
  });
}

int qt_mcwf_advance_async(qt_ens* e, const double* intervals, int n_int, qt_stream stream) {
  return guard([&] {
    if (!e || !intervals) throw Fail(QT_EINVAL, "null argument");
    ens_launch(e, intervals, n_int, (hipStream_t)stream, false);
  });
}

int qt_mcwf_set_graph(qt_ens* e, int enable, long long* captures, long long* replays) {
  return guard([&] {
    if (!e) throw Fail(QT_EINVAL, "null argument");
    if (enable >= 0) {
      e->useGraph = enable != 0;
      if (!e->useGraph) e->dropGraph();
    }
    if (captures) *captures = e->graphCaptures;
    if (replays) *replays = e->graphLaunches;
  });
}

int qt_mcwf_traj_averages(qt_ens* e, double* out) {
  return guard([&] {
    if (!e || !out) throw Fail(QT_EINVAL, "null argument");
    const qt::System& s = *e->sys->h;
    const int no = (int)s.O.size();
    if (!e->lastIntervals) throw Fail(QT_EINVAL, "no interval advanced yet");
    if (no) hipCheck(hipMemcpy(out, e->lastObs.p, (size_t)e->n * no * sizeof(double), hipMemcpyDeviceToHost), "D2H");
  });
}

int qt_mcwf_rho(qt_ens* e, double* rho) {
  return guard([&] {
    if (!e || !rho) throw Fail(QT_EINVAL, "null argument");
    const int d = (int)e->sys->h->dim, n = e->n;
    // split the trajectories so that about 1024 work groups share the sum (fixed split order)
    const int tiles = (d + 63) / 64;
    const int splits = std::max(1, std::min((1024 + tiles * tiles - 1) / (tiles * tiles), (n + 15) / 16));
    e->rhoDev.alloc((size_t)d * d);
    if (splits > 1) e->rhoPart.alloc((size_t)splits * d * d);
    hipCheck(qt::launch_ensemble_rho(e->psi.p, n, d, splits, e->rhoPart.p, e->rhoDev.p, nullptr), "ensemble rho");
    hipCheck(hipMemcpy(rho, e->rhoDev.p, (size_t)d * d * sizeof(double2), hipMemcpyDeviceToHost), "D2H");
  });
}

int qt_mcwf_state(qt_ens* e, int i, double* psi, double* t) {
  return guard([&] {
    if (!e) throw Fail(QT_EINVAL, "null argument");
    if (i < 0 || i >= e->n) throw Fail(QT_EINVAL, "trajectory index out of range");
    const long d = e->sys->h->dim;
    if (psi) hipCheck(hipMemcpy(psi, e->psi.p + (size_t)i * d, d * sizeof(double2), hipMemcpyDeviceToHost), "D2H");
    if (t) hipCheck(hipMemcpy(t, e->t.p + i, sizeof(double), hipMemcpyDeviceToHost), "D2H");
  });
}

int qt_mcwf_jumplog(qt_ens* e, int i, double* t, int* ch, int cap, int* n) {
  return guard([&] {
    if (!e) throw Fail(QT_EINVAL, "null argument");
    if (i < 0 || i >= e->n) throw Fail(QT_EINVAL, "trajectory index out of range");
    long long c[qt::kCounters];
    hipCheck(hipMemcpy(c, e->cnt.p + (size_t)i * qt::kCounters, sizeof(c), hipMemcpyDeviceToHost), "D2H");
    int nj = (int)std::min<long long>(c[qt::kJumps], e->p.jump_log_cap);
    int m = std::min(nj, cap);
    if (m > 0) {
      if (t) hipCheck(hipMemcpy(t, e->jt.p + (size_t)i * e->p.jump_log_cap, m * sizeof(double), hipMemcpyDeviceToHost), "D2H");
      if (ch) hipCheck(hipMemcpy(ch, e->jch.p + (size_t)i * e->p.jump_log_cap, m * sizeof(int), hipMemcpyDeviceToHost), "D2H");
    }
    if (n) *n = (int)c[qt::kJumps];
  });
}

int qt_mcwf_counters(qt_ens* e, long long* per, long long* totals) {
  return guard([&] {
    if (!e) throw Fail(QT_EINVAL, "null argument");
    std::vector<long long> c((size_t)e->n * qt::kCounters);
    hipCheck(hipMemcpy(c.data(), e->cnt.p, c.size() * sizeof(long long), hipMemcpyDeviceToHost), "D2H");
    if (per) std::memcpy(per, c.data(), c.size() * sizeof(long long));
    if (totals) {
      for (int k = 0; k < qt::kCounters; ++k) totals[k] = 0;
      for (int i = 0; i < e->n; ++i) for (int k = 0; k < qt::kCounters; ++k) totals[k] += c[(size_t)i * qt::kCounters + k];
    }
  });
}

// Checkpoint / resume (MCWF_Trajectory::stateIO, MCWF_Trajectory.h:134-147): psi, time,
// dtDid, dtTry, RNG position, counters and jump records of trajectories [first, first+count).
int qt_mcwf_get_states(qt_ens* e, int first, int count, qt_traj_state* st, double* psi, double* jt, int* jch) {
  return guard([&] {
    if (!e || !st || first < 0 || count < 0 || first + count > e->n) throw Fail(QT_EINVAL, "bad trajectory range");
    if (!count) return;
    const size_t d = e->sys->h->dim, cap = e->p.jump_log_cap, n = count;
    hipCheck(hipDeviceSynchronize(), "hipDeviceSynchronize");
    std::vector<double> t(n), dd(n), dt(n);
    std::vector<unsigned long long> ord(n);
    std::vector<long long> c(n * qt::kCounters);
    hipCheck(hipMemcpy(t.data(), e->t.p + first, n * sizeof(double), hipMemcpyDeviceToHost), "D2H");
    hipCheck(hipMemcpy(dd.data(), e->dtDid.p + first, n * sizeof(double), hipMemcpyDeviceToHost), "D2H");
    hipCheck(hipMemcpy(dt.data(), e->dtTry.p + first, n * sizeof(double), hipMemcpyDeviceToHost), "D2H");
    hipCheck(hipMemcpy(ord.data(), e->ord.p + first, n * sizeof(unsigned long long), hipMemcpyDeviceToHost), "D2H");
    hipCheck(hipMemcpy(c.data(), e->cnt.p + (size_t)first * qt::kCounters, c.size() * sizeof(long long),
                       hipMemcpyDeviceToHost), "D2H");
    for (size_t i = 0; i < n; ++i) {
      st[i].t = t[i]; st[i].dtDid = dd[i]; st[i].dtTry = dt[i]; st[i].ordinal = ord[i];
      for (int k = 0; k < qt::kCounters; ++k) st[i].counters[k] = c[i * qt::kCounters + k];
    }
    if (psi) hipCheck(hipMemcpy(psi, e->psi.p + (size_t)first * d, n * d * sizeof(double2), hipMemcpyDeviceToHost), "D2H");
    if (cap && jt) hipCheck(hipMemcpy(jt, e->jt.p + (size_t)first * cap, n * cap * sizeof(double), hipMemcpyDeviceToHost), "D2H");
    if (cap && jch) hipCheck(hipMemcpy(jch, e->jch.p + (size_t)first * cap, n * cap * sizeof(int), hipMemcpyDeviceToHost), "D2H");
  });
}

int qt_mcwf_set_states(qt_ens* e, int first, int count, const qt_traj_state* st, const double* psi, const double* jt,
                       const int* jch) {
  return guard([&] {
    if (!e || !st || !psi || first < 0 || count < 0 || first + count > e->n) throw Fail(QT_EINVAL, "bad trajectory range");
    if (!count) return;
    const size_t d = e->sys->h->dim, cap = e->p.jump_log_cap, n = count;
    std::vector<double> t(n), dd(n), dt(n);
    std::vector<unsigned long long> ord(n);
    std::vector<long long> c(n * qt::kCounters);
    for (size_t i = 0; i < n; ++i) {
      if (!std::isfinite(st[i].t) || !std::isfinite(st[i].dtDid) || !(st[i].dtTry > 0.) || !std::isfinite(st[i].dtTry))
        throw Fail(QT_EINVAL, "state with a non-finite time or a non-positive dtTry");
      t[i] = st[i].t; dd[i] = st[i].dtDid; dt[i] = st[i].dtTry; ord[i] = st[i].ordinal;
      for (int k = 0; k < qt::kCounters; ++k) c[i * qt::kCounters + k] = st[i].counters[k];
    }
    // the trajectories of an ensemble share one time (they all stop at each stream point): a load
    // of the whole ensemble sets it, a partial load must match it
    for (size_t i = 1; i < n; ++i)
      if (t[i] != t[0]) throw Fail(QT_EINVAL, "the trajectories of an ensemble share one time");
    const bool whole = first == 0 && count == e->n;
    if (!whole && t[0] != e->time) throw Fail(QT_EINVAL, "a partial load must match the ensemble's time");
    hipCheck(hipDeviceSynchronize(), "hipDeviceSynchronize");
    hipCheck(hipMemcpy(e->psi.p + (size_t)first * d, psi, n * d * sizeof(double2), hipMemcpyHostToDevice), "H2D");
    hipCheck(hipMemcpy(e->t.p + first, t.data(), n * sizeof(double), hipMemcpyHostToDevice), "H2D");
    // the picture origin coincides with the time at every saved instant (MCWF_Trajectory.h:140)
    hipCheck(hipMemcpy(e->t0.p + first, t.data(), n * sizeof(double), hipMemcpyHostToDevice), "H2D");
    hipCheck(hipMemcpy(e->dtDid.p + first, dd.data(), n * sizeof(double), hipMemcpyHostToDevice), "H2D");
    hipCheck(hipMemcpy(e->dtTry.p + first, dt.data(), n * sizeof(double), hipMemcpyHostToDevice), "H2D");
    hipCheck(hipMemcpy(e->ord.p + first, ord.data(), n * sizeof(unsigned long long), hipMemcpyHostToDevice), "H2D");
    hipCheck(hipMemcpy(e->cnt.p + (size_t)first * qt::kCounters, c.data(), c.size() * sizeof(long long),
                       hipMemcpyHostToDevice), "H2D");
    if (cap && jt) hipCheck(hipMemcpy(e->jt.p + (size_t)first * cap, jt, n * cap * sizeof(double), hipMemcpyHostToDevice), "H2D");
    if (cap && jch) hipCheck(hipMemcpy(e->jch.p + (size_t)first * cap, jch, n * cap * sizeof(int), hipMemcpyHostToDevice), "H2D");
    // a resumed ensemble starts without the failure of an earlier launch
    const int zero = 0;
    hipCheck(hipMemcpy(e->errflag.p, &zero, sizeof(int), hipMemcpyHostToDevice), "H2D");
    if (whole) e->time = t[0];
  });
}

int qt_mcwf_kernel_ms(qt_ens* e, double* total, int* launches) {
  if (!e) return QT_EINVAL;
  if (total) *total = e->kernelMs;
  if (launches) *launches = e->launches;
  return QT_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ mixed batch

namespace {

struct MixVariant { int nwb, rb, whb, rw, whw; };

std::vector<MixVariant> mix_variants() {
  int n = qt::mixed_variants(nullptr, 0);
  std::vector<int> t(5 * n);
  qt::mixed_variants(t.data(), n);
  std::vector<MixVariant> v;
  for (int i = 0; i < n; ++i) v.push_back({t[5 * i], t[5 * i + 1], t[5 * i + 2], t[5 * i + 3], t[5 * i + 4]});
  return v;
}

// Team kinds of the mixed launch: 0 = a whole work group, 1 = one wavefront, 2 = a quarter
// of the work group.  Can member (dim, width) run as `team` under variant V?  Register-cached
// ELL rows need width <= the variant's width; uncached ones read any width from the table.
bool team_fits(const MixVariant& V, int team, long dim, int width) {
  if (team == 1) return dim <= 64L * V.rw && (V.rw * V.whw > 24 || width <= V.whw);
  const int nw = team == 2 ? V.nwb / 4 : V.nwb;
  return nw >= 1 && dim <= 64L * nw * V.rb && (V.rb * V.whb > 24 || width <= V.whb);
}

// Default team: the smallest that holds the state (wavefront, then quarter, then work group);
// policy 1 leaves quarters out, policy 2 puts every member on whole work groups.
bool assign_teams(const MixVariant& V, const std::vector<qt_ens*>& mem, const int* req, std::vector<int>& team,
                  int policy = 0) {
  team.assign(mem.size(), 0);
  static const int order[3][3] = {{1, 2, 0}, {1, 0, -1}, {0, -1, -1}};
  for (size_t m = 0; m < mem.size(); ++m) {
    const qt::System& s = *mem[m]->sys->h;
    const int want = req ? req[m] : -1;
    if (want >= 0) {
      if (want > 2 || !team_fits(V, want, s.dim, s.H.width)) return false;
      team[m] = want;
      continue;
    }
    int got = -1;
    for (int k = 0; k < 3 && got < 0; ++k)
      if (order[policy][k] >= 0 && team_fits(V, order[policy][k], s.dim, s.H.width)) got = order[policy][k];
    if (got < 0) return false;
    team[m] = got;
  }
  return true;
}

}  // namespace

struct qt_batch_s {
  std::vector<qt_ens*> mem;
  int variant = 0;
  std::vector<int> team;
  bool f32 = false;
  bool td = false;              // some member runs in the interaction picture
  bool gsl = false;             // some member uses GSL's controller
  DevBuf<qt::MixDev> mix;
  DevBuf<int> queue;            // [3] queue heads, one per team kind
  qt::MixDev* host = nullptr;   // pinned staging for the descriptor upload
  // the launch is two kernels: work-group teams on the caller's stream (grid1), quarter and
  // wavefront teams on a stream of the batch's own (grid6), joined by events
  int grid1 = 0, grid6 = 0;
  int cus = 0;                  // compute units of the device (cached)
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipStream_t side = nullptr;
  hipEvent_t fork = nullptr, join = nullptr;
  double kernelMs = 0.;
  int launches = 0;
  ~qt_batch_s() {
    if (host) (void)hipHostFree(host);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (fork) (void)hipEventDestroy(fork);
    if (join) (void)hipEventDestroy(join);
    if (side) (void)hipStreamDestroy(side);
  }
  void choose(int v, const int* req, int policy = 0) {
    td = gsl = false;
    for (auto* e : mem) {
      td = td || e->sys->interactionPicture();
      gsl = gsl || e->p.controller == QT_CONTROLLER_GSL;
    }
    auto V = mix_variants();
    if (v >= (int)V.size()) throw Fail(QT_EINVAL, "unknown batch layout variant");
    if (v >= 0) {
      if (!assign_teams(V[v], mem, req, team, policy)) throw Fail(QT_ENOTSUP, "the requested layout does not fit every member");
      variant = v;
    } else {
      // the fitting variant that occupies the fewest wavefronts per launch: sum over members of
      // trajectories x wavefronts of their team (a dim-80 member takes one wavefront with two rows
      // per lane, against a four-wavefront quarter team with one row per lane); ties: the lowest
      // index
      int found = -1;
      double best = 0.;
      std::vector<int> t;
      for (size_t i = 0; i < V.size(); ++i) {
        if (!assign_teams(V[i], mem, req, t, policy)) continue;
        double cost = 0.;
        for (size_t m = 0; m < mem.size(); ++m)
          cost += (double)mem[m]->n * (t[m] == 1 ? 1 : t[m] == 2 ? V[i].nwb / 4 : V[i].nwb);
        if (found < 0 || cost < best) { found = (int)i; best = cost; team = t; }
      }
      if (found < 0) throw Fail(QT_ENOTSUP, "no compiled mixed-batch layout fits every member");
      variant = found;
    }
    const MixVariant& L = V[variant];
    if (!cus) {
      int dev = 0;
      hipCheck(hipGetDevice(&dev), "hipGetDevice");
      hipDeviceProp_t prop;
      hipCheck(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
      cus = prop.multiProcessorCount;
    }
    grid1 = cus * std::max(1, qt::mixed_blocks_per_cu(L.nwb, L.rb, L.whb, L.rw, L.whw, 1, f32, td, gsl));
    grid6 = cus * std::max(1, qt::mixed_blocks_per_cu(L.nwb, L.rb, L.whb, L.rw, L.whw, 6, f32, td, gsl));
  }
};

extern "C" {

int qt_batch_variants(int* table, int cap) { return qt::mixed_variants(table, cap); }

int qt_batch_create(qt_ens* const* members, int n_members, const qt_batch_params* p, qt_batch** out) {
  return guard([&] {
    if (!members || !out || n_members <= 0) throw Fail(QT_EINVAL, "bad arguments");
    if (n_members > qt::kMaxMembers) throw Fail(QT_ENOTSUP, "at most 16 members per batch");
    std::unique_ptr<qt_batch> b(new qt_batch);
    for (int m = 0; m < n_members; ++m) {
      if (!members[m]) throw Fail(QT_EINVAL, "null member");
      b->mem.push_back(members[m]);
    }
    b->f32 = members[0]->p.precision == QT_FP32_DERIV;
    for (auto* e : b->mem)
      if ((e->p.precision == QT_FP32_DERIV) != b->f32) throw Fail(QT_EINVAL, "members must share one precision");
    b->choose(p ? p->variant : -1, p ? p->team : nullptr);
    b->mix.alloc(1);
    b->queue.alloc(3);
    hipCheck(hipHostMalloc((void**)&b->host, sizeof(qt::MixDev), hipHostMallocDefault), "hipHostMalloc");
    hipCheck(hipEventCreate(&b->e0), "hipEventCreate");
    hipCheck(hipEventCreate(&b->e1), "hipEventCreate");
    *out = b.release();
  });
}

int qt_batch_destroy(qt_batch* b) { delete b; return QT_OK; }

int qt_batch_layout(const qt_batch* b, int* variant, int* team) {
  if (!b) return QT_EINVAL;
  if (variant) *variant = b->variant;
  if (team) for (size_t m = 0; m < b->mem.size(); ++m) team[m] = b->team[m];
  return QT_OK;
}

int qt_batch_kernel_ms(const qt_batch* b, double* total, int* launches) {
  if (!b) return QT_EINVAL;
  if (total) *total = b->kernelMs;
  if (launches) *launches = b->launches;
  return QT_OK;
}

}  // extern "C"

// One mixed launch: prepares every member, uploads the descriptor, runs the kernel between
// the batch's events.  Does not synchronise.
static void batch_launch(qt_batch* b, const double* const* intervals, const int* n_int, hipStream_t st) {
  const auto V = mix_variants();
  const MixVariant& L = V[b->variant];
  qt::MixDev& M = *b->host;
  hipCheck(hipStreamSynchronize(st), "hipStreamSynchronize");  // the staging buffer may still feed a copy
  std::memset(&M, 0, sizeof(M));
  M.n_members = (int)b->mem.size();
  M.queue = b->queue.p;
  int items[3] = {0, 0, 0};
  for (size_t m = 0; m < b->mem.size(); ++m) {
    qt_ens* e = b->mem[m];
    if (!intervals[m]) throw Fail(QT_EINVAL, "null interval list");
    M.E[m] = ens_prepare(e, intervals[m], n_int[m], st, true);
    M.E[m].chunk = 0; M.E[m].prog = nullptr;  // the mixed kernel runs whole trajectories
    hipCheck(qt::launch_zero_u64(M.E[m].acc, ens_acc_words(e, n_int[m]), st), "zero stream sums");
    M.S[m] = e->sys->dev;
    M.team[m] = b->team[m];
    M.first[m] = items[b->team[m]];
    items[b->team[m]] += e->n;
  }
  for (int k = 0; k < 3; ++k) M.items[k] = items[k];
  hipCheck(hipMemcpyAsync(b->mix.p, &M, sizeof(M), hipMemcpyHostToDevice, st), "batch descriptor H2D");
  hipCheck(qt::launch_fill_int(b->queue.p, 3, 0, st), "reset queues");
  // every resident work group of each kernel, also when the work would fit fewer: wavefront and
  // quarter items then spread over all compute units instead of filling a few.  The work-group
  // teams go first on st; the smaller teams run beside them on the batch's own stream (default
  // priority) and take compute units as the long-running work-group items drain.
  hipCheck(hipEventRecord(b->e0, st), "hipEventRecord");
  const bool big = items[0] > 0, small = items[1] + items[2] > 0;
  if (small && big) {  // the side stream starts once the descriptor and the queues are in place
    if (!b->side) hipCheck(hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking), "hipStreamCreateWithFlags");
    if (!b->fork) hipCheck(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming), "hipEventCreate");
    if (!b->join) hipCheck(hipEventCreateWithFlags(&b->join, hipEventDisableTiming), "hipEventCreate");
    hipCheck(hipEventRecord(b->fork, st), "hipEventRecord");
  }
  if (big)
    hipCheck(qt::launch_mixed(L.nwb, L.rb, L.whb, L.rw, L.whw, 1, b->f32, b->td, b->gsl, std::max(1, b->grid1), b->mix.p,
                              st), "mixed launch (work-group teams)");
  if (small && big) {
    hipCheck(hipStreamWaitEvent(b->side, b->fork, 0), "hipStreamWaitEvent");
    hipCheck(qt::launch_mixed(L.nwb, L.rb, L.whb, L.rw, L.whw, 6, b->f32, b->td, b->gsl, std::max(1, b->grid6), b->mix.p,
                              b->side), "mixed launch (quarter and wavefront teams)");
    hipCheck(hipEventRecord(b->join, b->side), "hipEventRecord");
    hipCheck(hipStreamWaitEvent(st, b->join, 0), "hipStreamWaitEvent");
  } else if (small) {
    hipCheck(qt::launch_mixed(L.nwb, L.rb, L.whb, L.rw, L.whw, 6, b->f32, b->td, b->gsl, std::max(1, b->grid6), b->mix.p,
                              st), "mixed launch (quarter and wavefront teams)");
  }
  hipCheck(hipEventRecord(b->e1, st), "hipEventRecord");
  for (size_t m = 0; m < b->mem.size(); ++m) {
    b->mem[m]->lastIntervals = n_int[m];
  }
}

// waits for a batched launch (both of its kernels: the side stream joins st)
static void batch_sync(qt_batch*, hipStream_t st) {
  hipCheck(hipStreamSynchronize(st), "hipStreamSynchronize");
}

extern "C" {

int qt_batch_advance(qt_batch* b, const double* const* intervals, const int* n_int, double* const* rows,
                     qt_stream stream) {
  return guard([&] {
    if (!b || !intervals || !n_int) throw Fail(QT_EINVAL, "null argument");
    hipStream_t st = (hipStream_t)stream;
    std::vector<double> t0(b->mem.size());
    for (size_t m = 0; m < b->mem.size(); ++m) t0[m] = b->mem[m]->time;
    batch_launch(b, intervals, n_int, st);
    batch_sync(b, st);
    for (size_t m = 0; m < b->mem.size(); ++m) {
      qt_ens* e = b->mem[m];
      for (int k = 0; k < n_int[m]; ++k) e->time += intervals[m][k];
      ens_rows(e, t0[m], intervals[m], n_int[m], rows ? rows[m] : nullptr, st);
    }
    float ms = 0.f;
    hipCheck(hipEventElapsedTime(&ms, b->e0, b->e1), "hipEventElapsedTime");
    b->kernelMs += ms; b->launches += 1;
  });
}

int qt_batch_autotune(qt_batch* b, double probe_span, int probe_traj, int* variant_out, double* ms_out, int ms_cap) {
  return guard([&] {
    if (!b || !(probe_span > 0) || probe_traj <= 0) throw Fail(QT_EINVAL, "bad arguments");
    long total = 0;
    for (auto* e : b->mem) total += e->n;
    const double f = std::min(1.0, (double)probe_traj / (double)total);
    // probe members: fresh ensembles of the same systems, parameters and initial states
    std::vector<std::unique_ptr<qt_ens, int (*)(qt_ens*)>> probes;
    std::vector<qt_ens*> pm;
    for (auto* e : b->mem) {
      const long d = e->sys->h->dim;
      std::vector<double2> psi0(d);
      hipCheck(hipMemcpy(psi0.data(), e->psiInit.p, d * sizeof(double2), hipMemcpyDeviceToHost), "D2H");
      qt_ens* pe = nullptr;
      const int n = std::max(1, (int)std::lround(e->n * f));
      int rc = qt_mcwf_create(e->sys, n, &e->p, reinterpret_cast<const double*>(psi0.data()), &pe);
      if (rc) throw Fail(rc, g_err);
      probes.emplace_back(pe, qt_mcwf_destroy);
      pm.push_back(pe);
    }
    const auto V = mix_variants();
    std::vector<int> teamSave = b->team;
    const int variantSave = b->variant;
    double best = 1e300;
    int bestV = -1;
    std::vector<int> bestTeam;
    int nc = 0;
    const double iv = probe_span;
    std::vector<const double*> ivs(pm.size(), &iv);
    std::vector<int> nis(pm.size(), 1);
    for (size_t v = 0; v < V.size(); ++v) {
      // candidate team assignments: smallest team that holds each state, no quarter teams,
      // everything on whole work groups (duplicates skipped)
      std::vector<std::vector<int>> seen;
      for (int policy = 0; policy < 3; ++policy) {
        std::vector<int> tm;
        if (!assign_teams(V[v], pm, nullptr, tm, policy)) continue;
        if (std::find(seen.begin(), seen.end(), tm) != seen.end()) continue;
        seen.push_back(tm);
        qt_batch probe;
        probe.mem = pm; probe.f32 = b->f32;
        probe.choose((int)v, tm.data());
        probe.mix.alloc(1); probe.queue.alloc(3);
        hipCheck(hipHostMalloc((void**)&probe.host, sizeof(qt::MixDev), hipHostMallocDefault), "hipHostMalloc");
        hipCheck(hipEventCreate(&probe.e0), "hipEventCreate");
        hipCheck(hipEventCreate(&probe.e1), "hipEventCreate");
        for (auto* e : pm) { int rc = qt_mcwf_reset(e, nullptr); if (rc) throw Fail(rc, g_err); }
        batch_launch(&probe, ivs.data(), nis.data(), nullptr);
        batch_sync(&probe, nullptr);
        float ms = 0.f;
        hipCheck(hipEventElapsedTime(&ms, probe.e0, probe.e1), "hipEventElapsedTime");
        if (ms_out && nc < ms_cap) ms_out[nc] = ms;
        ++nc;
        if (ms < best) { best = ms; bestV = (int)v; bestTeam = probe.team; }
      }
    }
    if (bestV < 0) { b->team = teamSave; b->variant = variantSave; throw Fail(QT_ENOTSUP, "no layout fits the members"); }
    b->choose(bestV, bestTeam.data());
    if (variant_out) *variant_out = bestV;
  });
}

}  // extern "C"

extern "C" {

int qt_master_create(qt_sys* const* systems, int n_points, const qt_master_params* p, const double* rho0,
                     qt_master** out) {
  return guard([&] {
    if (!rho0 || !out) throw Fail(QT_EINVAL, "bad arguments");
    check_master(systems, n_points, p);
    const qt::System& s0 = *systems[0]->h;
    const int d = (int)s0.dim;
    const int NB = (d + 15) / 16, DP = 16 * NB, KC = DP / 4, NN = DP * DP;
    const bool sparse = p->kernel == QT_MASTER_SPARSE || (p->kernel == QT_MASTER_AUTO && !qt::master_supported(NB));
    const int nj = (int)s0.J.size();
    int wj = 1;
    for (int k = 0; k < n_points; ++k)
      for (auto& j : systems[k]->h->J) wj = std::max(wj, j.width);
    std::unique_ptr<qt_master> m(new qt_master);
    m->n = n_points; m->d = d; m->nb = NB; m->p = *p;
    m->sys0 = systems[0];
    m->sparse = sparse;
    // H_nH/i A fragments: [point][NB][KC][64] {re, im}, H[16 r + l%16][4 kk + l/16]
    // per point: the part of H_nH/i that rotates with the Exact exponents (all of it in the
    // Schroedinger picture) as a dense matrix -- the MFMA product's A operand -- and the remainder,
    // entries applied beside the product with their own frequencies (master_entry_rotates)
    struct RestEntry { int r, c; qt::cd v, f; };
    std::vector<std::vector<qt::cd>> Hrot(n_points);
    std::vector<std::vector<RestEntry>> rest(n_points);
    std::vector<int> jc((size_t)n_points * std::max(1, nj) * wj * d, 0);
    std::vector<double2> jv((size_t)n_points * std::max(1, nj) * wj * d, make_double2(0., 0.));
    std::vector<double> dtTry(n_points);
    for (int k = 0; k < n_points; ++k) {
      const qt::System& s = *systems[k]->h;
      if (!sparse) Hrot[k].assign((size_t)DP * DP, 0.);
      for (int r = 0; r < d; ++r)
        for (int w = 0; w < s.H.width; ++w) {
          const size_t i = (size_t)r * s.H.width + w;
          if (s.H.vals[i] == qt::cd(0.)) continue;
          // the sparse kernel takes every entry with its own frequency
          if (!sparse && master_entry_rotates(s, r, i)) Hrot[k][(size_t)r * DP + s.H.cols[i]] += s.H.vals[i];
          else rest[k].push_back({r, s.H.cols[i], s.H.vals[i], s.H.freqs.empty() ? qt::cd(0.) : s.H.freqs[i]});
        }
      for (int mI = 0; mI < nj; ++mI) {
        const qt::Ell& e = s.J[mI];
        for (int w = 0; w < e.width; ++w)
          for (int r = 0; r < d; ++r) {
            const size_t o = (((size_t)k * nj + mI) * wj + w) * d + r;
            jc[o] = e.cols[(size_t)r * e.width + w];
            const qt::cd z = e.vals[(size_t)r * e.width + w];
            jv[o] = make_double2(z.real(), z.imag());
          }
      }
      dtTry[k] = p->dtInit > 0 ? p->dtInit : 1. / (10. * s.highestFrequency);
    }
    for (int k = 0; k < n_points; ++k) m->td = m->td || systems[k]->h->hasEx;
    auto widthOf = [&](const std::vector<RestEntry>& v) {
      std::vector<int> per(d, 0);
      int w = 0;
      for (auto& e : v) w = std::max(w, ++per[e.r]);
      return w;
    };
    // (one shared copy of the fragments with the per-point differences as remainder was measured
    // slower on the 512-point sweep, 591 k vs 650 k tries/s: the fragments are not what bounds it)
    const bool shared = false;
    m->sharedH = shared;
    const int nfrag = sparse ? 0 : shared ? 1 : n_points;
    std::vector<double> hf(std::max<size_t>(2, (size_t)nfrag * NB * KC * 64 * 2), 0.);
    static_assert(qt::kMasterMaxNB * 4 <= 32, "a row block's k-steps fit the mask's bits");
    std::vector<unsigned> hm(std::max<size_t>(1, (size_t)nfrag * NB), 0u);
    for (int k = 0; k < nfrag; ++k)
      for (int rb = 0; rb < NB; ++rb)
        for (int kk = 0; kk < KC; ++kk)
          for (int l = 0; l < 64; ++l) {
            const qt::cd z = Hrot[k][(size_t)(16 * rb + l % 16) * DP + 4 * kk + l / 16];
            const size_t o = ((((size_t)k * NB + rb) * KC + kk) * 64 + l) * 2;
            hf[o] = z.real(); hf[o + 1] = z.imag();
            if (z != qt::cd(0., 0.)) hm[(size_t)k * NB + rb] |= 1u << kk;
          }
    m->hfrag.upload(hf); m->hmask.upload(hm); m->jcol.upload(jc); m->jval.upload(jv);
    if (m->td) {  // the Exact frees' exponents per point (zero in the Schroedinger picture)
      std::vector<double2> z((size_t)n_points * d, make_double2(0., 0.));
      for (int k = 0; k < n_points; ++k) {
        const qt::System& sk = *systems[k]->h;
        if (!sk.hasEx) continue;
        for (int i = 0; i < d; ++i) z[(size_t)k * d + i] = make_double2(sk.exactZ[i].real(), sk.exactZ[i].imag());
      }
      m->exz.upload(z);
    }
    // the remainder as ELL [point][wr][d] with per-point frequency tables
    {
      int wr = 0, nrf = 1;
      std::vector<std::vector<qt::cd>> fl(n_points, std::vector<qt::cd>{qt::cd(0.)});
      for (int k = 0; k < n_points; ++k) {
        wr = std::max(wr, widthOf(rest[k]));
        for (auto& e : rest[k])
          if (std::find(fl[k].begin(), fl[k].end(), e.f) == fl[k].end()) fl[k].push_back(e.f);
        nrf = std::max(nrf, (int)fl[k].size());
      }
      if (nrf > qt::kMaxFreqs) throw Fail(QT_ENOTSUP, "interaction picture with more than 32 distinct frequencies");
      if (!m->td && nrf > 1) throw Fail(QT_EINVAL, "internal: a Schroedinger-picture remainder with frequencies");
      m->wr = wr;
      m->nrf = nrf;
      if (wr > 0) {
        std::vector<int> rc((size_t)n_points * wr * d, 0), rf((size_t)n_points * wr * d, 0);
        std::vector<double2> rv((size_t)n_points * wr * d, make_double2(0., 0.));
        std::vector<double2> rfq((size_t)n_points * nrf, make_double2(0., 0.));
        for (int k = 0; k < n_points; ++k) {
          std::vector<int> per(d, 0);
          for (auto& e : rest[k]) {
            const int a = per[e.r]++;
            const size_t o = ((size_t)k * wr + a) * d + e.r;
            rc[o] = e.c;
            rv[o] = make_double2(e.v.real(), e.v.imag());
            rf[o] = (int)(std::find(fl[k].begin(), fl[k].end(), e.f) - fl[k].begin());
          }
          for (size_t q = 0; q < fl[k].size(); ++q) rfq[(size_t)k * nrf + q] = make_double2(fl[k][q].real(), fl[k][q].imag());
        }
        m->rcol.upload(rc); m->rval.upload(rv); m->rfi.upload(rf); m->rfreq.upload(rfq);
      }
    }
    m->nj = nj; m->wj = wj;
    systems[0]->upload();  // observables of point 0
    // rho0 -> tile order: index (4 t + v) 64 + l holds (16 r + 4 v + l/16, 16 c + l%16), t = r NB + c
    std::vector<double2> x0(NN, make_double2(0., 0.)), xs((size_t)n_points * NN);
    for (int t = 0; t < NB * NB; ++t)
      for (int v = 0; v < 4; ++v)
        for (int l = 0; l < 64; ++l) {
          const int i = 16 * (t / NB) + 4 * v + l / 16, j = 16 * (t % NB) + l % 16;
          if (i < d && j < d) x0[(4 * t + v) * 64 + l] = make_double2(rho0[2 * ((size_t)i * d + j)], rho0[2 * ((size_t)i * d + j) + 1]);
        }
    for (int k = 0; k < n_points; ++k) std::copy(x0.begin(), x0.end(), xs.begin() + (size_t)k * NN);
    m->x.upload(xs);
    m->k.alloc((size_t)n_points * (sparse ? 9 : 7) * NN);
    m->t.upload(std::vector<double>(n_points, 0.));
    m->dtDid.upload(std::vector<double>(n_points, 0.));
    m->dtTry.upload(dtTry);
    m->done.upload(std::vector<int>(n_points, 0));
    m->tss.upload(std::vector<double>(n_points, -1.));
    m->cnt.upload(std::vector<long long>((size_t)n_points * qt::kCounters, 0));
    if (p->autostop_rep > 0) {
      const int no = (int)s0.O.size();
      m->asbuf.upload(std::vector<double>((size_t)n_points * (p->autostop_rep + 1) * std::max(1, no), 0.));
      // the stream point at t = 0 of trajectory::run was the handler's first call (Trajectory.h:576-589)
      std::vector<int> meta((size_t)n_points * 3, 0);
      for (int k = 0; k < n_points; ++k) meta[3 * (size_t)k] = 1;
      m->asmeta.upload(meta);
      std::vector<int> pv(std::max(1, no), -1), ps(std::max(1, no), 0);
      for (int q = 0; q < no; ++q) { pv[q] = s0.processVarOf[q]; ps[q] = s0.processSqrt[q]; }
      m->pvar.upload(pv);
      m->psqrt.upload(ps);
    }
    hipCheck(hipEventCreate(&m->e0), "hipEventCreate");
    hipCheck(hipEventCreate(&m->e1), "hipEventCreate");
    *out = m.release();
  });
}

int qt_master_destroy(qt_master* m) { delete m; return QT_OK; }

// qt_master_advance in two halves, so that a caller can have several sweeps (and other work)
// in flight before waiting: the launch half uploads the intervals, launches and queues the copies
// of the stream outputs into pinned host memory, all asynchronous on st; the finish half waits
// for st and turns the outputs into rows.
static void master_launch(qt_master* m, const double* intervals, int n_int, hipStream_t st, int dc = 0,
                          double t_end = 0.) {
    const qt::System& s0 = *m->sys0->h;
    const int no = (int)s0.O.size();
    m->intervals.alloc(n_int);
    hipCheck(hipMemcpyAsync(m->intervals.p, intervals, n_int * sizeof(double), hipMemcpyHostToDevice, st), "H2D");
    m->obs.alloc((size_t)m->n * n_int * std::max(1, no));
    m->dtd.alloc((size_t)m->n * n_int);
    m->dn.alloc((size_t)m->n * n_int);
    qt::MasterDev M{};
    M.n_points = m->n; M.dim = m->d; M.hfrag = m->hfrag.p; M.hmask = m->hmask.p;
    M.nj = m->nj; M.wj = m->wj; M.jcol = m->jcol.p; M.jval = m->jval.p;
    M.no = no; M.wo = m->sys0->dev.wo; M.ocol = m->sys0->dev.ocol; M.oval = m->sys0->dev.oval; M.oimag = m->sys0->dev.oimag; M.owid = m->sys0->dev.owid;
    M.x = m->x.p; M.k = m->k.p; M.t = m->t.p; M.dtDid = m->dtDid.p; M.dtTry = m->dtTry.p;
    M.done = m->done.p; M.tss = m->tss.p; M.counters = m->cnt.p;
    M.epsAbs = m->p.epsRel; M.epsRel = m->p.epsAbs;  // ODE.h:205
    M.ss_tol = m->p.ss_tol;
    M.as_rep = m->p.autostop_rep; M.as_eps = m->p.autostop_eps; M.as_epsAbs = m->p.autostop_eps_abs;
    M.as_buf = m->asbuf.p; M.as_meta = m->asmeta.p; M.proc_var = m->pvar.p; M.proc_sqrt = m->psqrt.p;
    M.intervals = m->intervals.p; M.n_intervals = n_int;
    M.obs_out = m->obs.p; M.dtdid_out = m->dtd.p; M.dnorm_out = m->dn.p;
    M.exz = m->td ? m->exz.p : nullptr;
    M.shared_h = m->sharedH ? 1 : 0;
    M.wr = m->wr; M.rcol = m->rcol.p; M.rval = m->rval.p; M.rfi = m->rfi.p; M.nrf = m->nrf; M.rfreq = m->rfreq.p;
    M.dc = dc; M.t_end = t_end;
    if (dc > 0) { m->tout.alloc((size_t)m->n * n_int); M.time_out = m->tout.p; }
    hipCheck(hipEventRecord(m->e0, st), "hipEventRecord");
    hipCheck(m->sparse ? qt::launch_master_sparse(M, st) : qt::launch_master(m->nb, M, st), "master launch");
    hipCheck(hipEventRecord(m->e1, st), "hipEventRecord");
    const size_t nob = (size_t)m->n * n_int * std::max(1, no), nd = (size_t)m->n * n_int;
    if (m->hcap < nob + 2 * nd) {
      if (m->hbuf) hipCheck(hipHostFree(m->hbuf), "hipHostFree");
      m->hbuf = nullptr; m->hcap = 0;
      hipCheck(hipHostMalloc((void**)&m->hbuf, (nob + 2 * nd) * sizeof(double), hipHostMallocDefault), "hipHostMalloc");
      m->hcap = nob + 2 * nd;
    }
    hipCheck(hipMemcpyAsync(m->hbuf, m->obs.p, nob * sizeof(double), hipMemcpyDeviceToHost, st), "D2H");
    hipCheck(hipMemcpyAsync(m->hbuf + nob, m->dtd.p, nd * sizeof(double), hipMemcpyDeviceToHost, st), "D2H");
    hipCheck(hipMemcpyAsync(m->hbuf + nob + nd, m->dn.p, nd * sizeof(double), hipMemcpyDeviceToHost, st), "D2H");
}

static void master_finish(qt_master* m, const double* intervals, int n_int, double* rows, double* dnorm, hipStream_t st) {
    hipCheck(hipStreamSynchronize(st), "hipStreamSynchronize");
    const qt::System& s0 = *m->sys0->h;
    const int no = (int)s0.O.size();
    const size_t nob = (size_t)m->n * n_int * std::max(1, no), nd = (size_t)m->n * n_int;
    const double* ob = m->hbuf;
    const double* dd = m->hbuf + nob;
    const double* dnv = m->hbuf + nob + nd;
    float ms = 0.f;
    hipCheck(hipEventElapsedTime(&ms, m->e0, m->e1), "hipEventElapsedTime");
    m->kernelMs += ms; m->launches += 1;
    const double tStart = m->time;
    for (int k = 0; k < n_int; ++k) m->time += intervals[k];
    for (int pt = 0; pt < m->n; ++pt) {
      double tt = tStart;
      for (int k = 0; k < n_int; ++k) {
        tt += intervals[k];
        if (rows) {
          double* r = rows + ((size_t)pt * n_int + k) * (2 + no);
          r[0] = tt; r[1] = dd[(size_t)pt * n_int + k];
          for (int q = 0; q < no; ++q) r[2 + q] = ob[((size_t)pt * n_int + k) * no + q];
          process_rows(s0, r + 2);
        }
        if (dnorm) dnorm[(size_t)pt * n_int + k] = dnv[(size_t)pt * n_int + k];
      }
    }
}

int qt_master_advance(qt_master* m, const double* intervals, int n_int, double* rows, double* dnorm, qt_stream stream) {
  return guard([&] {
    if (!m || !intervals || n_int <= 0) throw Fail(QT_EINVAL, "bad arguments");
    master_launch(m, intervals, n_int, (hipStream_t)stream);
    master_finish(m, intervals, n_int, rows, dnorm, (hipStream_t)stream);
  });
}

int qt_master_advance_dc(qt_master* m, double t_end, int dc, int max_rows, double* rows, int* n_rows, double* dnorm) {
  return guard([&] {
    if (!m || !rows || !n_rows || dc < 1 || max_rows < 1 || !std::isfinite(t_end)) throw Fail(QT_EINVAL, "bad arguments");
    if (m->n != 1) throw Fail(QT_EINVAL, "dc mode drives a single density operator (one sweep point)");
    if (t_end < m->time) throw Fail(QT_EINVAL, "t_end lies before the point's time");
    *n_rows = 0;
    if (m->time >= t_end) return;
    const int no = (int)m->sys0->h->O.size();
    std::vector<double> iv(max_rows, 0.);
    master_launch(m, iv.data(), max_rows, nullptr, dc, t_end);
    hipCheck(hipStreamSynchronize(nullptr), "hipStreamSynchronize");
    float ms = 0.f;
    hipCheck(hipEventElapsedTime(&ms, m->e0, m->e1), "hipEventElapsedTime");
    m->kernelMs += ms; m->launches += 1;
    std::vector<double> tt(max_rows);
    hipCheck(hipMemcpy(tt.data(), m->tout.p, tt.size() * sizeof(double), hipMemcpyDeviceToHost), "D2H");
    const size_t nob = (size_t)max_rows * std::max(1, no);
    const double* ob = m->hbuf;
    const double* dd = m->hbuf + nob;
    const double* dnv = m->hbuf + nob + max_rows;
    int n = 0;
    while (n < max_rows) {
      double* r = rows + (size_t)n * (2 + no);
      r[0] = tt[n]; r[1] = dd[n];
      for (int q = 0; q < no; ++q) r[2 + q] = ob[(size_t)n * no + q];
      for (int q = 0; q < no; ++q) if (!std::isfinite(r[2 + q])) throw Fail(QT_ENUMERIC, "Infinite detected in averages");
      process_rows(*m->sys0->h, r + 2);
      if (dnorm) dnorm[n] = dnv[n];
      ++n;
      if (tt[n - 1] >= t_end) break;
    }
    m->time = tt[n - 1];
    *n_rows = n;
  });
}

int qt_master_rho(qt_master* m, int point, double* rho) {
  return guard([&] {
    if (!m || point < 0 || point >= m->n || !rho) throw Fail(QT_EINVAL, "bad arguments");
    const int d = m->d, NB = m->nb, DP = 16 * NB, NN = DP * DP;
    std::vector<double2> x(NN);
    hipCheck(hipMemcpy(x.data(), m->x.p + (size_t)point * NN, NN * sizeof(double2), hipMemcpyDeviceToHost), "D2H");
    for (int t = 0; t < NB * NB; ++t)
      for (int v = 0; v < 4; ++v)
        for (int l = 0; l < 64; ++l) {
          const int i = 16 * (t / NB) + 4 * v + l / 16, j = 16 * (t % NB) + l % 16;
          if (i < d && j < d) { rho[2 * ((size_t)i * d + j)] = x[(4 * t + v) * 64 + l].x; rho[2 * ((size_t)i * d + j) + 1] = x[(4 * t + v) * 64 + l].y; }
        }
  });
}

// Checkpoint / resume of one sweep point (Master::stateIO, Master.h:89-102)
int qt_master_get_point(qt_master* m, int point, qt_master_state* st, double* rho) {
  return guard([&] {
    if (!m || !st || point < 0 || point >= m->n) throw Fail(QT_EINVAL, "bad arguments");
    hipCheck(hipDeviceSynchronize(), "hipDeviceSynchronize");
    hipCheck(hipMemcpy(&st->t, m->t.p + point, sizeof(double), hipMemcpyDeviceToHost), "D2H");
    hipCheck(hipMemcpy(&st->dtDid, m->dtDid.p + point, sizeof(double), hipMemcpyDeviceToHost), "D2H");
    hipCheck(hipMemcpy(&st->dtTry, m->dtTry.p + point, sizeof(double), hipMemcpyDeviceToHost), "D2H");
    hipCheck(hipMemcpy(st->counters, m->cnt.p + (size_t)point * qt::kCounters, qt::kCounters * sizeof(long long),
                       hipMemcpyDeviceToHost), "D2H");
    hipCheck(hipMemcpy(&st->done, m->done.p + point, sizeof(int), hipMemcpyDeviceToHost), "D2H");
    hipCheck(hipMemcpy(&st->t_ss, m->tss.p + point, sizeof(double), hipMemcpyDeviceToHost), "D2H");
    if (rho) {
      const int rc = qt_master_rho(m, point, rho);
      if (rc != QT_OK) throw Fail(rc, "reading rho");
    }
  });
}

int qt_master_set_point(qt_master* m, int point, const qt_master_state* st, const double* rho) {
  return guard([&] {
    if (!m || !st || !rho || point < 0 || point >= m->n) throw Fail(QT_EINVAL, "bad arguments");
    if (!std::isfinite(st->t) || !std::isfinite(st->dtDid) || !(st->dtTry > 0.) || !std::isfinite(st->dtTry))
      throw Fail(QT_EINVAL, "state with a non-finite time or a non-positive dtTry");
    hipCheck(hipDeviceSynchronize(), "hipDeviceSynchronize");
    // all points of a sweep share the stream times: point 0 sets the time, the others match it
    if (point != 0) {
      double t0 = 0.;
      hipCheck(hipMemcpy(&t0, m->t.p, sizeof(double), hipMemcpyDeviceToHost), "D2H");
      if (t0 != st->t) throw Fail(QT_EINVAL, "the points of a sweep share one time: load point 0 first");
    }
    const int d = m->d, NB = m->nb, DP = 16 * NB, NN = DP * DP;
    std::vector<double2> x(NN, make_double2(0., 0.));
    for (int t = 0; t < NB * NB; ++t)
      for (int v = 0; v < 4; ++v)
        for (int l = 0; l < 64; ++l) {
          const int i = 16 * (t / NB) + 4 * v + l / 16, j = 16 * (t % NB) + l % 16;
          if (i < d && j < d) x[(4 * t + v) * 64 + l] = make_double2(rho[2 * ((size_t)i * d + j)], rho[2 * ((size_t)i * d + j) + 1]);
        }
    hipCheck(hipMemcpy(m->x.p + (size_t)point * NN, x.data(), NN * sizeof(double2), hipMemcpyHostToDevice), "H2D");
    hipCheck(hipMemcpy(m->t.p + point, &st->t, sizeof(double), hipMemcpyHostToDevice), "H2D");
    hipCheck(hipMemcpy(m->dtDid.p + point, &st->dtDid, sizeof(double), hipMemcpyHostToDevice), "H2D");
    hipCheck(hipMemcpy(m->dtTry.p + point, &st->dtTry, sizeof(double), hipMemcpyHostToDevice), "H2D");
    hipCheck(hipMemcpy(m->cnt.p + (size_t)point * qt::kCounters, st->counters, qt::kCounters * sizeof(long long),
                       hipMemcpyHostToDevice), "H2D");
    hipCheck(hipMemcpy(m->done.p + point, &st->done, sizeof(int), hipMemcpyHostToDevice), "H2D");
    hipCheck(hipMemcpy(m->tss.p + point, &st->t_ss, sizeof(double), hipMemcpyHostToDevice), "H2D");
    if (m->p.autostop_rep > 0) {
      // a continuing run streams no row at its start: the handler starts afresh (Trajectory.h:583)
      const int zero[3] = {0, 0, 0};
      hipCheck(hipMemcpy(m->asmeta.p + 3 * (size_t)point, zero, sizeof(zero), hipMemcpyHostToDevice), "H2D");
    }
    if (point == 0) m->time = st->t;
  });
}

int qt_master_status(qt_master* m, int* done, double* tss, double* t, long long* counters) {
  return guard([&] {
    if (!m) throw Fail(QT_EINVAL, "null argument");
    if (done) hipCheck(hipMemcpy(done, m->done.p, m->n * sizeof(int), hipMemcpyDeviceToHost), "D2H");
    if (tss) hipCheck(hipMemcpy(tss, m->tss.p, m->n * sizeof(double), hipMemcpyDeviceToHost), "D2H");
    if (t) hipCheck(hipMemcpy(t, m->t.p, m->n * sizeof(double), hipMemcpyDeviceToHost), "D2H");
    if (counters) hipCheck(hipMemcpy(counters, m->cnt.p, (size_t)m->n * qt::kCounters * sizeof(long long), hipMemcpyDeviceToHost), "D2H");
  });
}

int qt_master_kernel_ms(qt_master* m, double* total, int* launches) {
  if (!m) return QT_EINVAL;
  if (total) *total = m->kernelMs;
  if (launches) *launches = m->launches;
  return QT_OK;
}

int qt_mcwf_plain_run(qt_sys* sys, const qt_mcwf_params* p, unsigned long long traj, const double* psi0,
                      const double* intervals, int n_int, double* obs_host, double* psi_out, double* jt, int* jch,
                      int jcap, long long* counters) {
  return guard([&] {
    if (!sys || !psi0 || !intervals || n_int <= 0) throw Fail(QT_EINVAL, "bad arguments");
    check_params(p);
    const qt::System& s = *sys->h;
    sys->checkPictureSupported();
    if ((int)s.J.size() > qt::kMaxJumps) throw Fail(QT_ENOTSUP, "too many jump channels");
    sys->upload();
    const bool td = sys->interactionPicture(), gsl = p->controller == QT_CONTROLLER_GSL;
    LayoutChoice L = pick_layout(s.dim, s.H.width, p->layout_rows, p->layout_waves);
    const long d = s.dim;
    const int nj = (int)s.J.size(), no = (int)s.O.size();
    hipStream_t st = nullptr;
    DevBuf<double2> x, xn, xc;
    DevBuf<double> err, post, obs;
    x.alloc(d); xn.alloc(d); xc.alloc(d); err.alloc(1); post.alloc(1 + std::max(1, nj)); obs.alloc(std::max(1, no));
    hipCheck(hipMemcpy(x.p, psi0, d * sizeof(double2), hipMemcpyHostToDevice), "H2D");
    std::vector<qt::cd> psi(d);
    for (long i = 0; i < d; ++i) psi[i] = qt::cd(psi0[2 * i], psi0[2 * i + 1]);
    double t = 0., t0 = 0., dtDid = 0., dtTry = p->dtInit > 0 ? p->dtInit : 1. / (10. * s.highestFrequency);
    const double epsAbs = p->epsRel, epsRel = p->epsAbs;  // ODE.h:205
    long long c[qt::kCounters] = {0};
    if (nj) {
      auto r = host_rates(s, psi);
      double tot = 0.; for (double v : r) tot += v;
      if (!std::isfinite(tot)) throw Fail(QT_ENUMERIC, "Infinite detected in averages");  // LiouvilleanAveragedCommon.h:82-93
      dtTry = std::min(dtTry, p->dpLimit / tot);
    }
    unsigned long long ord = 0;
    std::vector<double> hp(1 + std::max(1, nj));
    auto coherent = [&](double Dt) {
      const double timeBefore = t;
      const double nextDtTry = (std::fabs(Dt) < std::fabs(dtTry / 10.) ? dtTry : 0.);
      if ((Dt > 0.) != (dtTry > 0.)) dtTry = -dtDid;
      if (std::fabs(dtTry) > std::fabs(Dt)) dtTry = Dt;
      for (;;) {
        qt::PlainTry P{(int)d, x.p, xn.p, err.p, t, dtTry, epsAbs, epsRel, t0};
        hipCheck(qt::launch_plain_try(L.rows, L.nw, L.wh, td, gsl, sys->dev, P, st), "plain try");
        double e = 0.;
        hipCheck(hipMemcpy(&e, err.p, sizeof(double), hipMemcpyDeviceToHost), "D2H");
        c[qt::kTries]++; c[qt::kDerivs] += gsl ? 7 : 6;
        if (gsl) {
          // ControlledErrorStepperGSL::try_step (ODE_GSL.h:44-84): accepted, then
          // gsl_odeiv2_control_hadjust with the order 5 of GSL's rkck
          const double hOld = dtTry;
          t += dtTry;
          if (e > 1.1) dtTry = std::max(0.9 * std::pow(e, -1. / 5.), 0.2) * hOld;
          else if (e < 0.5) dtTry = std::min(std::max(0.9 * std::pow(e, -1. / 6.), 1.), 5.) * hOld;
          std::swap(x.p, xn.p);
          break;
        }
        if (e > 1.0) { dtTry *= std::max(9. / 10. * std::pow(e, -1. / (4 - 1)), 1. / 5.); c[qt::kFailed]++; continue; }
        t += dtTry;
        if (e < 0.5) { e = std::max(std::pow(5.0, -5.0), e); dtTry *= 9. / 10. * std::pow(e, -1. / 5); }
        std::swap(x.p, xn.p);
        break;
      }
      dtDid = t - timeBefore;
      if (nextDtTry != 0.) dtTry = nextDtTry;
      c[qt::kOdeSteps]++;
      // Exact::actWithU and t0 = t (MCWF_Trajectory.h:269-272), then renorm and the rates
      hipCheck(qt::launch_plain_post(L.rows, L.nw, L.wh, sys->dev, x.p, post.p, td ? t - t0 : 0., st), "plain post");
      if (td) t0 = t;
      hipCheck(hipMemcpy(hp.data(), post.p, (1 + nj) * sizeof(double), hipMemcpyDeviceToHost), "D2H");
    };
    double now = 0.;
    for (int k = 0; k < n_int; ++k) {
      const double endTime = now + intervals[k];
      for (;;) {
        const double Dt = endTime - t;
        if (Dt == 0.) break;
        hipCheck(hipMemcpy(xc.p, x.p, d * sizeof(double2), hipMemcpyDeviceToDevice), "D2D");
        const double tCache = t, dtDidCache = dtDid;
        coherent(Dt);
        if (nj) {
          for (;;) {
            double tot = 0.;
            for (int m = 0; m < nj; ++m) tot += hp[1 + m];
            const double sugg = p->dpLimit / tot;
            if (tot * dtDid > p->overshootTol * p->dpLimit) {
              t = tCache; dtDid = dtDidCache; dtTry = sugg; c[qt::kStepBacks]++;
              hipCheck(hipMemcpy(x.p, xc.p, d * sizeof(double2), hipMemcpyDeviceToDevice), "D2D");
              coherent(Dt);
              continue;
            } else if (tot * dtTry > p->dpLimit) c[qt::kOvershoots]++;
            dtTry = std::min(dtTry, sugg);
            break;
          }
          double random = qt::jump_uniform(p->seed, traj, ord++) / dtDid;
          int no_ = 0;
          for (; random > 0. && no_ != nj; random -= hp[1 + no_++]) ;
          if (random < 0.) {
            --no_;
            const double nf = std::sqrt(hp[1 + no_]);
            if (!std::isfinite(nf)) throw Fail(QT_ENUMERIC, "Infinite detected in MCWF_Trajectory::performJump");
            hipCheck(qt::launch_plain_jump(L.rows, L.nw, L.wh, sys->dev, x.p, no_, nf, st), "plain jump");
            if (jt && c[qt::kJumps] < jcap) jt[c[qt::kJumps]] = t;
            if (jch && c[qt::kJumps] < jcap) jch[c[qt::kJumps]] = no_;
            c[qt::kJumps]++;
          }
        }
        c[qt::kMcwfSteps]++;
      }
      now = endTime;
      if (obs_host && no) {
        hipCheck(qt::launch_plain_obs(L.rows, L.nw, L.wh, sys->dev, x.p, obs.p, st), "plain obs");
        hipCheck(hipMemcpy(obs_host + (size_t)k * no, obs.p, no * sizeof(double), hipMemcpyDeviceToHost), "D2H");
      }
    }
    if (psi_out) hipCheck(hipMemcpy(psi_out, x.p, d * sizeof(double2), hipMemcpyDeviceToHost), "D2H");
    if (counters) std::memcpy(counters, c, sizeof(c));
  });
}

}  // extern "C"

// ------------------------------------------------------------------ serving loop

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// stream intervals of trajectory::run in Dt mode (Trajectory.h:560-595): advance(min(Dt, T - t))
std::vector<double> dt_intervals(double T, double Dt) {
  std::vector<double> v;
  double t = 0.;
  while (t < T) { const double d = std::min(Dt, T - t); v.push_back(d); t = t + d; }
  return v;
}

struct Request {
  long long id = 0;
  qt_sys* sys = nullptr;
  int n = 0;
  std::vector<double> psi0;
  std::vector<double> iv;
  size_t next = 0;
  qt_mcwf_params p{};
  qt_ens* ens = nullptr;
  size_t bytes = 0;
  std::vector<double> rows;
  double tSubmit = 0., tAdmit = 0., tDone = 0.;
  long long tries = 0;
  // a Master-equation parameter sweep (qt_server_submit_sweep) instead of an MCWF ensemble
  bool sweep = false;
  std::vector<qt_sys*> systems;
  std::vector<double> rho0;
  qt_master_params mp{};
  qt_master* m = nullptr;
  hipStream_t st = nullptr;               // its own non-blocking stream while admitted
  std::vector<std::vector<double>> prow;  // per point, the streamed rows so far
  bool admitted = false;
  int status = QT_OK;                     // a failed request keeps its error code and message
  std::string error;
  ~Request() {
    if (ens) qt_mcwf_destroy(ens);
    if (m) qt_master_destroy(m);
    if (st) (void)hipStreamDestroy(st);
  }
};

// device bytes a request holds while admitted (qt_ens buffers for one chunk of `chunk` intervals)
size_t request_bytes(const qt::System& s, int n, int chunk, int jcap, double span) {
  const size_t d = s.dim, no = std::max<size_t>(1, s.O.size());
  const size_t cap = (size_t)std::min(2048.0, std::max(64.0, std::ceil(span / 0.02 * 1.25) + 8));
  size_t b = 2 * (size_t)n * d * 16;                  // psi, psiInit
  b += (size_t)n * (5 * 8 + qt::kCounters * 8);        // t, dtDid, dtTry, dtTryInit, ordinal, counters
  b += (size_t)n * jcap * 12;                          // jump log
  b += (size_t)qt::kAccStripes * chunk * (no + 1) * 16 + (size_t)n * no * 8 + (size_t)chunk * (no + 2) * 8;  // stream sums, rows
  b += s.J.empty() ? 0 : (size_t)n * cap * 8;          // jump draws
  return b;
}

}  // namespace

struct qt_server_s {
  qt_server_params P{};
  long long nextId = 1;
  std::deque<Request*> waiting;
  std::vector<Request*> active;
  std::map<long long, std::unique_ptr<Request>> all;
  size_t used = 0;
  qt_batch_s batch;
  long long steps = 0, tries = 0, finished = 0, failed = 0;
  double kernelMs = 0.;

  // Retires r with an error: its device work is waited for and its memory freed; the loop and
  // the other requests go on.  r must not be in `waiting` or `active` afterwards (the caller's job).
  void fail(Request* r, int code, const std::string& msg) {
    if (r->st) (void)hipStreamSynchronize(r->st);
    if (r->ens) { qt_mcwf_destroy(r->ens); r->ens = nullptr; }
    if (r->m) { qt_master_destroy(r->m); r->m = nullptr; }
    if (r->st) { (void)hipStreamDestroy(r->st); r->st = nullptr; }
    if (r->admitted) used -= r->bytes;
    r->admitted = false;
    r->status = code ? code : QT_EDEVICE;
    r->error = msg;
    r->rows.clear(); r->prow.clear();
    r->tDone = now_s();
    ++failed;
  }
};

extern "C" {

int qt_server_create(const qt_server_params* p, qt_server** out) {
  return guard([&] {
    if (!p || !out) throw Fail(QT_EINVAL, "null argument");
    if (p->max_active < 1 || p->max_active > qt::kMaxMembers) throw Fail(QT_EINVAL, "max_active must be 1..16");
    if (p->chunk_intervals < 1) throw Fail(QT_EINVAL, "chunk_intervals must be positive");
    if (p->hbm_budget == 0) throw Fail(QT_EINVAL, "zero memory budget");
    check_params(&p->mcwf);
    std::unique_ptr<qt_server> s(new qt_server);
    s->P = *p;
    s->batch.f32 = p->mcwf.precision == QT_FP32_DERIV;
    s->batch.mix.alloc(1);
    s->batch.queue.alloc(3);
    hipCheck(hipHostMalloc((void**)&s->batch.host, sizeof(qt::MixDev), hipHostMallocDefault), "hipHostMalloc");
    hipCheck(hipEventCreate(&s->batch.e0), "hipEventCreate");
    hipCheck(hipEventCreate(&s->batch.e1), "hipEventCreate");
    *out = s.release();
  });
}

int qt_server_destroy(qt_server* s) { delete s; return QT_OK; }

int qt_server_submit(qt_server* s, qt_sys* sys, int n_traj, const double* psi0, double T, double Dt,
                     unsigned long long seed, long long* id) {
  return guard([&] {
    if (!s || !sys || !psi0 || n_traj <= 0 || !(T > 0) || !(Dt > 0)) throw Fail(QT_EINVAL, "bad arguments");
    const qt::System& h = *sys->h;
    sys->checkPictureSupported();
    std::unique_ptr<Request> r(new Request);
    r->id = s->nextId++;
    r->sys = sys; r->n = n_traj;
    r->psi0.assign(psi0, psi0 + 2 * h.dim);
    r->iv = dt_intervals(T, Dt);
    r->p = s->P.mcwf;
    r->p.seed = seed; r->p.traj0 = 0;
    r->bytes = request_bytes(h, n_traj, s->P.chunk_intervals, std::max(1, r->p.jump_log_cap), Dt * s->P.chunk_intervals);
    if (r->bytes > s->P.hbm_budget) throw Fail(QT_ENOMEM, "request larger than the server's memory budget");
    // every compiled mixed layout must be able to take the system
    {
      auto V = mix_variants();
      bool ok = false;
      for (auto& v : V) ok = ok || team_fits(v, 1, h.dim, h.H.width) || team_fits(v, 0, h.dim, h.H.width);
      if (!ok) throw Fail(QT_ENOTSUP, "system does not fit the mixed-batch layouts");
    }
    r->tSubmit = now_s();
    if (id) *id = r->id;
    s->waiting.push_back(r.get());
    s->all[r->id] = std::move(r);
  });
}

int qt_server_submit_sweep(qt_server* s, qt_sys* const* systems, int n_points, const double* rho0, double T,
                           double Dt, const qt_master_params* p, long long* id) {
  return guard([&] {
    if (!s || !systems || !rho0 || !p || n_points <= 0 || !(T > 0) || !(Dt > 0)) throw Fail(QT_EINVAL, "bad arguments");
    // what qt_master_create would refuse is refused here, not at admission inside a step
    check_master(systems, n_points, p);
    const qt::System& h = *systems[0]->h;
    std::unique_ptr<Request> r(new Request);
    r->id = s->nextId++;
    r->sweep = true;
    r->sys = systems[0];
    r->systems.assign(systems, systems + n_points);
    r->n = n_points;
    r->rho0.assign(rho0, rho0 + 2 * (size_t)h.dim * h.dim);
    r->mp = *p;
    r->iv = dt_intervals(T, Dt);
    r->prow.resize(n_points);
    // device bytes while admitted: per point rho, the seven RK slots and the H fragments on the
    // padded 16 x 16 tile grid, plus a chunk of rows
    const size_t dp = ((size_t)h.dim + 15) / 16 * 16;
    r->bytes = (size_t)n_points * (10 * dp * dp * 16 + (size_t)s->P.chunk_intervals * (3 + h.O.size()) * 8);
    if (r->bytes > s->P.hbm_budget) throw Fail(QT_ENOMEM, "request larger than the server's memory budget");
    r->tSubmit = now_s();
    if (id) *id = r->id;
    s->waiting.push_back(r.get());
    s->all[r->id] = std::move(r);
  });
}

// One continuous-batching iteration: admit waiting requests in arrival order while slots and
// the memory budget allow (a request that does not fit lets later, smaller ones pass), advance
// every admitted request by its next chunk of stream intervals in ONE mixed launch, append the
// streamed rows, retire finished requests and free their memory.  A request that fails (at
// admission, or a non-finite state in its part of a launch) is retired with its error; the
// others go on.
int qt_server_step(qt_server* s, int* n_active, int* n_waiting) {
  return guard([&] {
    if (!s) throw Fail(QT_EINVAL, "null argument");
    auto code_of = [](const std::exception& ex) {
      const Fail* f = dynamic_cast<const Fail*>(&ex);
      return f ? f->code : QT_EINVAL;
    };
    for (auto it = s->waiting.begin(); it != s->waiting.end();) {
      Request* r = *it;
      if ((int)s->active.size() >= s->P.max_active) break;
      if (s->used + r->bytes > s->P.hbm_budget) { ++it; continue; }
      int rc = r->sweep ? qt_master_create(r->systems.data(), r->n, &r->mp, r->rho0.data(), &r->m)
                        : qt_mcwf_create(r->sys, r->n, &r->p, r->psi0.data(), &r->ens);
      it = s->waiting.erase(it);
      if (rc) { s->fail(r, rc, g_err); continue; }
      r->tAdmit = now_s();
      r->admitted = true;
      s->used += r->bytes;
      s->active.push_back(r);
    }
    if (!s->active.empty()) {
      std::vector<Request*> ens, sw;
      for (auto* r : s->active) (r->sweep ? sw : ens).push_back(r);
      std::vector<Request*> dead;  // failed during this step
      const size_t na = ens.size();
      std::vector<const double*> ivp(na);
      std::vector<int> ni(na);
      std::vector<double> t0(na);
      for (size_t m = 0; m < na; ++m) {
        Request* r = ens[m];
        ivp[m] = r->iv.data() + r->next;
        ni[m] = (int)std::min<size_t>(s->P.chunk_intervals, r->iv.size() - r->next);
        t0[m] = r->ens->time;
      }
      // the ensembles in one mixed launch (asynchronous), the sweeps beside it on their own stream
      bool launched = false;
      if (na) {
        try {
          s->batch.mem.clear();
          for (auto* r : ens) s->batch.mem.push_back(r->ens);
          s->batch.choose(-1, nullptr);
          batch_launch(&s->batch, ivp.data(), ni.data(), nullptr);
          launched = true;
        } catch (std::exception& ex) {
          (void)hipDeviceSynchronize();
          for (auto* r : ens) { s->fail(r, code_of(ex), ex.what()); dead.push_back(r); }
        }
      }
      // every sweep launched on its own non-blocking stream before anything is waited for
      std::vector<int> nks(sw.size(), 0);
      for (size_t i = 0; i < sw.size(); ++i) {
        Request* r = sw[i];
        try {
          if (!r->st) hipCheck(hipStreamCreateWithFlags(&r->st, hipStreamNonBlocking), "hipStreamCreateWithFlags");
          nks[i] = (int)std::min<size_t>(s->P.chunk_intervals, r->iv.size() - r->next);
          master_launch(r->m, r->iv.data() + r->next, nks[i], r->st);
        } catch (std::exception& ex) {
          s->fail(r, code_of(ex), ex.what());
          dead.push_back(r);
          nks[i] = 0;
        }
      }
      for (size_t i = 0; i < sw.size(); ++i) {
        Request* r = sw[i];
        if (!nks[i]) continue;
        try {
          const int nk = nks[i], no = (int)r->sys->h->O.size();
          std::vector<double> rows((size_t)r->n * nk * (2 + no));
          master_finish(r->m, r->iv.data() + r->next, nk, rows.data(), nullptr, r->st);
          for (double v : rows) if (!std::isfinite(v)) throw Fail(QT_ENUMERIC, "Infinite detected in averages");
          for (int k = 0; k < r->n; ++k)
            r->prow[k].insert(r->prow[k].end(), rows.begin() + (size_t)k * nk * (2 + no),
                              rows.begin() + (size_t)(k + 1) * nk * (2 + no));
          r->next += nk;
        } catch (std::exception& ex) {
          s->fail(r, code_of(ex), ex.what());
          dead.push_back(r);
        }
      }
      if (launched) {
        bool ok = true;
        try {
          batch_sync(&s->batch, nullptr);
        } catch (std::exception& ex) {  // the launch failed: every member is suspect
          for (auto* r : ens) { s->fail(r, code_of(ex), ex.what()); dead.push_back(r); }
          ok = false;
        }
        for (size_t m = 0; ok && m < na; ++m) {
          Request* r = ens[m];
          try {
            const int no = (int)r->sys->h->O.size();
            std::vector<double> rows((size_t)ni[m] * (2 + no));
            for (int k = 0; k < ni[m]; ++k) r->ens->time += ivp[m][k];
            ens_rows(r->ens, t0[m], ivp[m], ni[m], rows.data(), nullptr);
            r->rows.insert(r->rows.end(), rows.begin(), rows.end());
            r->next += ni[m];
          } catch (std::exception& ex) {
            s->fail(r, code_of(ex), ex.what());
            dead.push_back(r);
          }
        }
        if (ok) {
          float ms = 0.f;
          hipCheck(hipEventElapsedTime(&ms, s->batch.e0, s->batch.e1), "hipEventElapsedTime");
          s->kernelMs += ms;
        }
      }
      ++s->steps;
      // retire
      std::vector<Request*> keep;
      for (auto* r : s->active) {
        if (std::find(dead.begin(), dead.end(), r) != dead.end()) continue;
        if (r->next < r->iv.size()) { keep.push_back(r); continue; }
        try {
          if (r->sweep) {
            std::vector<int> done(r->n);
            std::vector<double> tss(r->n), tt(r->n);
            std::vector<long long> c((size_t)r->n * qt::kCounters);
            int rc = qt_master_status(r->m, done.data(), tss.data(), tt.data(), c.data());
            if (rc) throw Fail(rc, g_err);
            r->tries = 0;
            for (int k = 0; k < r->n; ++k) r->tries += c[(size_t)k * qt::kCounters + qt::kTries];
            double ms = 0.; int nl = 0;
            qt_master_kernel_ms(r->m, &ms, &nl);
            s->kernelMs += ms;
            qt_master_destroy(r->m);
            r->m = nullptr;
            if (r->st) hipCheck(hipStreamDestroy(r->st), "hipStreamDestroy");
            r->st = nullptr;
            for (auto& pr : r->prow) r->rows.insert(r->rows.end(), pr.begin(), pr.end());  // point-major
            r->prow.clear();
          } else {
            long long tot[qt::kCounters];
            int rc = qt_mcwf_counters(r->ens, nullptr, tot);
            if (rc) throw Fail(rc, g_err);
            r->tries = tot[qt::kTries];
            qt_mcwf_destroy(r->ens);
            r->ens = nullptr;
          }
        } catch (std::exception& ex) {
          s->fail(r, code_of(ex), ex.what());
          continue;
        }
        s->tries += r->tries;
        s->used -= r->bytes;
        r->admitted = false;
        r->tDone = now_s();
        ++s->finished;
      }
      s->active.swap(keep);
    }
    if (n_active) *n_active = (int)s->active.size();
    if (n_waiting) *n_waiting = (int)s->waiting.size();
  });
}

int qt_server_result(qt_server* s, long long id, double* rows, int cap_rows, int* n_rows, double* latency,
                     double* queue_wait, long long* tries) {
  return guard([&] {
    if (!s) throw Fail(QT_EINVAL, "null argument");
    auto it = s->all.find(id);
    if (it == s->all.end()) throw Fail(QT_EINVAL, "unknown request id");
    Request* r = it->second.get();
    if (r->tDone == 0.) throw Fail(QT_EINVAL, "request not finished");
    if (latency) *latency = r->tDone - r->tSubmit;
    if (r->status != QT_OK) {
      if (n_rows) *n_rows = 0;
      throw Fail(r->status, "request " + std::to_string(id) + " failed: " + r->error);
    }
    const int w = 2 + (int)r->sys->h->O.size();
    const int n = (int)(r->rows.size() / w);
    if (n_rows) *n_rows = n;
    if (rows) std::memcpy(rows, r->rows.data(), (size_t)std::min(n, cap_rows) * w * sizeof(double));
    if (latency) *latency = r->tDone - r->tSubmit;
    if (queue_wait) *queue_wait = r->tAdmit - r->tSubmit;
    if (tries) *tries = r->tries;
  });
}

int qt_server_release(qt_server* s, long long id) {
  return guard([&] {
    if (!s) throw Fail(QT_EINVAL, "null argument");
    auto it = s->all.find(id);
    if (it == s->all.end()) throw Fail(QT_EINVAL, "unknown request id");
    if (it->second->tDone == 0.) throw Fail(QT_EINVAL, "request not finished");
    s->all.erase(it);
  });
}

int qt_server_stats(const qt_server* s, qt_server_stats_t* st) {
  if (!s || !st) return QT_EINVAL;
  st->steps = s->steps;
  st->finished = s->finished;
  st->failed = s->failed;
  st->tries = s->tries;
  st->kernel_ms = s->kernelMs;
  st->active = (int)s->active.size();
  st->waiting = (int)s->waiting.size();
  st->bytes_in_use = s->used;
  return QT_OK;
}

}  // extern "C"
