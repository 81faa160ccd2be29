// qt_master_common.hpp -- device pieces shared by the Master-equation kernels (the MFMA kernel
// for small dimensions and the sparse one beyond).
#pragma once

#include "qt_device.hpp"
#include "qt_kernels.hpp"
#include "qt_master_dev.hpp"

namespace qt {

// Cash-Karp tableau by stage index (stage s = 0..5 computes K_{s+1}): a_{s+1, i+1} and c_{s+1}
__device__ __forceinline__ double ck_a(int s, int i) {
  constexpr double A[6][5] = {{0., 0., 0., 0., 0.},
                              {CK::a21, 0., 0., 0., 0.},
                              {CK::a31, CK::a32, 0., 0., 0.},
                              {CK::a41, CK::a42, CK::a43, 0., 0.},
                              {CK::a51, CK::a52, CK::a53, CK::a54, 0.},
                              {CK::a61, CK::a62, CK::a63, CK::a64, CK::a65}};
  return A[s][i];
}
__device__ __forceinline__ double ck_c(int s) {
  constexpr double C[6] = {0., CK::c2, CK::c3, CK::c4, CK::c5, CK::c6};
  return C[s];
}

// AutostopHandlerGeneric::operator() (CPPQEDutils/Trajectory.h:187-214) for one point, on the
// processed stream row (Mode.cc:198-201 processing of the raw averages `raw`).  The first call
// only starts the running average; the queue then fills to as_rep rows; from there on a row
// within as_eps (relative, max over the averages) of the running average stops the point,
// otherwise it replaces the oldest row.  Returns 1 to stop.  One thread per point.
__device__ inline int autostop(const MasterDev& M, int p, const double* raw) {
  const int no = M.no, rep = M.as_rep;
  double buf[kMaxAutostopAvr];
  for (int q = 0; q < no; ++q) {
    double v = raw[q];
    const int vq = M.proc_var[q];
    if (vq >= 0) {
      v -= raw[vq] * raw[vq];
      if (M.proc_sqrt[q]) v = sqrt(v);
    }
    buf[q] = fabs(v) < M.as_epsAbs ? M.as_epsAbs : v;
  }
  int* meta = M.as_meta + 3 * (size_t)p;
  double* avg = M.as_buf + (size_t)p * (rep + 1) * no;
  double* ring = avg + no;
  if (!meta[0]) {
    meta[0] = 1;
    for (int q = 0; q < no; ++q) avg[q] = 0.;
    return 0;
  }
  int qs = meta[1], qh = meta[2];
  if (qs == rep) {
    double m = -INFINITY;
    for (int q = 0; q < no; ++q) {
      const double r = fabs(avg[q] - buf[q]) / (fabs(avg[q]) + fabs(buf[q]));
      m = (m < r) ? r : m;
    }
    if (m < M.as_eps) return 1;
    double* front = ring + (size_t)qh * no;
    for (int q = 0; q < no; ++q) { avg[q] = avg[q] + (buf[q] - front[q]) / (double)rep; front[q] = buf[q]; }
    qh = (qh + 1) % rep;
  } else {
    double* slot = ring + (size_t)((qh + qs) % rep) * no;
    for (int q = 0; q < no; ++q) { avg[q] = ((double)qs * avg[q] + buf[q]) / (double)(qs + 1); slot[q] = buf[q]; }
    ++qs;
  }
  meta[1] = qs; meta[2] = qh;
  return 0;
}

}  // namespace qt
