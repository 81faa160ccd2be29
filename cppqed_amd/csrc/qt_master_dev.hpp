// qt_master_dev.hpp -- argument block of the batched Master-equation kernel.
#pragma once

#include <hip/hip_runtime.h>

namespace qt {

constexpr int kMaxAutostopAvr = 64;  // averages a point's autostop can watch

struct MasterDev {
  int n_points;
  int dim;               // true dimension d
  const double* hfrag;   // [point][NB][DP/4][64] x {re, im}: A fragments of H_nH/i
  int nj, wj;
  const int* jcol;       // [point][nj][wj][dim] ELL of the jump operators
  const double2* jval;
  int no, wo;            // observables (shared by all points)
  const int* ocol;       // [no][wo][dim]
  const double2* oval;
  const int* owid;       // [no]
  const int* oimag;
  double2* x;            // [point][DP*DP] tile order
  double2* k;            // [point][7][DP*DP]: six stage derivatives and the 5th-order proposal
  double* t;
  double* dtDid;
  double* dtTry;
  int* done;             // 1 once ||drho/dt||_F < ss_tol was seen at a stream point
  double* tss;           // time at which that happened
  long long* counters;   // [point][kCounters]
  double epsAbs, epsRel;
  double ss_tol;         // steady-state threshold on ||drho/dt||_F (<= 0: never stop)
  const double* intervals;
  int n_intervals;
  // autostop (AutostopHandlerGeneric, CPPQEDutils/Trajectory.h:179-222) on the processed averages
  int as_rep;            // queue length; 0: off
  double as_eps, as_epsAbs;
  double* as_buf;        // [point][(as_rep + 1) * no]: running average, then the queue as a ring
  int* as_meta;          // [point][3]: started, queue size, queue head
  const int* proc_var;   // [no]: index of the mean whose square is subtracted, or -1
  const int* proc_sqrt;  // [no]: square root after the subtraction
  double* obs_out;       // [point][n_intervals][no]
  double* dtdid_out;     // [point][n_intervals]
  double* dnorm_out;     // [point][n_intervals]: ||drho/dt||_F at the stream point
};

}  // namespace qt
