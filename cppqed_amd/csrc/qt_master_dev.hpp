// qt_master_dev.hpp -- argument block of the batched Master-equation kernel.
#pragma once

#include <hip/hip_runtime.h>

namespace qt {

constexpr int kMaxAutostopAvr = 64;  // averages a point's autostop can watch
constexpr int kMasterMaxNB = 6;      // the MFMA kernel's largest padded dimension, in 16-row blocks

struct MasterDev {
  int n_points;
  int dim;               // true dimension d
  const double* hfrag;   // [point][NB][DP/4][64] x {re, im}: A fragments of H_nH/i ([1][...] when shared_h)
  int shared_h;          // 1: every point's product uses the one copy (point 0's); its differences are remainder
  const unsigned* hmask; // [point][NB] ([1][NB] when shared_h): bit kk set when the A fragment of row block
                         // rb and k-step kk holds a non-zero entry (the product skips the others)
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
  // interaction picture (the unitary one, Master.h:73, QM_Picture.cc:45-47): H_nH/i at picture time
  // tau is E (H_nH/i) E^-1 with E = diag(exp(exz tau)) (an entry |r><c| rotates with exz[r] - exz[c],
  // the Exact frees' frequencies), and after each step rho <- U rho U^dagger with
  // U = diag(exp(-exz (t - t0))) (Master.h:186-193); null in the Schroedinger picture
  const double2* exz;    // [point][dim]
  // entries of H_nH/i that do not rotate so (a term of an Exact free with frequencies the reference
  // leaves out, Tridiagonal.tcc:61-67): a sparse remainder applied beside the product, entry w of
  // row i is rval * exp(rfreq[rfi] tau) at column rcol
  int wr;                // remainder width (0: none)
  const int* rcol;       // [point][wr][dim]
  const double2* rval;
  const int* rfi;
  int nrf;               // distinct remainder frequencies per point (<= kMaxFreqs)
  const double2* rfreq;  // [point][nrf]
  // dc mode (trajectory::run with dc > 0): each stream point follows up to dc Master steps toward t_end
  int dc;
  double t_end;
  double* time_out;      // [point][n_intervals]: the point's time at each stream point
  double* obs_out;       // [point][n_intervals][no]
  double* dtdid_out;     // [point][n_intervals]
  double* dnorm_out;     // [point][n_intervals]: ||drho/dt||_F at the stream point
};

}  // namespace qt
