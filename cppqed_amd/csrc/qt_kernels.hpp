// qt_kernels.hpp -- argument blocks shared by the HIP kernels and the host launchers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace qt {

constexpr int kMaxJumps = 8;      // jump channels handled by the batched kernel
constexpr int kCounters = 8;      // per-trajectory counters, see Counter
constexpr int kMaxFreqs = 32;     // distinct interaction-picture frequencies of one H_nH/i
constexpr int kAccStripes = 16;   // stream accumulators are striped by trajectory index (less contention)
constexpr int kMaxDict = 512;     // entries of the H table's coefficient dictionary (SysDev::hdict)
constexpr int kMaxDictSlots = 12; // ELL slots of a table in dictionary form

enum Counter {
  kMcwfSteps = 0,   // MCWF_Trajectory::step calls
  kOdeSteps = 1,    // successful ODE steps (including redone ones after a step-back)
  kFailed = 2,      // rejected RKCK tries
  kDerivs = 3,      // H_nH/i applications
  kTries = 4,       // RKCK try_step calls (accepted + rejected): the throughput unit
  kStepBacks = 5,   // dpLimit overshoot step-backs (MCWF_Trajectory.h:304-308)
  kOvershoots = 6,  // dtTry capped by dpLimit (MCWF_Trajectory.h:309-311)
  kJumps = 7,
};

// Device view of a lowered system (ELL tables, channel-major for jumps/observables).
struct SysDev {
  int dim;
  int wh;                  // H row width (padded)
  const int* hcol;         // [wh][dim] (column-major ELL: lanes read consecutive rows)
  const double2* hval;     // [wh][dim]
  const float2* hval32;    // [wh][dim]
  int nj, wj;
  const int* jcol;         // [nj][wj][dim]
  const double2* jval;
  // rate weights [nj][dim] when every J_m^dagger J_m is diagonal (each row of J_m has at most one
  // entry and no two rows share a column: a, sigma, ...): rate_m = sum_c jw[m][c] |psi_c|^2, the
  // reference's own form of the rates (Mode_.h:180-184); null otherwise
  const double* jw;
  int no, wo;
  const int* ocol;         // [no][wo][dim]
  const double2* oval;
  const int* oimag;        // [no]
  const int* owid;         // [no]: entries of observable q per row (<= wo)
  // interaction picture (nfreq = 0 in the Schroedinger picture): H entry (w, row) at time t is
  // hval * exp(freq[hfidx] (t - t0)); after each ODE step psi_row *= exp(-exz[row] (t - t0))
  int nfreq;
  const double2* freq;     // [nfreq], freq[0] = 0
  const int* hfidx;        // [wh][dim]
  const double2* exz;      // [dim] (null without an Exact part)
  // processing of the streamed averages (Averaged::process): average q becomes
  // av[q] - av[pvar[q]]^2 when pvar[q] >= 0, then its square root when psqrt[q]
  const int* pvar;         // [no]
  const int* psqrt;        // [no]
  // constant-offset form of the H table (Schroedinger picture): when non-null, slot w of every
  // row holds the entry at column row + hoff[w] (zero where the row has none, its hcol clamped
  // into range), so a gather may compute the column instead of loading it
  const int* hoff;         // [wh]
  // dictionary form of the constant-offset table (with hoff, when the table has at most
  // kMaxDictSlots slots, each slot takes at most 256 distinct values and all slots together at most
  // kMaxDict): slot w of row r holds hdict[hdbase[w] + hidx[w][r]] (hdict32: the same values in
  // FP32).  Null without one.
  int ndict;
  const double2* hdict;
  const float2* hdict32;
  const int* hdbase;           // [kMaxDictSlots]
  const unsigned char* hidx;   // [wh][dim]
};

// Device view of an MCWF ensemble and the per-launch work description.
struct EnsDev {
  int n_traj;
  double2* psi;                 // [n_traj][dim]
  double* t;                    // [n_traj]
  double* dtDid;
  double* dtTry;
  double* t0;                   // time at which the pictures coincide (interaction picture)
  unsigned long long* ordinal;  // RNG counter (MCWF step ordinal)
  unsigned long long seed;
  unsigned long long traj0;     // global index of trajectory 0 (RNG key)
  long long* counters;          // [n_traj][kCounters]
  double* jt;                   // [n_traj][jcap]
  int* jch;
  int jcap;
  double epsAbs, epsRel;        // error-checker constants (already in odeint's order)
  double dpLimit, overshootTol;
  // per launch
  const double* intervals;      // advance lengths (Trajectory.h:43-50 semantics)
  int n_intervals;
  unsigned long long* acc;      // [kAccStripes][n_intervals][no + 1][2]: fixed-point sums over the
                                // trajectories of each stream point's raw averages and dtDid (acc_add)
  double acc_max;               // |value| bound that keeps the sums in range: 2^43 / n_traj
  double* last_obs;             // [n_traj][no]: raw averages at the launch's last stream point
  int* work;                    // work-queue head (trajectories are claimed in order)
  int gsl;                      // 1: GSL's controller (kernel variant chosen on the host), 0: odeint's
  int* err_flag;                // bit 0: a non-finite jump norm; bit 1: non-finite rates or averages;
                                // bit 2: an average beyond the stream sums' range; bit 3: host draws ran out
  const double* draws;          // [n_traj][draw_cap]: jump draws for ordinals base + i (may be null), base =
                                // draw_base[traj] if draw_base, else the trajectory's ordinal at launch
  int draw_cap;
  const unsigned long long* draw_base;  // [n_traj] or null
  // dc mode (trajectory::run with dc > 0): each of the n_intervals stream points follows up to dc
  // MCWF steps step(t_end - t); 0: Dt mode (the intervals)
  int dc;
  double t_end;
  double* rows_out;             // [n_traj][n_intervals][2 + no] raw rows (time, dtDid, averages) or null
  int draws_strict;             // 1: draws are host-supplied, running out of them fails the launch (bit 3)
  // interleaved scheduling (one-wavefront engines): the work items are (chunk c, trajectory) in
  // chunk-major order, chunk c = stream intervals [c chunk, (c + 1) chunk); prog[traj] counts the
  // trajectory's finished chunks, and the claimer of chunk c waits until chunk c - 1 is done.
  // chunk = 0: one item per trajectory covering every interval of the launch
  int chunk;
  int* prog;                    // [n_traj]
};

// Mixed batch: up to kMaxMembers ensembles of different systems advanced by one launch.
constexpr int kMaxMembers = 16;

struct alignas(16) MixDev {
  int n_members;
  int items[3];                 // trajectories per team kind: 0 work group, 1 wavefront, 2 quarter work group
  int* queue;                   // [3] queue heads, one per team kind
  int team[kMaxMembers];        // team kind of member m
  int first[kMaxMembers];       // first item of member m in its kind's queue
  SysDev S[kMaxMembers];
  EnsDev E[kMaxMembers];
};
static_assert(sizeof(MixDev) % 16 == 0, "MixDev is copied to LDS in 16-byte words");

// One RKCK try of one trajectory (the plain form): x -> xn, err
struct PlainTry {
  int dim;
  const double2* x;
  double2* xn;
  double* err;       // [1]
  double t, dt;
  double epsAbs, epsRel;
  double t0;         // picture origin (interaction picture)
};

}  // namespace qt
