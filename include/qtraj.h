/* qtraj.h -- C ABI of the MI355X quantum-trajectory hot path.
 *
 * Plain C: opaque handles, plain pointers and sizes, int status (0 = ok),
 * thread-local error text from qt_last_error().  Complex arrays are
 * interleaved (re, im) doubles, row-major over the subsystem dimensions
 * (the reference's Blitz++ StateVectorLow / DensityOperatorLow layout).
 *
 * Each entry point names the reference interface it replaces
 * (paths relative to the vukics/cppqed repository).
 */
#ifndef QTRAJ_H
#define QTRAJ_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct qt_sys qt_sys;          /* a lowered quantum system (host tables + device copy) */
typedef struct qt_ens qt_ens;          /* a device-resident MCWF ensemble */
typedef struct qt_master qt_master;    /* a device-resident batch of density operators */
typedef void* qt_stream;               /* hipStream_t; NULL = the default stream */

enum { QT_OK = 0, QT_EINVAL = 1, QT_EDEVICE = 2, QT_ENOMEM = 3, QT_ENUMERIC = 4, QT_ENOTSUP = 5 };
enum { QT_FP64 = 0, QT_FP32_DERIV = 1 };
enum { QT_CONTROLLER_ODEINT = 0, QT_CONTROLLER_GSL = 1 };

const char* qt_last_error(void);

/* ---------------------------------------------------------------- systems */

/* Builds a system from an element spec (one statement per line or ';'; frees are the axes
 * 0, 1, ... in order, interactions name them):
 *   free mode cutoff=N delta=D kappa=K eta=E [etaI=..] [picture=Sch|IP|UIP]
 *   free qbit deltaA=D gamma=G eta=E [etaI=..] [picture=Sch|IP|UIP]
 *   free ptla gamma=G deltaA=D eta=E [etaI=..]
 *   free particle fin=F omrec=W [vclass=V kpart=K mf=sin|cos|plus|minus] [picture=Sch|IP]
 *   ia jc q m g=G [gI=..]              ia bs m1 m2 g=U [gI=..]          ia zz q1 q2 J=J
 *   ia pcav m p u0=U kcav=K mf=cos|sin|plus|minus [vclass=V]
 *   ia pcav_orth m p u0=U
 *   ia ptm m0 m1 p u0=U0 u1=U1 phi=PHI k0=K k1=K mf0=.. mf1=..
 *   system binary|composite
 * Replaces the element makers and composition of the reference:
 *   mode::make  CPPQEDelements/frees/Mode.tcc:40-60,   qbit::make  Qbit.cc:217-246,
 *   PumpedTwoLevelAtomSch  PumpedTwoLevelAtom.cc:35-49, (Pumped)Particle  Particle.cc:79-110, 210-252,
 *   jaynescummings::make  JaynesCummings_.h:85-90, CoupledModes  CoupledModes.cc:36-41,
 *   ParticleAlongCavity / ParticleOrthogonalToCavity  ParticleCavity.cc:26-140 (a pumped particle of
 *   another wavenumber than the cavity's: the chained product term, refused with only the particle
 *   in the interaction picture),
 *   ParticleTwoModes  ParticleTwoModes.cc:35-46 (equal wavenumbers); zz has no reference
 *   counterpart (Qbit_.h:68-70 declares sigmazop without defining it);
 *   binary::make  CPPQEDcore/composites/BinarySystem.cc:323-333, composite::make  Composite.h:363-367.
 * Elements outside this catalogue come in through qt_sys_build_terms.  The device copy of the
 * tables is made on first use. */
int qt_sys_build(const char* spec, qt_sys** out);
int qt_sys_destroy(qt_sys* sys);

/* A system from its operators, for elements outside the spec catalogue: the same lowering as
 * qt_sys_build, fed with the data of the reference's quantumoperator::Tridiagonal terms.
 *
 * qt_tridiag is one Tridiagonal<R> (CPPQEDcore/quantumoperator/Tridiagonal.h:71-287) acting on
 * the system axes axes[0..rank) -- the V of its Act<V...> (structure/Interaction.h), or the free's
 * own axis.  k[a] is axis a's off-diagonal offset (Tridiagonal::getDifferences).  diags[i],
 * i < 3^rank, is the diagonal Tridiagonal::get()(i): digit a of i in base 3 (axis 0 the most
 * significant, as the direct product numbers them, Tridiagonal.tcc:40-50) selects on axis a the
 * main diagonal (0, |n><n|), the lower one (1, |n+K><n|) or the upper one (2, |n><n+K|); the
 * array is row-major over those axes with extent dims[a] (digit 0) or dims[a] - k[a], as
 * interleaved (re, im) doubles, or NULL for an empty diagonal.  freqs (may be NULL) likewise
 * holds Tridiagonal::getFreqs(): an entry's value at picture time tau is diag exp(freq tau)
 * (Tridiagonal::propagate, Tridiagonal.tcc:154-163); the main diagonal never rotates.
 * qt_coo is a general sparse operator on axes (an observable that is no Tridiagonal, e.g. the
 * particle's position operator, or a product of Tridiagonals such as ParticleAlongCavity's chained
 * term, ParticleCavity.cc:113-140): entry e is vals[e] |rows[e]><cols[e]| with row-major local
 * indices over the axes' dimensions, rotating as vals[e] exp(freqs[e] tau) when freqs (interleaved
 * (re, im) [2 nnz]) is given, not at all when it is NULL. */
typedef struct {
  int rank;
  const int* axes;              /* [rank] */
  const int* k;                 /* [rank] */
  const double* const* diags;   /* [3^rank] */
  const double* const* freqs;   /* [3^rank] or NULL */
} qt_tridiag;

typedef struct {
  int rank;
  const int* axes;              /* [rank] */
  long long nnz;
  const long long* rows;        /* [nnz] */
  const long long* cols;        /* [nnz] */
  const double* vals;           /* [2 nnz] */
  const double* freqs;          /* [2 nnz] or NULL */
} qt_coo;

/* an operator: the sum of its terms (Tridiagonals first, in order, then the sparse ones) */
typedef struct {
  int n_tri;
  const qt_tridiag* tri;
  int n_coo;
  const qt_coo* coo;
} qt_operator;

typedef struct {
  int dim;
  /* Exact part (structure::Exact, Exact.h): U = exp(-exact[n] (t - t0)) on level n, interleaved
   * (re, im) [2 dim]; NULL: the free has none (Schroedinger picture) */
  const double* exact;
} qt_free_desc;

typedef struct {
  int n_frees;
  const qt_free_desc* frees;    /* the system's axes in order (binary / composite frees) */
  qt_operator h;                /* H_nH/i: every element's TridiagonalHamiltonian terms
                                   (TridiagonalHamiltonian.h:21-87), applied in this order */
  int n_jumps;
  const qt_operator* jumps;     /* J_m in the reference's channel order (frees first) */
  int n_avr;
  const qt_operator* avr;       /* raw linear averages <psi|O_q|psi> */
  const int* avr_imag;          /* [n_avr]: 1 streams Im<O_q>, 0 Re<O_q> */
  const int* avr_var_of;        /* [n_avr]: q' whose square is subtracted (VAR), or -1 */
  const int* avr_sqrt;          /* [n_avr]: square root after the subtraction (DEV) */
  double highest_frequency;     /* QuantumSystem::highestFrequency (DynamicsBase.cc:49-66) */
} qt_system_desc;

/* Builds a system from its operators (copied; the caller keeps its arrays).  Replaces the
 * element classes' construction of their Tridiagonals and their binary / composite assembly
 * (structure/Free.h, composites/BinarySystem.cc:323-333, Composite.h:363-367). */
int qt_sys_build_terms(const qt_system_desc* desc, qt_sys** out);

typedef struct {
  long dim;                 /* total Hilbert-space dimension */
  int rank;
  int dims[16];
  int n_jumps;              /* Liouvillean::nAvr<LA_Li>  (Liouvillean.h:54) */
  int n_avr;                /* Averaged::nAvr            (streamed columns after time, dtDid) */
  int h_width, j_width, o_width;
  int time_dependent;       /* 1 in interaction picture */
  double highest_frequency; /* QuantumSystem::highestFrequency, QuantumSystem.h:21-41 */
  long long h_nnz;          /* stored nonzeros of H_nH/i (merged) */
  long long m_terms;        /* complex multiply-adds of one H_nH/i application counted the reference's
                               way: stored entries of every term's operator times its slice count
                               (Tridiagonal<R>::apply over fullRange, Tridiagonal.tcc:220-263) */
  int h_offsets;            /* slots of H's constant-offset form (every slot one fixed col - row, the
                               device gathers then compute columns), 0 when H has none */
} qt_sys_info_t;
int qt_sys_info(const qt_sys* sys, qt_sys_info_t* info);

/* Host copies of the lowered tables (testing aid): *width receives the row width;
 * cols[dim*width] row-major, vals/freqs[2*dim*width] (any may be NULL).
 * which: 0 = H_nH/i, 1+m = J_m, 1+n_jumps+q = observable q, -1 = H_nH/i in its
 * constant-offset form (as the device holds it; QT_EINVAL when h_offsets is 0). */
int qt_sys_table(const qt_sys* sys, int which, int* width, int* cols, double* vals, double* freqs);
/* The lowered system's Exact exponents per basis state (the sum over the frees of their exact[n],
 * interleaved, [2*dim]; zeros without an Exact part) and the streaming rules of its averages (as
 * qt_system_desc's avr_imag, avr_var_of, avr_sqrt, [n_avr] each).  Any pointer may be NULL. */
int qt_sys_rules(const qt_sys* sys, double* exact_z, int* avr_imag, int* avr_var_of, int* avr_sqrt);

/* dpsidt += (H_nH/i) psi for n state vectors, device pointers, on `stream`.
 * Replaces structure::Hamiltonian<RANK>::addContribution(t, psi, dpsidt, t0)
 * (CPPQEDcore/structure/Hamiltonian.h:38-43) as dispatched by
 * binary::Hamiltonian::addContribution_v (BinarySystem.cc:194-207) /
 * composite (Composite.tcc:178-197).  In the interaction picture the entries are taken at
 * picture time t - t0 (TDH_Base::addContribution_v with Tridiagonal::propagate,
 * TridiagonalHamiltonian.h:42-47). */
int qt_apply_heff(qt_sys* sys, double t, double t0, const double* psi_dev, double* dpsidt_dev, int n,
                  int precision, qt_stream stream);

/* ------------------------------------------------------------ MCWF ensemble */

typedef struct {
  double epsRel, epsAbs;     /* user values: the error checker gets (abs=epsRel, rel=epsAbs), ODE.h:205 */
  double dpLimit, overshootTol;   /* mcwf::Pars, MCWF_Trajectory.h:28-47 */
  double dtInit;             /* <= 0: 1/(10*highest_frequency), Trajectory.h:126 */
  int precision;             /* QT_FP64 | QT_FP32_DERIV */
  int controller;            /* QT_CONTROLLER_ODEINT (ODE.h:189-223) or QT_CONTROLLER_GSL (ODE_GSL.h:44-84:
                                every step accepted, gsl_odeiv2_control_hadjust; every form but
                                the sub-wave layouts) */
  unsigned long long seed;   /* counter-RNG key word 0 */
  unsigned long long traj0;  /* RNG key word 1 of trajectory 0 (trajectory k uses traj0+k) */
  int jump_log_cap;          /* jump records kept per trajectory */
  int layout_rows, layout_waves;  /* 0 = pick automatically (autotuned table) */
  int layout_occupancy;      /* minimum waves per SIMD asked of the register allocator; 0 = table */
  int layout_group_lanes;    /* 16 | 32 | 64: several trajectories per wavefront (G lanes each); 0 = table */
  int sched_chunk;           /* stream intervals per scheduling chunk of the one-wavefront layouts: with
                                more trajectories than resident wavefronts a launch runs every
                                trajectory's chunk c before any chunk c+1.  0 = automatic (about
                                n_intervals/8), < 0 = whole trajectories.  Results are the same
                                either way. */
} qt_mcwf_params;

/* nTraj copies of psi0 (host) on one system: mcwf::makeEnsemble, MCWF_Trajectory.h:213-234,
 * including the constructor's initial dpLimit check (:114-117). */
int qt_mcwf_create(qt_sys* sys, int n_traj, const qt_mcwf_params* p, const double* psi0_host, qt_ens** out);
int qt_mcwf_destroy(qt_ens* ens);

/* Advances every trajectory through n_intervals stream intervals of the given lengths
 * (cppqedutils::advance, Trajectory.h:43-50, for each trajectory: Ensemble::advance,
 * StochasticTrajectory.h:220) in one persistent launch.  At each interval end the
 * ensemble-averaged streamed row is written to rows_host[k*(2+n_avr)]: time, mean dtDid,
 * processed averages (AverageTrajectories + DensityOperatorStreamer,
 * MCWF_Trajectory.h:406-420).  rows_host may be NULL. */
int qt_mcwf_advance(qt_ens* ens, const double* intervals, int n_intervals, double* rows_host, qt_stream stream);

/* dc mode of the run loop for a single trajectory (cppqedutils::run with dc > 0,
 * CPPQEDutils/Trajectory.h:555-600): MCWF steps MCWF_Trajectory::step(t_end - t)
 * (MCWF_Trajectory.h:347-374), a streamed row after every dc steps and at t_end, up to max_rows
 * rows in this call.  rows_host[r*(2+n_avr)] = time, dtDid, processed averages; *n_rows = rows
 * written (fewer than max_rows only when t_end was reached; 0 when the trajectory is already at
 * t_end).  An ensemble of one trajectory only: the reference's Ensemble has no step()
 * (Trajectory.h:317-323), so its dc mode is the Dt mode of qt_mcwf_advance. */
int qt_mcwf_advance_dc(qt_ens* ens, double t_end, int dc, int max_rows, double* rows_host, int* n_rows,
                       qt_stream stream);

/* Jump draws from the host instead of the counter RNG: trajectory i's uniforms in [0, 1) for its
 * next n_per MCWF steps, draws_host[i * n_per + j] for the step whose RNG ordinal is j past the
 * ordinal it has now -- e.g. the reference's pcg64 streams through libstdc++'s
 * generate_canonical (randomutils::Pars / incrementForNextStream, CPPQEDutils/Random.h:137-161;
 * MCWF_Trajectory::sampleRandom, MCWF_Trajectory.h:161, 191), so that the device replays the
 * reference's own runs.  Later advances (also after qt_mcwf_reset, which returns the ordinals to
 * 0) take their draws from this table; an advance in which a trajectory needs a draw beyond it
 * fails with QT_EINVAL.  n_per = 0 returns to the counter RNG.  Not for sub-wave layouts. */
int qt_mcwf_set_draws(qt_ens* ens, const double* draws_host, int n_per);

/* Allocation-free form of qt_mcwf_advance.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  The same kernels and results as qt_mcwf_advance (rows_host as there,
 * may be NULL), but every per-launch buffer -- the uploaded intervals, the stream sums and the
 * device rows -- lives in the library's static device scratch array of 8 MiB per GPU, and the
 * jump draws come from the counter RNG inside the kernel instead of a bulk draw buffer.  It calls
 * no hipMalloc / hipFree (the ensemble's own buffers are made by qt_mcwf_create).  Limit: the
 * stream sums take 256 n_intervals (n_avr + 1) bytes; beyond 8 MiB in all the call is refused
 * with QT_EINVAL.  Every call on one GPU shares that GPU's scratch, which keeps the last call's
 * contents (the stream sums are cleared by each call): calls that overlap on one GPU -- from
 * other threads or on other streams -- race.  Synchronises `stream`. */
int qt_mcwf_advance_noalloc(qt_ens* ens, const double* intervals, int n_intervals, double* rows_host, qt_stream stream);

/* Asynchronous form for serving/graph replay: no host synchronisation, no rows. */
int qt_mcwf_advance_async(qt_ens* ens, const double* intervals, int n_intervals, qt_stream stream);

/* Returns every trajectory to its state right after qt_mcwf_create (t = 0, psi0, initial dtTry,
 * RNG ordinal 0, counters 0), device-side, asynchronously on `stream`. */
int qt_mcwf_reset(qt_ens* ens, qt_stream stream);

/* Graph replay: with enable = 1 the launch sequence of an advance (jump draws, queue reset,
 * batched kernel) is captured once into a hipGraph and replayed by later advances whose launch
 * description (interval count, buffers, draw budget) is unchanged -- interval values are
 * uploaded before each replay.  enable = 0 returns to direct launches, -1 only queries.
 * captures / replays (may be NULL) report how often each happened. */
int qt_mcwf_set_graph(qt_ens* ens, int enable, long long* captures, long long* replays);

/* Per-trajectory raw (unprocessed) averages at the last streamed point: out[n_traj*n_avr]. */
int qt_mcwf_traj_averages(qt_ens* ens, double* out);
/* The ensemble's density operator, on request (AverageTrajectories<MCWF_Trajectory>,
 * MCWF_Trajectory.h:406-420, with StateVector::addTo, StateVector.h:173-180):
 * rho = (1/N) sum_k |psi_k><psi_k| of the trajectories' current states, formed on the device in FP64
 * (a fixed-order sum, so repeated calls agree bit for bit; the order differs from the reference's
 * sequential addTo at rounding level).  rho_host: d*d complex, row-major.  Synchronising. */
int qt_mcwf_rho(qt_ens* ens, double* rho_host);
/* State of trajectory i (host copy) and its time. */
int qt_mcwf_state(qt_ens* ens, int i, double* psi_host, double* t);
/* Jump record of trajectory i (MCWF_TrajectoryLogger jump list). */
int qt_mcwf_jumplog(qt_ens* ens, int i, double* t, int* ch, int cap, int* n);
/* Per-trajectory counters [n_traj][8]: MCWF steps, ODE steps, failed tries, derivative
 * calls, RKCK tries, step-backs, dpLimit overshoots, jumps (MCWF_TrajectoryLogger.h:59-101,
 * ODE.h:99-122).  totals[8] (may be NULL) receives their sums. */
int qt_mcwf_counters(qt_ens* ens, long long* per_traj, long long* totals);
/* Average duration (ms) of the batched kernel launches so far, from HIP events on the launch stream. */
int qt_mcwf_kernel_ms(qt_ens* ens, double* total_ms, int* launches);

/* The complete dynamical state of one trajectory: what MCWF_Trajectory::stateIO archives
 * (MCWF_Trajectory.h:134-147: array, time, ODE logger, dtDid, dtTry, random engine, MCWF logger). */
typedef struct {
  double t, dtDid, dtTry;
  unsigned long long ordinal;   /* RNG position: counter of the trajectory's next MCWF step draw */
  long long counters[8];        /* as qt_mcwf_counters; counters[7] = jumps so far */
} qt_traj_state;

/* Checkpoint of trajectories [first, first+count): st[count]; psi_host[count][dim] complex (may be
 * NULL); jt/jch[count][jump_log_cap] jump records (may be NULL; the first min(counters[7], cap) are
 * valid).  Waits for the device. */
int qt_mcwf_get_states(qt_ens* ens, int first, int count, qt_traj_state* st, double* psi_host, double* jt, int* jch);
/* Resume: overwrites trajectories [first, first+count) with archived states (the reading half of
 * MCWF_Trajectory::stateIO, including t0 = t, MCWF_Trajectory.h:136-143).  All share one time; a
 * load of the whole ensemble sets the ensemble's time, a partial one must match it.  jt/jch as
 * above (may be NULL: the jump records are left as they are). */
int qt_mcwf_set_states(qt_ens* ens, int first, int count, const qt_traj_state* st, const double* psi_host,
                       const double* jt, const int* jch);

/* ---------------------------------------------------------------- mixed batch */

/* A mixed batch advances several ensembles (each its own system, trajectory count, state and
 * parameters) in ONE persistent launch: the reference runs each ensemble's trajectories
 * serially (trajectory::Ensemble::advance, StochasticTrajectory.h:220) and separate ensembles
 * one after the other.  Each member runs one trajectory per work group (team 0), per quarter
 * of a work group (team 2) or per wavefront (team 1); the layout variant sets the work-group
 * width and the rows per lane of the teams.  All members share one precision; each keeps its
 * own controller (odeint's or GSL's) and picture (Schroedinger or interaction picture). */
typedef struct qt_batch_s qt_batch;

typedef struct {
  int variant;                 /* index into qt_batch_variants' table; -1: first that fits */
  int team[16];                /* per member: 0 work group, 1 wavefront, 2 quarter of a work group,
                                  -1 chosen by dimension (smallest team that holds the state) */
} qt_batch_params;

/* Compiled layout variants, 5 ints each: waves per work group, rows per lane (team 0), ELL
 * register width (team 0), rows per lane (team 1), ELL register width (team 1).  Returns the count. */
int qt_batch_variants(int* table, int cap);
/* members: up to 16 ensembles; p may be NULL (defaults) */
int qt_batch_create(qt_ens* const* members, int n_members, const qt_batch_params* p, qt_batch** out);
int qt_batch_destroy(qt_batch* b);
/* Chosen layout: *variant and team[n_members]. */
int qt_batch_layout(const qt_batch* b, int* variant, int* team);
/* Advances member m through intervals[m][0..n_intervals[m]) (each member as qt_mcwf_advance),
 * all members together: the work-group-team members in one persistent kernel on `stream`, the
 * quarter- and wavefront-team members in a second one on a stream of the batch's own that
 * `stream` then waits for; rows[m] (may be NULL) receives member m's ensemble rows. */
int qt_batch_advance(qt_batch* b, const double* const* intervals, const int* n_intervals, double* const* rows,
                     qt_stream stream);
/* Times every layout that fits the members on a probe (each member advanced by probe_span on a
 * copy of at most probe_traj trajectories) and keeps the fastest; *variant_out and
 * ms_out[n_candidates] (may be NULL) report the outcome.  Members' states are left unchanged. */
int qt_batch_autotune(qt_batch* b, double probe_span, int probe_traj, int* variant_out, double* ms_out, int ms_cap);
int qt_batch_kernel_ms(const qt_batch* b, double* total_ms, int* launches);

/* ---------------------------------------------------------------- serving loop */

/* A request loop over the mixed batch (continuous batching): requests -- MCWF ensembles of
 * different systems -- are submitted at any time; each qt_server_step admits waiting requests
 * (arrival order, smaller ones may pass one that does not fit) while member slots and the
 * device-memory budget allow, advances every admitted request by its next chunk of stream
 * intervals in ONE launch, and retires finished requests, freeing their memory.  Request state
 * stays resident on the device between steps.  The reference runs one trajectory driver at a
 * time (cppqedutils::run, Trajectory.h:462-610); it has no counterpart of this loop. */
typedef struct qt_server_s qt_server;

typedef struct {
  size_t hbm_budget;       /* device bytes the admitted requests may hold */
  int max_active;          /* requests advanced per launch, 1..16 */
  int chunk_intervals;     /* stream intervals each request advances per step */
  qt_mcwf_params mcwf;     /* ODE / MCWF parameters of every request (seed set per request) */
} qt_server_params;

typedef struct {
  long long steps, finished, tries;
  double kernel_ms;
  int active, waiting;
  size_t bytes_in_use;
  long long failed;        /* requests retired with an error (qt_server_result returns it) */
} qt_server_stats_t;

int qt_server_create(const qt_server_params* p, qt_server** out);
int qt_server_destroy(qt_server* s);
/* Queues an ensemble of n_traj trajectories of sys (either picture) from psi0_host, run to T
 * with stream interval Dt (trajectory::run Dt mode) and RNG key (seed, trajectory index), under
 * the server's ODE / MCWF parameters (controller included).  sys must outlive the request.
 * Returns the request id. */
int qt_server_submit(qt_server* s, qt_sys* sys, int n_traj, const double* psi0_host, double T, double Dt,
                     unsigned long long seed, long long* id);
/* One iteration of the loop (admit, one launch, retire); reports active and waiting counts.
 * A request that fails -- at admission (qt_mcwf_create / qt_master_create refuse it, e.g. a
 * non-finite initial rate) or in its part of a launch (a non-finite rate, jump norm or streamed
 * average, the reference's "Infinite detected ..." exceptions) -- is retired with that error and
 * its memory freed; the other requests and the loop go on. */
int qt_server_step(qt_server* s, int* n_active, int* n_waiting);
/* Rows (time, mean dtDid, processed averages) of a finished request, its latency (submit to
 * retirement, seconds), its wait before admission and its RKCK tries.  For a failed request it
 * returns the request's error code and message (latency still reported, *n_rows = 0). */
int qt_server_result(qt_server* s, long long id, double* rows, int cap_rows, int* n_rows, double* latency,
                     double* queue_wait, long long* tries);
/* Forgets a finished request. */
int qt_server_release(qt_server* s, long long id);
int qt_server_stats(const qt_server* s, qt_server_stats_t* stats);

/* ------------------------------------------------------- Master-equation sweep */

typedef struct {
  double epsRel, epsAbs;   /* user values, as for MCWF (tolerance swap of ODE.h:205 applied inside) */
  double dtInit;           /* <= 0: 1/(10*highest_frequency) of each point's system */
  double ss_tol;           /* stop a point once ||drho/dt||_F < ss_tol at a stream point (<= 0: never) */
  /* the reference's steady-state criterion, AutostopHandlerGeneric (CPPQEDutils/Trajectory.h:179-222,
   * --autoStopEpsilon / --autoStopEpsilonAbs / --autoStopRepetition): a point stops at the stream
   * point whose processed averages lie within autostop_eps (relative) of the moving average of
   * the previous autostop_rep rows; averages below autostop_eps_abs count as autostop_eps_abs.
   * autostop_rep = 0: off.  Either criterion stops a point (qt_master_status: done, t_ss). */
  double autostop_eps, autostop_eps_abs;
  int autostop_rep;
  int kernel;              /* QT_MASTER_AUTO (MFMA up to d = 96, sparse beyond), QT_MASTER_MFMA or
                              QT_MASTER_SPARSE (any d up to 4096) */
} qt_master_params;

enum { QT_MASTER_AUTO = 0, QT_MASTER_MFMA = 1, QT_MASTER_SPARSE = 2 };

/* A batch of density operators, point k evolving under systems[k] (same dimension and jump
 * structure), all from rho0_host (d*d complex, row-major): master::make + Master
 * (CPPQEDcore/quantumtrajectory/Master.h:130-136, 50-208) for every sweep point.  Up to d = 96
 * the (H_nH/i) rho products run on FP64 MFMA with the stage matrix in LDS; beyond (up to 4096,
 * or on request) a sparse kernel applies the superoperator from the ELL rows of H_nH/i and the
 * jump operators with every matrix in HBM.  Schroedinger picture, or the
 * unitary interaction picture (UIP; the reference turns IP into UIP for the Master,
 * QM_Picture.cc:45-47, and refuses non-unitary ones, Master.h:73): then each step ends with
 * rho <- U rho U^dagger (Master.h:186-193). */
int qt_master_create(qt_sys* const* systems, int n_points, const qt_master_params* p, const double* rho0_host,
                     qt_master** out);
int qt_master_destroy(qt_master* m);
/* Advances every point through the stream intervals (Master::step via trajectory::run's Dt
 * loop, Trajectory.h:560-595); rows_host[(point*n_intervals + k)*(2+n_avr)] = time, dtDid,
 * processed averages (DensityOperatorStreamer); dnorm_host[point*n_intervals + k] =
 * ||drho/dt||_F at that stream point.  Either may be NULL. */
int qt_master_advance(qt_master* m, const double* intervals, int n_intervals, double* rows_host, double* dnorm_host,
                      qt_stream stream);
/* dc mode of the run loop for one density operator (a sweep of one point; cppqedutils::run with
 * dc > 0, Trajectory.h:555-600): Master steps step(t_end - t) (Master.h:152-208), a row after every
 * dc steps and at t_end, at most max_rows in this call; rows as qt_master_advance's (time, dtDid,
 * processed averages), dnorm_host[r] = ||drho/dt||_F (may be NULL); *n_rows = rows written. */
int qt_master_advance_dc(qt_master* m, double t_end, int dc, int max_rows, double* rows_host, int* n_rows,
                         double* dnorm_host);
/* rho of one point (d*d complex, row-major) */
int qt_master_rho(qt_master* m, int point, double* rho_host);
/* per point: steady-state flag, the time it was reached, current time, counters[8] */
int qt_master_status(qt_master* m, int* done, double* t_ss, double* t, long long* counters);
int qt_master_kernel_ms(qt_master* m, double* total_ms, int* launches);

/* A parameter-sweep solve for the request loop: n_points Master-equation points (systems[k], the
 * same dimension and jump structure, all from rho0_host, d*d complex row-major; parameters as
 * qt_master_create) advanced to T in stream intervals of Dt, a chunk per qt_server_step, on its
 * own stream beside the step's mixed MCWF launch.  Admitted by slots and memory like the
 * ensembles.  qt_server_result's rows are point-major: [point][interval][2 + n_avr], as
 * qt_master_advance writes them.  No counterpart in the reference (its Master runs one point
 * per process, Master.h:50-208; sweeps are separate runs). */
int qt_server_submit_sweep(qt_server* s, qt_sys* const* systems, int n_points, const double* rho0_host, double T,
                           double Dt, const qt_master_params* p, long long* id);

/* One sweep point's dynamical state: what Master::stateIO archives (Master.h:89-102: array,
 * time, ODE logger, dtDid, dtTry) plus the sweep's steady-state flag. */
typedef struct {
  double t, dtDid, dtTry;
  long long counters[8];   /* as qt_master_status */
  int done;                /* steady state reached */
  double t_ss;             /* time it was reached (-1: not yet) */
} qt_master_state;

/* Checkpoint of one point; rho_host (d*d complex, row-major) may be NULL.  Waits for the device. */
int qt_master_get_point(qt_master* m, int point, qt_master_state* st, double* rho_host);
/* Resume: overwrites one point (the reading half of Master::stateIO).  The points of a sweep
 * share the stream times, so loading point 0 sets the sweep's time and the others must match it. */
int qt_master_set_point(qt_master* m, int point, const qt_master_state* st, const double* rho_host);

/* -------------------------------------------------------------- plain form */

/* The unbatched GPU form: one trajectory at a time, the MCWF/ODE control on the host
 * (as MCWF_Trajectory::step, MCWF_Trajectory.h:347-374), one device launch per RKCK
 * try and per post-step phase, FP64.  Advances trajectory psi through the intervals,
 * writing the raw averages after each interval into obs_host[k*n_avr] and the jump
 * record into jt/jch (cap jcap).  counters[8] as qt_mcwf_counters. */
int qt_mcwf_plain_run(qt_sys* sys, const qt_mcwf_params* p, unsigned long long traj, const double* psi0_host,
                      const double* intervals, int n_intervals, double* obs_host, double* psi_out,
                      double* jt, int* jch, int jcap, long long* counters);

#ifdef __cplusplus
}
#endif

#endif /* QTRAJ_H */
