"""cppqed_amd -- MI355X-native quantum-trajectory hot path (MCWF ensembles, Master sweeps).

Thin Python mirror of the reference's element / trajectory-driver interface
(C++QED: ``mode::make``, ``qbit::make``, ``jaynescummings::make``,
``binary::make``, ``composite::make``, ``evolve``), over the C ABI in
``include/qtraj.h`` implemented by ``libqtraj.so`` (HIP for gfx950 + host C++).
The library must be built (``python -m cppqed_amd.build``); nothing here falls
back to a CPU path.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "libqtraj.so")

QT_FP64, QT_FP32_DERIV = 0, 1
QT_CONTROLLER_ODEINT, QT_CONTROLLER_GSL = 0, 1
QT_OK, QT_EINVAL, QT_EDEVICE, QT_ENOMEM, QT_ENUMERIC, QT_ENOTSUP = 0, 1, 2, 3, 4, 5

_lib = None


class QtError(RuntimeError):
    """An error of the C ABI: .code is the status (QT_EINVAL, QT_ENUMERIC, ...)."""

    def __init__(self, msg, code=0):
        super().__init__(msg)
        self.code = code


class SysInfo(C.Structure):
    _fields_ = [("dim", C.c_long), ("rank", C.c_int), ("dims", C.c_int * 16), ("n_jumps", C.c_int),
                ("n_avr", C.c_int), ("h_width", C.c_int), ("j_width", C.c_int), ("o_width", C.c_int),
                ("time_dependent", C.c_int), ("highest_frequency", C.c_double), ("h_nnz", C.c_longlong),
                ("m_terms", C.c_longlong), ("h_offsets", C.c_int)]


class McwfParams(C.Structure):
    _fields_ = [("epsRel", C.c_double), ("epsAbs", C.c_double), ("dpLimit", C.c_double),
                ("overshootTol", C.c_double), ("dtInit", C.c_double), ("precision", C.c_int),
                ("controller", C.c_int), ("seed", C.c_ulonglong), ("traj0", C.c_ulonglong),
                ("jump_log_cap", C.c_int), ("layout_rows", C.c_int), ("layout_waves", C.c_int),
                ("layout_occupancy", C.c_int), ("layout_group_lanes", C.c_int), ("sched_chunk", C.c_int)]


class MasterParams(C.Structure):
    _fields_ = [("epsRel", C.c_double), ("epsAbs", C.c_double), ("dtInit", C.c_double), ("ss_tol", C.c_double),
                ("autostop_eps", C.c_double), ("autostop_eps_abs", C.c_double), ("autostop_rep", C.c_int),
                ("kernel", C.c_int)]


MASTER_KERNELS = {"auto": 0, "mfma": 1, "sparse": 2}   # QT_MASTER_AUTO / _MFMA / _SPARSE


class BatchParams(C.Structure):
    _fields_ = [("variant", C.c_int), ("team", C.c_int * 16)]


class ServerParams(C.Structure):
    _fields_ = [("hbm_budget", C.c_size_t), ("max_active", C.c_int), ("chunk_intervals", C.c_int),
                ("mcwf", McwfParams)]


class ServerStats(C.Structure):
    _fields_ = [("steps", C.c_longlong), ("finished", C.c_longlong), ("tries", C.c_longlong),
                ("kernel_ms", C.c_double), ("active", C.c_int), ("waiting", C.c_int), ("bytes_in_use", C.c_size_t),
                ("failed", C.c_longlong)]


class TrajState(C.Structure):
    _fields_ = [("t", C.c_double), ("dtDid", C.c_double), ("dtTry", C.c_double), ("ordinal", C.c_ulonglong),
                ("counters", C.c_longlong * 8)]


class MasterState(C.Structure):
    _fields_ = [("t", C.c_double), ("dtDid", C.c_double), ("dtTry", C.c_double), ("counters", C.c_longlong * 8),
                ("done", C.c_int), ("t_ss", C.c_double)]


# counter slots (qt_mcwf_counters)
C_MCWF, C_ODE, C_FAILED, C_DERIVS, C_TRIES, C_STEPBACKS, C_OVERSHOOTS, C_JUMPS = range(8)


def lib():
    """Loads libqtraj.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise QtError("libqtraj.so is not built (python -m cppqed_amd.build)")
        l = C.CDLL(LIBPATH)
        vp, ip, dp = C.c_void_p, C.c_int, C.c_double
        l.qt_last_error.restype = C.c_char_p
        l.qt_sys_build.argtypes = [C.c_char_p, C.POINTER(vp)]
        l.qt_sys_destroy.argtypes = [vp]
        l.qt_sys_info.argtypes = [vp, C.POINTER(SysInfo)]
        l.qt_sys_table.argtypes = [vp, ip, C.POINTER(C.c_int), vp, vp, vp]
        l.qt_sys_rules.argtypes = [vp, vp, vp, vp, vp]
        l.qt_apply_heff.argtypes = [vp, dp, dp, vp, vp, ip, ip, vp]
        l.qt_mcwf_create.argtypes = [vp, ip, C.POINTER(McwfParams), vp, C.POINTER(vp)]
        l.qt_mcwf_destroy.argtypes = [vp]
        l.qt_mcwf_advance.argtypes = [vp, vp, ip, vp, vp]
        l.qt_mcwf_advance_async.argtypes = [vp, vp, ip, vp]
        l.qt_mcwf_advance_noalloc.argtypes = [vp, vp, ip, vp, vp]
        l.qt_mcwf_set_draws.argtypes = [vp, vp, ip]
        l.qt_mcwf_advance_dc.argtypes = [vp, dp, ip, ip, vp, C.POINTER(C.c_int), vp]
        l.qt_mcwf_reset.argtypes = [vp, vp]
        l.qt_mcwf_set_graph.argtypes = [vp, ip, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
        l.qt_mcwf_traj_averages.argtypes = [vp, vp]
        l.qt_mcwf_rho.argtypes = [vp, vp]
        l.qt_mcwf_state.argtypes = [vp, ip, vp, vp]
        l.qt_mcwf_jumplog.argtypes = [vp, ip, vp, vp, ip, C.POINTER(C.c_int)]
        l.qt_mcwf_counters.argtypes = [vp, vp, vp]
        l.qt_mcwf_kernel_ms.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_int)]
        l.qt_mcwf_get_states.argtypes = [vp, ip, ip, C.POINTER(TrajState), vp, vp, vp]
        l.qt_mcwf_set_states.argtypes = [vp, ip, ip, C.POINTER(TrajState), vp, vp, vp]
        l.qt_master_get_point.argtypes = [vp, ip, C.POINTER(MasterState), vp]
        l.qt_master_set_point.argtypes = [vp, ip, C.POINTER(MasterState), vp]
        l.qt_mcwf_plain_run.argtypes = [vp, C.POINTER(McwfParams), C.c_ulonglong, vp, vp, ip, vp, vp, vp, vp, ip, vp]
        l.qt_batch_variants.argtypes = [vp, ip]
        l.qt_batch_create.argtypes = [vp, ip, C.POINTER(BatchParams), C.POINTER(vp)]
        l.qt_batch_destroy.argtypes = [vp]
        l.qt_batch_layout.argtypes = [vp, C.POINTER(C.c_int), vp]
        l.qt_batch_advance.argtypes = [vp, vp, vp, vp, vp]
        l.qt_batch_autotune.argtypes = [vp, dp, ip, C.POINTER(C.c_int), vp, ip]
        l.qt_batch_kernel_ms.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_int)]
        l.qt_server_create.argtypes = [C.POINTER(ServerParams), C.POINTER(vp)]
        l.qt_server_destroy.argtypes = [vp]
        l.qt_server_submit.argtypes = [vp, vp, ip, vp, dp, dp, C.c_ulonglong, C.POINTER(C.c_longlong)]
        l.qt_server_step.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        l.qt_server_result.argtypes = [vp, C.c_longlong, vp, ip, C.POINTER(C.c_int), C.POINTER(C.c_double),
                                       C.POINTER(C.c_double), C.POINTER(C.c_longlong)]
        l.qt_server_release.argtypes = [vp, C.c_longlong]
        l.qt_server_submit_sweep.argtypes = [vp, vp, ip, vp, dp, dp, C.POINTER(MasterParams), C.POINTER(C.c_longlong)]
        l.qt_server_stats.argtypes = [vp, C.POINTER(ServerStats)]
        l.qt_master_create.argtypes = [vp, ip, C.POINTER(MasterParams), vp, C.POINTER(vp)]
        l.qt_master_destroy.argtypes = [vp]
        l.qt_master_advance.argtypes = [vp, vp, ip, vp, vp, vp]
        l.qt_master_advance_dc.argtypes = [vp, dp, ip, ip, vp, C.POINTER(C.c_int), vp]
        l.qt_master_rho.argtypes = [vp, ip, vp]
        l.qt_master_status.argtypes = [vp, vp, vp, vp, vp]
        l.qt_master_kernel_ms.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_int)]
        _lib = l
    return _lib


def check(rc):
    if rc:
        raise QtError(lib().qt_last_error().decode(), rc)


def _cbuf(a):
    """complex128 C-contiguous copy and its pointer."""
    a = np.ascontiguousarray(np.asarray(a, dtype=np.complex128))
    return a, a.ctypes.data


# ---------------------------------------------------------------- elements (spec builders)

class Element:
    """A free element or an interaction, as a spec statement."""

    def __init__(self, stmt, dim=None):
        self.stmt = stmt
        self.dim = dim


class mode:  # noqa: N801  (mirrors the reference namespace mode::)
    @staticmethod
    def make(cutoff, delta=-10.0, kappa=0.0, eta=0.0, picture="Sch"):
        """mode::make(ParsPumpedLossy, QM_Picture), CPPQEDelements/frees/Mode.tcc:40-60."""
        eta = complex(eta)
        return Element(f"free mode cutoff={int(cutoff)} delta={delta!r} kappa={kappa!r} "
                       f"eta={eta.real!r} etaI={eta.imag!r} picture={picture}", int(cutoff))


class qbit:  # noqa: N801
    @staticmethod
    def make(deltaA=-10.0, gamma=0.0, eta=0.0, picture="Sch"):
        """qbit::make(ParsPumpedLossy, QM_Picture), CPPQEDelements/frees/Qbit.cc:217-246."""
        eta = complex(eta)
        return Element(f"free qbit deltaA={deltaA!r} gamma={gamma!r} eta={eta.real!r} etaI={eta.imag!r} "
                       f"picture={picture}", 2)


def pumped_two_level_atom_sch(gamma, deltaA, eta):
    """PumpedTwoLevelAtomSch, CPPQEDelements/frees/PumpedTwoLevelAtom.cc:35-49."""
    eta = complex(eta)
    return Element(f"free ptla gamma={gamma!r} deltaA={deltaA!r} eta={eta.real!r} etaI={eta.imag!r}", 2)


class Interaction:
    def __init__(self, kind, frees, **pars):
        self.kind, self.frees, self.pars = kind, list(frees), pars


class jaynescummings:  # noqa: N801
    @staticmethod
    def make(q, m, g):
        """jaynescummings::make, CPPQEDelements/interactions/JaynesCummings_.h:85-90."""
        g = complex(g)
        return Interaction("jc", [q, m], g=g.real, gI=g.imag)


class coupledmodes:  # noqa: N801
    @staticmethod
    def make(m1, m2, u):
        """coupledmodes (CM_XX_RWA), CPPQEDelements/interactions/CoupledModes.cc:36-41."""
        u = complex(u)
        return Interaction("bs", [m1, m2], g=u.real, gI=u.imag)


def zz(q1, q2, J):
    """J sigma_z (x) sigma_z between two qubits (new element: the reference lacks sigma_z)."""
    return Interaction("zz", [q1, q2], J=J)


class Act:
    """composite::_<i,j>(interaction): the interaction acts on free indices (i, j)."""

    def __init__(self, i, j, ia):
        self.axes, self.ia = (i, j), ia


def _ia_stmt(ia, axes):
    kv = " ".join(f"{k}={v!r}" for k, v in ia.pars.items())
    return f"ia {ia.kind} {axes[0]} {axes[1]} {kv}"


class binary:  # noqa: N801
    @staticmethod
    def make(ia):
        """binary::make, CPPQEDcore/composites/BinarySystem.cc:323-333."""
        f0, f1 = ia.frees
        return System("\n".join([f0.stmt, f1.stmt, _ia_stmt(ia, (0, 1)), "system binary"]))


class composite:  # noqa: N801
    @staticmethod
    def make(frees, acts=()):
        """composite::make(frees..., acts...), CPPQEDcore/composites/Composite.h:363-367."""
        lines = [f.stmt for f in frees] + [_ia_stmt(a.ia, a.axes) for a in acts] + ["system composite"]
        return System("\n".join(lines))


# ---------------------------------------------------------------- system

class System:
    """A lowered system: host tables plus device copy (qt_sys)."""

    def __init__(self, spec):
        if isinstance(spec, Element):
            spec = spec.stmt
        self.spec = spec
        h = C.c_void_p()
        check(lib().qt_sys_build(spec.encode(), C.byref(h)))
        self._init(h)

    @classmethod
    def adopt(cls, handle):
        """A System owning an already built qt_sys handle (e.g. from qt_sys_build_terms)."""
        s = cls.__new__(cls)
        s.spec = None
        s._init(handle)
        return s

    def _init(self, h):
        self.h = h
        info = SysInfo()
        check(lib().qt_sys_info(self.h, C.byref(info)))
        self.info = info
        self.dim = info.dim
        self.dims = tuple(info.dims[:info.rank])
        self.n_jumps = info.n_jumps
        self.n_avr = info.n_avr
        self.highest_frequency = info.highest_frequency

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.qt_sys_destroy(self.h)
            self.h = None

    def table(self, which=0):
        """(cols[dim, w], vals[dim, w], freqs[dim, w]) of H_nH/i (0), J_m (1+m), observable (1+n_jumps+q),
        or H_nH/i in its constant-offset form (-1)."""
        w = C.c_int()
        check(lib().qt_sys_table(self.h, which, C.byref(w), None, None, None))
        n, w = self.dim, w.value
        cols = np.zeros(n * w, np.int32)
        vals = np.zeros(n * w, np.complex128)
        freqs = np.zeros(n * w, np.complex128)
        check(lib().qt_sys_table(self.h, which, None, cols.ctypes.data, vals.ctypes.data, freqs.ctypes.data))
        return cols.reshape(n, w), vals.reshape(n, w), freqs.reshape(n, w)

    def rules(self):
        """(Exact exponents per basis state, avr_imag, avr_var_of, avr_sqrt) (qt_sys_rules)"""
        z = np.zeros(self.dim, np.complex128)
        im, var, sq = (np.zeros(self.n_avr, np.int32) for _ in range(3))
        check(lib().qt_sys_rules(self.h, z.ctypes.data, im.ctypes.data, var.ctypes.data, sq.ctypes.data))
        return z, im, var, sq

    def dense(self, which=0):
        """Dense matrix of table `which` assembled from its ELL rows (testing aid)."""
        cols, vals, _ = self.table(which)
        m = np.zeros((self.dim, self.dim), np.complex128)
        rows = np.repeat(np.arange(self.dim), cols.shape[1])
        np.add.at(m, (rows, cols.ravel()), vals.ravel())
        return m

    def apply_heff(self, psi_dev_ptr, dpsidt_dev_ptr, n=1, precision=QT_FP64, stream=None, t=0.0, t0=0.0):
        check(lib().qt_apply_heff(self.h, t, t0, psi_dev_ptr, dpsidt_dev_ptr, n, precision, stream))


# ---------------------------------------------------------------- run-loop intervals

def run_intervals(T=None, Dt=0.1, NDt=0):
    """Stream-interval lengths of trajectory::run in Dt mode (CPPQEDutils/Trajectory.h:560-595):
    NDt > 0: NDt intervals of Dt; otherwise advance(min(Dt, T - t)) while t < T."""
    if NDt:
        return np.full(int(NDt), float(Dt))
    out, t = [], 0.0
    while t < T:
        dt = min(Dt, T - t)
        out.append(dt)
        t = t + dt
    return np.array(out)


# ---------------------------------------------------------------- MCWF ensemble

def mcwf_pars(epsRel=1e-6, epsAbs=1e-12, dpLimit=0.01, overshootTol=10.0, dtInit=0.0, precision=QT_FP64,
              seed=1001, traj0=0, jump_log_cap=64, layout=(0, 0), occupancy=0, group_lanes=0, controller="odeint",
              sched_chunk=0):
    """mcwf::Pars + ode_engine::Pars with the reference defaults (MCWF_Trajectory.h:28-47, ODE.h:31-63).
    controller: "odeint" (ODE_EngineBoost, the default) or "gsl" (ODE_EngineGSL, ODE_GSL.h).
    sched_chunk: stream intervals per scheduling chunk (0 automatic, < 0 whole trajectories)."""
    ctl = {"odeint": QT_CONTROLLER_ODEINT, "gsl": QT_CONTROLLER_GSL}[controller]
    return McwfParams(epsRel, epsAbs, dpLimit, overshootTol, dtInit, precision, ctl, seed, traj0,
                      jump_log_cap, layout[0], layout[1], occupancy, group_lanes, sched_chunk)


class Ensemble:
    """Device-resident ensemble of MCWF trajectories (mcwf::makeEnsemble + trajectory::Ensemble)."""

    def __init__(self, system, n_traj, psi0, pars=None):
        self.system = system
        self.n = int(n_traj)
        self.pars = pars or mcwf_pars()
        psi0, p = _cbuf(psi0)
        if psi0.size != system.dim:
            raise ValueError("initial state has the wrong dimension")
        h = C.c_void_p()
        check(lib().qt_mcwf_create(system.h, self.n, C.byref(self.pars), p, C.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.qt_mcwf_destroy(self.h)
            self.h = None

    def advance(self, intervals):
        """Advances through the stream intervals; returns rows (time, mean dtDid, processed averages)."""
        iv = np.ascontiguousarray(np.asarray(intervals, dtype=np.float64))
        rows = np.zeros((len(iv), 2 + self.system.n_avr))
        check(lib().qt_mcwf_advance(self.h, iv.ctypes.data, len(iv), rows.ctypes.data, None))
        return rows

    def reset(self, stream=None):
        check(lib().qt_mcwf_reset(self.h, stream))

    def advance_dc(self, t_end, dc, max_rows=1024):
        """qt_mcwf_advance_dc (one trajectory): rows after every dc MCWF steps and at t_end, until
        t_end is reached; shape (rows, 2 + n_avr)."""
        out = []
        while True:
            rows = np.zeros((max_rows, 2 + self.system.n_avr))
            n = C.c_int()
            check(lib().qt_mcwf_advance_dc(self.h, float(t_end), int(dc), int(max_rows), rows.ctypes.data,
                                           C.byref(n), None))
            out.append(rows[:n.value])
            if n.value < max_rows or (n.value and rows[n.value - 1, 0] >= t_end):
                break
        return np.concatenate(out)

    def set_draws(self, draws):
        """qt_mcwf_set_draws: draws[i, j] is trajectory i's uniform for its j-th MCWF step from now
        (None: back to the counter RNG)."""
        if draws is None:
            check(lib().qt_mcwf_set_draws(self.h, None, 0))
            return
        d = np.ascontiguousarray(np.asarray(draws, dtype=np.float64))
        if d.ndim != 2 or d.shape[0] != self.n:
            raise ValueError(f"draws must have shape ({self.n}, n_per)")
        check(lib().qt_mcwf_set_draws(self.h, d.ctypes.data, d.shape[1]))

    def advance_noalloc(self, intervals):
        """qt_mcwf_advance_noalloc: as advance(), with every per-launch device buffer in the library's
        static scratch (no device allocation in the call)."""
        iv = np.ascontiguousarray(np.asarray(intervals, dtype=np.float64))
        rows = np.zeros((len(iv), 2 + self.system.n_avr))
        check(lib().qt_mcwf_advance_noalloc(self.h, iv.ctypes.data, len(iv), rows.ctypes.data, None))
        return rows

    def advance_async(self, intervals, stream=None):
        iv = np.ascontiguousarray(np.asarray(intervals, dtype=np.float64))
        check(lib().qt_mcwf_advance_async(self.h, iv.ctypes.data, len(iv), stream))

    def set_graph(self, enable=True):
        """Replay the advance's launch sequence from a captured hipGraph (qt_mcwf_set_graph);
        returns (captures, replays) so far."""
        c, r = C.c_longlong(), C.c_longlong()
        check(lib().qt_mcwf_set_graph(self.h, -1 if enable is None else int(bool(enable)), C.byref(c), C.byref(r)))
        return c.value, r.value

    def traj_averages(self):
        out = np.zeros((self.n, self.system.n_avr))
        check(lib().qt_mcwf_traj_averages(self.h, out.ctypes.data))
        return out

    def rho(self):
        """The ensemble's density operator (1/N) sum_k |psi_k><psi_k| of the current states (qt_mcwf_rho)."""
        d = self.system.dim
        out = np.zeros((d, d), np.complex128)
        check(lib().qt_mcwf_rho(self.h, out.ctypes.data))
        return out

    def state(self, i):
        psi = np.zeros(self.system.dim, np.complex128)
        t = C.c_double()
        check(lib().qt_mcwf_state(self.h, i, psi.ctypes.data, C.byref(t)))
        return psi, t.value

    def jumplog(self, i, cap=4096):
        t = np.zeros(cap)
        ch = np.zeros(cap, np.int32)
        n = C.c_int()
        check(lib().qt_mcwf_jumplog(self.h, i, t.ctypes.data, ch.ctypes.data, cap, C.byref(n)))
        m = min(n.value, cap, self.pars.jump_log_cap)
        return t[:m], ch[:m], n.value

    def counters(self):
        per = np.zeros((self.n, 8), np.int64)
        tot = np.zeros(8, np.int64)
        check(lib().qt_mcwf_counters(self.h, per.ctypes.data, tot.ctypes.data))
        return per, tot

    def kernel_ms(self):
        ms = C.c_double()
        n = C.c_int()
        check(lib().qt_mcwf_kernel_ms(self.h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def time(self):
        """The ensemble's time (every trajectory stops at each stream point)."""
        t = C.c_double()
        check(lib().qt_mcwf_state(self.h, 0, None, C.byref(t)))
        return t.value

    # ------------------------------------------------------------ checkpoint / resume

    def get_states(self, first=0, count=None):
        """(TrajState array, psi[count, dim], jump times[count, cap], jump channels[count, cap])."""
        count = self.n - first if count is None else count
        st = (TrajState * count)()
        cap = max(0, self.pars.jump_log_cap)
        psi = np.zeros((count, self.system.dim), np.complex128)
        jt = np.zeros((count, max(cap, 1)))
        jc = np.zeros((count, max(cap, 1)), np.int32)
        check(lib().qt_mcwf_get_states(self.h, first, count, st, psi.ctypes.data, jt.ctypes.data, jc.ctypes.data))
        return st, psi, jt, jc

    def set_states(self, st, psi, jt=None, jc=None, first=0):
        count = len(st)
        psi = np.ascontiguousarray(np.asarray(psi, dtype=np.complex128).reshape(count, self.system.dim))
        arr = (TrajState * count)(*st)
        jt = None if jt is None else np.ascontiguousarray(jt, dtype=np.float64)
        jc = None if jc is None else np.ascontiguousarray(jc, dtype=np.int32)
        check(lib().qt_mcwf_set_states(self.h, first, count, arr, psi.ctypes.data,
                                       None if jt is None else jt.ctypes.data, None if jc is None else jc.ctypes.data))

    def records(self):
        """The ensemble as state-file records (one per trajectory), in archive order.  The device
        keeps no maximal-overshoot or norm-deviation figures (logging only): they are written as 0."""
        from . import statefile as SF
        st, psi, jt, jc = self.get_states()
        out = []
        for i in range(self.n):
            c = st[i].counters
            nj = min(int(c[C_JUMPS]), self.pars.jump_log_cap)
            out.append(SF.TrajRecord(
                psi=psi[i].reshape(self.system.dims), t=st[i].t, dtDid=st[i].dtDid, dtTry=st[i].dtTry,
                ode_derivs=c[C_DERIVS], ode_steps=c[C_ODE], ode_failed=c[C_FAILED],
                rng=SF.format_philox(self.pars.seed, self.pars.traj0 + i, st[i].ordinal),
                mcwf_steps=c[C_MCWF], overshot=c[C_OVERSHOOTS], tolerance_overshot=c[C_STEPBACKS],
                jumps=[(float(jt[i, k]), int(jc[i, k])) for k in range(nj)]))
        return out

    def save_state(self, path, append=True):
        """Appends the ensemble as one archive to a state file (writeViaSStream,
        CPPQEDutils/Trajectory.h:395-410): an 'Ensemble of MCWF_Trajectory' archive, or an
        'MCWF_Trajectory' archive for a single trajectory."""
        from . import statefile as SF
        recs = self.records()
        kind = SF.MCWF_ID if self.n == 1 else SF.ENSEMBLE_ID
        SF.write(path, [SF.Frame(trajectory_id=kind, rank=len(self.system.dims), trajs=recs)], append=append)

    def load_state(self, path, rng="keep"):
        """Continues from the last archive of a state file (readViaSStream, Trajectory.h:392-405).
        rng='keep' requires the archived Philox records of this ensemble's own streams (seed and
        trajectory keys) and continues them; rng='restart' takes psi, times and counters from any
        archive (also the reference's pcg64 ones) and restarts this ensemble's own streams."""
        from . import statefile as SF
        fr = SF.read_last(path)
        if fr.trajectory_id not in (SF.MCWF_ID, SF.ENSEMBLE_ID):
            raise SF.StateFileError("not an MCWF archive: %r" % fr.trajectory_id)
        if len(fr.trajs) != self.n:
            raise SF.StateFileError("archive holds %d trajectories, the ensemble %d" % (len(fr.trajs), self.n))
        cap = max(0, self.pars.jump_log_cap)
        st = (TrajState * self.n)()
        psi = np.zeros((self.n, self.system.dim), np.complex128)
        jt = np.zeros((self.n, max(cap, 1)))
        jc = np.zeros((self.n, max(cap, 1)), np.int32)
        for i, r in enumerate(fr.trajs):
            if r.psi.size != self.system.dim or tuple(r.psi.shape) != tuple(self.system.dims):
                raise SF.StateFileError("archived state has shape %s, the system %s" % (r.psi.shape, self.system.dims))
            if rng == "keep":
                ph = SF.parse_philox(r.rng)
                if ph is None or ph[:2] != (self.pars.seed, self.pars.traj0 + i):
                    raise SF.StateFileError("trajectory %d: the archive's random stream is not this ensemble's "
                                            "(use rng='restart' to continue with the ensemble's own streams)" % i)
                st[i].ordinal = ph[2]
            elif rng == "restart":
                st[i].ordinal = 0
            else:
                raise ValueError("rng must be 'keep' or 'restart'")
            st[i].t, st[i].dtDid, st[i].dtTry = r.t, r.dtDid, r.dtTry
            c = st[i].counters
            c[C_MCWF], c[C_ODE], c[C_FAILED], c[C_DERIVS] = r.mcwf_steps, r.ode_steps, r.ode_failed, r.ode_derivs
            c[C_TRIES] = r.ode_steps + r.ode_failed
            c[C_STEPBACKS], c[C_OVERSHOOTS], c[C_JUMPS] = r.tolerance_overshot, r.overshot, len(r.jumps)
            psi[i] = r.psi.reshape(-1)
            for k, (tj, ch) in enumerate(r.jumps[:cap]):
                jt[i, k], jc[i, k] = tj, ch
        self.set_states(list(st), psi, jt, jc)
        return fr


def batch_variants():
    """Compiled mixed-batch layouts: (waves per work group, rows/lane and ELL register width of the
    work-group team, rows/lane and width of the wavefront team)."""
    n = lib().qt_batch_variants(None, 0)
    t = np.zeros(5 * n, np.int32)
    lib().qt_batch_variants(t.ctypes.data, n)
    return [tuple(int(x) for x in t[5 * i:5 * i + 5]) for i in range(n)]


class Batch:
    """Several ensembles (different systems) advanced by one batched launch (qt_batch_*): the
    work-group-team members in one persistent kernel, the quarter- and wavefront-team members in a
    second beside it."""

    def __init__(self, members, variant=-1, team=None):
        self.members = list(members)
        arr = (C.c_void_p * len(self.members))(*[m.h for m in self.members])
        p = BatchParams(variant, (C.c_int * 16)(*([-1] * 16)))
        if team is not None:
            for i, t in enumerate(team):
                p.team[i] = t
        h = C.c_void_p()
        check(lib().qt_batch_create(arr, len(self.members), C.byref(p), C.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.qt_batch_destroy(self.h)
            self.h = None

    def layout(self):
        v = C.c_int()
        team = np.zeros(len(self.members), np.int32)
        check(lib().qt_batch_layout(self.h, C.byref(v), team.ctypes.data))
        return v.value, team.tolist()

    def advance(self, intervals):
        """intervals: one list for every member, or a list of per-member lists; returns per-member rows."""
        if np.ndim(intervals[0]) == 0:
            intervals = [intervals] * len(self.members)
        ivs = [np.ascontiguousarray(np.asarray(iv, dtype=np.float64)) for iv in intervals]
        rows = [np.zeros((len(iv), 2 + m.system.n_avr)) for iv, m in zip(ivs, self.members)]
        ivp = (C.c_void_p * len(ivs))(*[iv.ctypes.data for iv in ivs])
        nis = np.array([len(iv) for iv in ivs], np.int32)
        rp = (C.c_void_p * len(rows))(*[r.ctypes.data for r in rows])
        check(lib().qt_batch_advance(self.h, ivp, nis.ctypes.data, rp, None))
        return rows

    def autotune(self, probe_span=0.5, probe_traj=2048):
        v = C.c_int()
        ms = np.zeros(64)
        check(lib().qt_batch_autotune(self.h, probe_span, probe_traj, C.byref(v), ms.ctypes.data, 64))
        return v.value, ms[ms > 0]

    def kernel_ms(self):
        ms = C.c_double()
        n = C.c_int()
        check(lib().qt_batch_kernel_ms(self.h, C.byref(ms), C.byref(n)))
        return ms.value, n.value


class Server:
    """Continuous-batching request loop over the mixed batch (qt_server_*): MCWF ensemble requests
    of different systems are admitted by member slots and device memory, advanced together one
    chunk of stream intervals per launch, and retired when done."""

    def __init__(self, hbm_budget=8 << 30, max_active=16, chunk_intervals=10, pars=None):
        self.p = ServerParams(int(hbm_budget), max_active, chunk_intervals, pars or mcwf_pars())
        h = C.c_void_p()
        check(lib().qt_server_create(C.byref(self.p), C.byref(h)))
        self.h = h
        self._systems = {}

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.qt_server_destroy(self.h)
            self.h = None

    def submit(self, system, n_traj, psi0, T, Dt=0.1, seed=1001):
        if np.shape(psi0) != (system.dim,):
            raise ValueError(f"psi0 must have shape ({system.dim},), got {np.shape(psi0)}")
        psi0, p = _cbuf(psi0)
        rid = C.c_longlong()
        check(lib().qt_server_submit(self.h, system.h, int(n_traj), p, float(T), float(Dt), int(seed), C.byref(rid)))
        self._systems[rid.value] = system
        return rid.value

    def submit_sweep(self, systems, rho0, T, Dt=0.1, epsRel=1e-6, epsAbs=1e-12, dtInit=0.0, ss_tol=0.0,
                     autostop=None, kernel="auto"):
        """A Master-equation parameter sweep (one point per system) as a request; its result rows
        are point-major, shape (points * intervals, 2 + n_avr).  autostop as for MasterSweep."""
        d = systems[0].dim
        if np.shape(rho0) != (d, d):
            raise ValueError(f"rho0 must have shape ({d}, {d}), got {np.shape(rho0)}")
        rho = np.ascontiguousarray(np.asarray(rho0, dtype=np.complex128))
        arr = (C.c_void_p * len(systems))(*[s.h for s in systems])
        a_eps, a_abs, a_rep = autostop if autostop else (0.0, 0.0, 0)
        mp = MasterParams(epsRel, epsAbs, dtInit, ss_tol, a_eps, a_abs, a_rep, MASTER_KERNELS[kernel])
        rid = C.c_longlong()
        check(lib().qt_server_submit_sweep(self.h, arr, len(systems), rho.ctypes.data, float(T), float(Dt),
                                           C.byref(mp), C.byref(rid)))
        self._systems[rid.value] = systems[0]
        self._keep = getattr(self, "_keep", {})
        self._keep[rid.value] = list(systems)
        return rid.value

    def step(self):
        a, w = C.c_int(), C.c_int()
        check(lib().qt_server_step(self.h, C.byref(a), C.byref(w)))
        return a.value, w.value

    def result(self, rid):
        """(rows, latency_s, queue_wait_s, tries) of a finished request, or None if not finished;
        raises QtError (with the request's error code) for a request that failed."""
        system = self._systems[rid]
        n = C.c_int()
        lat, wait, tries = C.c_double(), C.c_double(), C.c_longlong()
        rc = lib().qt_server_result(self.h, rid, None, 0, C.byref(n), C.byref(lat), C.byref(wait), C.byref(tries))
        if rc:
            msg = lib().qt_last_error().decode()
            if msg == "request not finished":
                return None
            raise QtError(msg, rc)
        rows = np.zeros((n.value, 2 + system.n_avr))
        check(lib().qt_server_result(self.h, rid, rows.ctypes.data, n.value, C.byref(n), C.byref(lat), C.byref(wait),
                                     C.byref(tries)))
        return rows, lat.value, wait.value, tries.value

    def release(self, rid):
        check(lib().qt_server_release(self.h, rid))
        self._systems.pop(rid, None)
        getattr(self, "_keep", {}).pop(rid, None)

    def stats(self):
        st = ServerStats()
        check(lib().qt_server_stats(self.h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}


COUNTER_NAMES = ("mcwf_steps", "ode_steps", "failed", "derivs", "tries", "step_backs", "overshoots", "jumps")


def mcwf_plain_run(system, psi0, intervals, pars=None, traj=0, jcap=4096):
    """The unbatched GPU form: one trajectory, one launch per RKCK try (qt_mcwf_plain_run)."""
    pars = pars or mcwf_pars()
    psi0, p = _cbuf(psi0)
    iv = np.ascontiguousarray(np.asarray(intervals, dtype=np.float64))
    obs = np.zeros((len(iv), system.n_avr))
    psi = np.zeros(system.dim, np.complex128)
    jt = np.zeros(jcap)
    jch = np.zeros(jcap, np.int32)
    cnt = np.zeros(8, np.int64)
    check(lib().qt_mcwf_plain_run(system.h, C.byref(pars), traj, p, iv.ctypes.data, len(iv), obs.ctypes.data,
                                  psi.ctypes.data, jt.ctypes.data, jch.ctypes.data, jcap, cnt.ctypes.data))
    nj = int(min(cnt[7], jcap))
    return dict(obs=obs, psi=psi, jump_t=jt[:nj], jump_ch=jch[:nj], counters=cnt)


# ---------------------------------------------------------------- Master sweep

class MasterSweep:
    """A batch of density operators on device, point k under systems[k] (master::make + Master)."""

    def __init__(self, systems, rho0, epsRel=1e-6, epsAbs=1e-12, dtInit=0.0, ss_tol=0.0, autostop=None,
                 kernel="auto"):
        """autostop = (epsilon, epsilonAbs, repetition): the reference's AutostopHandlerGeneric
        criterion per point (CPPQEDutils/Trajectory.h:179-222); None: off.  kernel: "auto" (the
        MFMA kernel up to d = 96, the sparse one beyond), "mfma" or "sparse"."""
        self.systems = list(systems)
        self.n = len(self.systems)
        self.dim = self.systems[0].dim
        self.n_avr = self.systems[0].n_avr
        rho0 = np.ascontiguousarray(np.asarray(rho0, dtype=np.complex128))
        if rho0.shape != (self.dim, self.dim):
            raise ValueError("rho0 has the wrong shape")
        arr = (C.c_void_p * self.n)(*[s_.h.value if hasattr(s_.h, "value") else s_.h for s_ in self.systems])
        a_eps, a_abs, a_rep = autostop if autostop else (0.0, 0.0, 0)
        self.pars = MasterParams(epsRel, epsAbs, dtInit, ss_tol, a_eps, a_abs, a_rep, MASTER_KERNELS[kernel])
        h = C.c_void_p()
        check(lib().qt_master_create(arr, self.n, C.byref(self.pars), rho0.ctypes.data, C.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.qt_master_destroy(self.h)
            self.h = None

    def advance(self, intervals):
        """rows[point, k] = (time, dtDid, processed averages); dnorm[point, k] = ||drho/dt||_F."""
        iv = np.ascontiguousarray(np.asarray(intervals, dtype=np.float64))
        rows = np.zeros((self.n, len(iv), 2 + self.n_avr))
        dn = np.zeros((self.n, len(iv)))
        check(lib().qt_master_advance(self.h, iv.ctypes.data, len(iv), rows.ctypes.data, dn.ctypes.data, None))
        return rows, dn

    def advance_dc(self, t_end, dc, max_rows=1024):
        """qt_master_advance_dc (a sweep of one point): rows after every dc steps and at t_end."""
        out = []
        while True:
            rows = np.zeros((max_rows, 2 + self.n_avr))
            n = C.c_int()
            check(lib().qt_master_advance_dc(self.h, float(t_end), int(dc), int(max_rows), rows.ctypes.data,
                                             C.byref(n), None))
            out.append(rows[:n.value])
            if n.value < max_rows or (n.value and rows[n.value - 1, 0] >= t_end):
                break
        return np.concatenate(out)

    def rho(self, point):
        r = np.zeros((self.dim, self.dim), np.complex128)
        check(lib().qt_master_rho(self.h, point, r.ctypes.data))
        return r

    def status(self):
        done = np.zeros(self.n, np.int32)
        tss = np.zeros(self.n)
        t = np.zeros(self.n)
        cnt = np.zeros((self.n, 8), np.int64)
        check(lib().qt_master_status(self.h, done.ctypes.data, tss.ctypes.data, t.ctypes.data, cnt.ctypes.data))
        return done, tss, t, cnt

    def kernel_ms(self):
        ms = C.c_double()
        n = C.c_int()
        check(lib().qt_master_kernel_ms(self.h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    # ------------------------------------------------------------ checkpoint / resume

    def get_point(self, point):
        st = MasterState()
        rho = np.zeros((self.dim, self.dim), np.complex128)
        check(lib().qt_master_get_point(self.h, point, C.byref(st), rho.ctypes.data))
        return st, rho

    def set_point(self, point, st, rho):
        rho = np.ascontiguousarray(np.asarray(rho, dtype=np.complex128).reshape(self.dim, self.dim))
        check(lib().qt_master_set_point(self.h, point, C.byref(st), rho.ctypes.data))

    def save_state(self, path, point=0, append=True):
        """Appends one point as a 'Master' archive (Master::stateIO, Master.h:89-102)."""
        from . import statefile as SF
        st, rho = self.get_point(point)
        dims = self.systems[point].dims
        c = st.counters
        rec = SF.MasterRecord(rho=rho.reshape(tuple(dims) + tuple(dims)), t=st.t, dtDid=st.dtDid, dtTry=st.dtTry,
                              ode_derivs=c[C_DERIVS], ode_steps=c[C_ODE], ode_failed=c[C_FAILED])
        SF.write(path, [SF.Frame(trajectory_id=SF.MASTER_ID, rank=2 * len(dims), master=rec)], append=append)

    def load_state(self, path, point=0):
        """Continues one point from the last archive of a state file ('Master', or 'ArrayOnly' for a
        bare density operator as an initial condition)."""
        from . import statefile as SF
        fr = SF.read_last(path)
        st = MasterState()
        if fr.trajectory_id == SF.MASTER_ID:
            m = fr.master
            rho = m.matrix()
            st.t, st.dtDid, st.dtTry = m.t, m.dtDid, m.dtTry
            c = st.counters
            c[C_MCWF], c[C_ODE], c[C_FAILED], c[C_DERIVS] = m.ode_steps, m.ode_steps, m.ode_failed, m.ode_derivs
            c[C_TRIES] = m.ode_steps + m.ode_failed
        elif fr.trajectory_id == SF.ARRAY_ONLY:
            cur, _ = self.get_point(point)
            d = self.dim
            rho = fr.array.reshape(d, d)
            st.t, st.dtDid, st.dtTry = cur.t, cur.dtDid, cur.dtTry
        else:
            raise SF.StateFileError("not a Master archive: %r" % fr.trajectory_id)
        if rho.shape != (self.dim, self.dim):
            raise SF.StateFileError("archived density operator has the wrong dimension")
        st.t_ss = -1.0
        self.set_point(point, st, rho)
        return fr


from . import statefile, trajectory  # noqa: E402,F401  (Q.statefile, Q.trajectory)
