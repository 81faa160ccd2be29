"""ctypes binding of the CPU oracle (oracle/liboracle.so).

Test infrastructure only: the oracle is the parity reference, never the product path.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(os.path.dirname(_HERE), "oracle", "liboracle.so")


class Params(C.Structure):
    _fields_ = [("epsRel", C.c_double), ("epsAbs", C.c_double), ("dpLimit", C.c_double),
                ("overshootTol", C.c_double), ("dtInit", C.c_double), ("rng", C.c_int),
                ("controller", C.c_int), ("seed", C.c_ulonglong), ("stream", C.c_ulonglong)]


class Run(C.Structure):
    _fields_ = [("mode", C.c_int), ("dc", C.c_int), ("Dt", C.c_double), ("NDt", C.c_long), ("T", C.c_double)]


class TrajState(C.Structure):
    _fields_ = [("t", C.c_double), ("dtDid", C.c_double), ("dtTry", C.c_double),
                ("dpMaxOvershoot", C.c_double), ("dpTolMaxOvershoot", C.c_double), ("normMaxDev", C.c_double),
                ("pcg", C.c_ulonglong * 4), ("ode", C.c_longlong * 3), ("mcwf", C.c_longlong * 3),
                ("ordinal", C.c_ulonglong)]


class MasterState(C.Structure):
    _fields_ = [("t", C.c_double), ("dtDid", C.c_double), ("dtTry", C.c_double), ("ode", C.c_longlong * 3)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        l = C.CDLL(_LIB)
        l.orc_last_error.restype = C.c_char_p
        l.orc_system_create.restype = C.c_void_p
        l.orc_system_create.argtypes = [C.c_char_p]
        l.orc_system_destroy.argtypes = [C.c_void_p]
        l.orc_system_dim.restype = C.c_long
        l.orc_system_dim.argtypes = [C.c_void_p]
        l.orc_system_njumps.argtypes = [C.c_void_p]
        l.orc_system_navr.argtypes = [C.c_void_p]
        l.orc_system_highest_frequency.argtypes = [C.c_void_p]
        l.orc_system_highest_frequency.restype = C.c_double
        dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
        l.orc_apply_heff.argtypes = [C.c_void_p, C.c_double, C.c_double, dp, dp]
        l.orc_rates.argtypes = [C.c_void_p, dp, dp]
        l.orc_act_jump.argtypes = [C.c_void_p, dp, C.c_int]
        l.orc_averages.argtypes = [C.c_void_p, dp, dp]
        l.orc_pcg64_uniform.argtypes = [C.c_ulonglong, C.c_ulonglong, C.c_long, dp, C.c_void_p]
        l.orc_philox4x64.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        l.orc_counter_uniform.argtypes = [C.c_ulonglong, C.c_ulonglong, C.c_ulonglong]
        l.orc_counter_uniform.restype = C.c_double
        l.orc_mcwf_run.restype = C.c_long
        l.orc_mcwf_run.argtypes = [C.c_void_p, dp, C.POINTER(Params), C.POINTER(Run), dp, C.c_long,
                                   C.c_void_p, C.c_void_p, C.c_long, C.c_void_p, C.c_void_p]
        l.orc_ensemble_run.restype = C.c_long
        l.orc_ensemble_run.argtypes = [C.c_void_p, dp, C.POINTER(Params), C.POINTER(Run), C.c_long, dp, C.c_long, C.c_void_p]
        l.orc_master_run.restype = C.c_long
        l.orc_master_run.argtypes = [C.c_void_p, dp, C.POINTER(Params), C.POINTER(Run), dp, C.c_long, C.c_void_p]
        l.orc_mcwf_resume.restype = C.c_long
        l.orc_mcwf_resume.argtypes = [C.c_void_p, dp, C.POINTER(Params), C.POINTER(Run), C.POINTER(TrajState),
                                      C.c_void_p, C.c_void_p, C.POINTER(C.c_long), C.c_long, dp, C.c_long, C.c_void_p]
        l.orc_master_resume.restype = C.c_long
        l.orc_master_resume.argtypes = [C.c_void_p, dp, C.POINTER(Params), C.POINTER(Run), C.POINTER(MasterState),
                                        dp, C.c_long, C.c_void_p]
        l.orc_master_run_autostop.restype = C.c_long
        l.orc_master_run_autostop.argtypes = [C.c_void_p, dp, C.POINTER(Params), C.POINTER(Run), C.c_double,
                                              C.c_double, C.c_uint, dp, C.c_long, C.c_void_p]
        l.orc_autostop_create.restype = C.c_void_p
        l.orc_autostop_create.argtypes = [C.c_double, C.c_double, C.c_uint]
        l.orc_autostop_destroy.argtypes = [C.c_void_p]
        l.orc_autostop_feed.argtypes = [C.c_void_p, dp, C.c_int]
        l.orc_simulated_run.restype = C.c_long
        l.orc_simulated_run.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                        C.c_double, C.c_int, C.c_void_p, C.c_long, dp, C.c_long, C.c_void_p]
        l.orc_mt19937_uniform.argtypes = [C.c_uint, C.c_long, C.c_double, C.c_double, dp]
        l.orc_slices.restype = C.c_long
        l.orc_slices.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_long, C.c_void_p]
        _lib = l
    return _lib


def _cplx(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.complex128)).view(np.float64)


class System:
    """An oracle system built from a spec string (same grammar as the product's)."""

    def __init__(self, spec):
        self.h = lib().orc_system_create(spec.encode())
        if not self.h:
            raise ValueError(lib().orc_last_error().decode())
        self.dim = lib().orc_system_dim(self.h)
        self.njumps = lib().orc_system_njumps(self.h)
        self.navr = lib().orc_system_navr(self.h)
        self.highest_frequency = lib().orc_system_highest_frequency(self.h)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_system_destroy(self.h)

    def apply_heff(self, psi, t=0.0, t0=0.0, dpsidt=None):
        out = np.zeros(self.dim, np.complex128) if dpsidt is None else np.array(dpsidt, np.complex128)
        o = out.view(np.float64)
        if lib().orc_apply_heff(self.h, t, t0, _cplx(psi), o):
            raise RuntimeError(lib().orc_last_error().decode())
        return o.view(np.complex128)

    def dense_heff(self):
        d = self.dim
        m = np.zeros((d, d), np.complex128)
        for j in range(d):
            e = np.zeros(d, np.complex128)
            e[j] = 1
            m[:, j] = self.apply_heff(e)
        return m

    def rates(self, psi):
        out = np.zeros(self.njumps)
        lib().orc_rates(self.h, _cplx(psi), out)
        return out

    def act_jump(self, psi, m):
        p = _cplx(psi).copy()
        lib().orc_act_jump(self.h, p, m)
        return p.view(np.complex128)

    def averages(self, psi):
        out = np.zeros(self.navr)
        lib().orc_averages(self.h, _cplx(psi), out)
        return out


def params(epsRel=1e-6, epsAbs=1e-12, dpLimit=0.01, overshootTol=10.0, dtInit=0.0, rng="philox",
           controller="odeint", seed=1001, stream=1):
    return Params(epsRel, epsAbs, dpLimit, overshootTol, dtInit, 0 if rng == "pcg64" else 1,
                  0 if controller == "odeint" else 1, seed, stream)


def run_spec(mode="dc", dc=10, Dt=0.1, NDt=0, T=1.0):
    return Run({"dc": 0, "Dt": 1, "NDt": 2}[mode], dc, Dt, NDt, T)


def mcwf_run(system, psi0, p, r, cap=100000, jcap=100000):
    nav = system.navr
    rows = np.zeros((cap, 2 + nav))
    jt = np.zeros(jcap)
    jc = np.zeros(jcap, np.int32)
    cnt = np.zeros(8, np.int64)
    psi = np.zeros(system.dim, np.complex128)
    n = lib().orc_mcwf_run(system.h, _cplx(psi0), C.byref(p), C.byref(r), rows, cap,
                           jt.ctypes.data, jc.ctypes.data, jcap, cnt.ctypes.data, psi.view(np.float64).ctypes.data)
    if n < 0:
        raise RuntimeError(lib().orc_last_error().decode())
    nj = int(cnt[7])
    return dict(rows=rows[:min(n, cap)], jump_t=jt[:nj], jump_ch=jc[:nj], counters=cnt, psi=psi)


def ensemble_run(system, psi0, p, r, ntraj, cap=100000):
    rows = np.zeros((cap, 2 + system.navr))
    tries = np.zeros(1, np.int64)
    n = lib().orc_ensemble_run(system.h, _cplx(psi0), C.byref(p), C.byref(r), ntraj, rows, cap, tries.ctypes.data)
    if n < 0:
        raise RuntimeError(lib().orc_last_error().decode())
    return rows[:min(n, cap)], int(tries[0])


def master_run(system, rho0, p, r, cap=100000):
    rows = np.zeros((cap, 2 + system.navr))
    d = system.dim
    rho = np.zeros((d, d), np.complex128)
    n = lib().orc_master_run(system.h, _cplx(rho0), C.byref(p), C.byref(r), rows, cap, rho.view(np.float64).ctypes.data)
    if n < 0:
        raise RuntimeError(lib().orc_last_error().decode())
    return rows[:min(n, cap)], rho


def counter_uniform(seed, traj, step):
    return lib().orc_counter_uniform(seed, traj, step)


def philox4x64(ctr, key):
    c = (C.c_ulonglong * 4)(*ctr)
    k = (C.c_ulonglong * 2)(*key)
    o = (C.c_ulonglong * 4)()
    lib().orc_philox4x64(C.byref(c), C.byref(k), C.byref(o))
    return list(o)


def pcg64_uniform(seed, stream, n):
    out = np.zeros(n)
    raw = np.zeros(n, np.uint64)
    lib().orc_pcg64_uniform(seed, stream, n, out, raw.ctypes.data)
    return out, raw


def _split128(x):
    return (x >> 64) & ((1 << 64) - 1), x & ((1 << 64) - 1)


def mcwf_resume(system, rec, p, r, cap=100000, jcap=100000):
    """Continues a trajectory from a state-file record (cppqed_amd.statefile.TrajRecord with a
    pcg64 random-engine record); returns the streamed rows and the final record."""
    from cppqed_amd import statefile as SF
    st = TrajState()
    st.t, st.dtDid, st.dtTry = rec.t, rec.dtDid, rec.dtTry
    st.dpMaxOvershoot, st.dpTolMaxOvershoot, st.normMaxDev = (rec.dp_max_overshoot, rec.dp_tolerance_max_overshoot,
                                                              rec.norm_max_deviation)
    state, inc = SF.parse_pcg64(rec.rng)
    st.pcg[0], st.pcg[1] = _split128(state)
    st.pcg[2], st.pcg[3] = _split128(inc)
    st.ode[0], st.ode[1], st.ode[2] = rec.ode_derivs, rec.ode_steps, rec.ode_failed
    st.mcwf[0], st.mcwf[1], st.mcwf[2] = rec.mcwf_steps, rec.overshot, rec.tolerance_overshot
    jt = np.zeros(jcap)
    jc = np.zeros(jcap, np.int32)
    for k, (tj, ch) in enumerate(rec.jumps):
        jt[k], jc[k] = tj, ch
    nj = C.c_long(len(rec.jumps))
    rows = np.zeros((cap, 2 + system.navr))
    psi = np.zeros(system.dim, np.complex128)
    n = lib().orc_mcwf_resume(system.h, _cplx(rec.psi.reshape(-1)), C.byref(p), C.byref(r), C.byref(st),
                              jt.ctypes.data, jc.ctypes.data, C.byref(nj), jcap, rows, cap,
                              psi.view(np.float64).ctypes.data)
    if n < 0:
        raise RuntimeError(lib().orc_last_error().decode())
    out = SF.TrajRecord(psi=psi.reshape(rec.psi.shape), t=st.t, dtDid=st.dtDid, dtTry=st.dtTry,
                        ode_derivs=st.ode[0], ode_steps=st.ode[1], ode_failed=st.ode[2],
                        rng=SF.format_pcg64((st.pcg[0] << 64) | st.pcg[1], (st.pcg[2] << 64) | st.pcg[3]),
                        mcwf_steps=st.mcwf[0], overshot=st.mcwf[1], tolerance_overshot=st.mcwf[2],
                        dp_max_overshoot=st.dpMaxOvershoot, dp_tolerance_max_overshoot=st.dpTolMaxOvershoot,
                        norm_max_deviation=st.normMaxDev,
                        jumps=[(float(jt[k]), int(jc[k])) for k in range(nj.value)])
    return rows[:min(n, cap)], out


def master_resume(system, rec, p, r, cap=100000):
    """Continues a Master evolution from a state-file record (cppqed_amd.statefile.MasterRecord)."""
    from cppqed_amd import statefile as SF
    st = MasterState()
    st.t, st.dtDid, st.dtTry = rec.t, rec.dtDid, rec.dtTry
    st.ode[0], st.ode[1], st.ode[2] = rec.ode_derivs, rec.ode_steps, rec.ode_failed
    d = system.dim
    rows = np.zeros((cap, 2 + system.navr))
    rho = np.zeros((d, d), np.complex128)
    n = lib().orc_master_resume(system.h, _cplx(rec.matrix()), C.byref(p), C.byref(r), C.byref(st), rows, cap,
                                rho.view(np.float64).ctypes.data)
    if n < 0:
        raise RuntimeError(lib().orc_last_error().decode())
    out = SF.MasterRecord(rho=rho.reshape(rec.rho.shape), t=st.t, dtDid=st.dtDid, dtTry=st.dtTry,
                          ode_derivs=st.ode[0], ode_steps=st.ode[1], ode_failed=st.ode[2])
    return rows[:min(n, cap)], out


class Autostop:
    """The oracle's AutostopHandlerGeneric; call with a streamed array, True = stop."""

    def __init__(self, eps, eps_abs, rep):
        self.h = lib().orc_autostop_create(eps, eps_abs, rep)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_autostop_destroy(self.h)

    def __call__(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        return bool(lib().orc_autostop_feed(self.h, v, len(v)))


def master_run_autostop(system, rho0, p, r, eps, eps_abs, rep, cap=100000):
    rows = np.zeros((cap, 2 + system.navr))
    d = system.dim
    rho = np.zeros((d, d), np.complex128)
    n = lib().orc_master_run_autostop(system.h, _cplx(rho0), C.byref(p), C.byref(r), eps, eps_abs, rep, rows, cap,
                                      rho.view(np.float64).ctypes.data)
    if n < 0:
        raise RuntimeError(lib().orc_last_error().decode())
    return rows[:min(n, cap)], rho


SIMULATED = {"evolved": 0, "oscillator": 1}


def simulated_run(which, controller="odeint", dtInit=0.02, epsRel=1e-6, epsAbs=1e-18, T=1.0, Dt=0.1, dc=0,
                  times=None, cap=200000):
    """A classical test equation on the oracle's ODE engine (the reference's Simulated trajectory):
    dc-mode rows (dc > 0: every dc ODE steps), Dt-mode rows, or one row after advancing to each of
    ``times``.  Columns: t, dtDid, y, y' (complex).  Returns (rows, stats = [derivative calls, steps,
    failed tries, tries])."""
    rows = np.zeros((cap, 6))
    st = np.zeros(4, np.int64)
    tp = None
    if times is not None:
        times = np.ascontiguousarray(times, dtype=np.float64)
        tp = times.ctypes.data
    n = lib().orc_simulated_run(SIMULATED[which], 1 if controller == "gsl" else 0, dtInit, epsRel, epsAbs, T, Dt, dc,
                                tp, 0 if times is None else len(times), rows, cap, st.ctypes.data)
    if n < 0:
        raise RuntimeError(lib().orc_last_error().decode())
    r = rows[:min(n, cap)]
    return np.column_stack([r[:, 0], r[:, 1], r[:, 2] + 1j * r[:, 3], r[:, 4] + 1j * r[:, 5]]), st


def mt19937_uniform(seed, n, a, b):
    out = np.zeros(n)
    lib().orc_mt19937_uniform(seed, n, a, b, out)
    return out


def slices(dims, V, cap=1 << 21):
    """The oracle's slice iteration: (offsets of the slices in iteration order, strides of the
    retained axes in V's order)."""
    d = np.ascontiguousarray(dims, dtype=np.int32)
    v = np.ascontiguousarray(V, dtype=np.int32)
    off = np.zeros(cap, np.int64)
    vs = np.zeros(max(len(V), 1), np.int64)
    n = lib().orc_slices(d.ctypes.data, len(d), v.ctypes.data, len(v), off.ctypes.data, cap, vs.ctypes.data)
    if n < 0:
        raise RuntimeError(lib().orc_last_error().decode())
    return off[:min(n, cap)], vs[:len(V)]
