"""Trajectory state files: the codec against the reference's own archives, and continuing
a run from them (CPU only).

Fixtures: tests/golden/state/*.state are the reference's Testing/expected/*.state files of
the PTLA, 1particle1mode and Ring runs (bzip2 streams of Boost binary archives written at
every third stream point and at the end, --sdf 3).  The streamed rows they belong to are the
golden files beside them.
"""
import glob
import os

import numpy as np
import pytest

import oracle_lib as O
from cppqed_amd import statefile as SF
from sysdefs import P1M1, P1M1_CONF, P1M1_MASTER, PTLA, RING, RING_MASTER
from test_oracle import GOLDEN_MODES, P1M1_MASTER_MODES, P1M1_MODES, RING_MASTER_MODES, load_golden

HERE = os.path.dirname(os.path.abspath(__file__))
STATES = sorted(glob.glob(os.path.join(HERE, "golden", "state", "*.state")))
SDF = 3  # state-dump frequency of the golden runs (Testing/run/testdriver.conf)


def test_fixtures_present():
    assert len(STATES) == 36


@pytest.mark.parametrize("path", STATES, ids=[os.path.basename(p) for p in STATES])
def test_archives_round_trip_byte_exact(path):
    """Every archive decodes and re-encodes to the same bytes (Trajectory.h:353-410)."""
    archives = SF.split_frames(SF.read_stream(path))
    assert archives
    for a in archives:
        fr = SF.decode_frame(a)
        assert SF.encode_frame(fr) == a
    # the file-level writer reproduces the decompressed stream
    assert SF.join_frames(archives) == SF.read_stream(path)


def test_write_read_file(tmp_path):
    fr = SF.read(os.path.join(HERE, "golden", "state", "run_PTLA_CPPQED_dt_ensemble.state"))
    p = str(tmp_path / "x.state")
    SF.write(p, fr[:2])
    SF.write(p, fr[2:])                       # appending adds a second bzip2 stream
    back = SF.read(p)
    assert [SF.encode_frame(f) for f in back] == [SF.encode_frame(f) for f in fr]
    assert SF.encode_frame(SF.read_last(p)) == SF.encode_frame(fr[-1])


def test_array_only_and_rejections():
    a = (np.arange(6) + 1j).reshape(2, 3)
    fr = SF.decode_frame(SF.encode_frame(SF.Frame(trajectory_id=SF.ARRAY_ONLY, array=a)))
    np.testing.assert_array_equal(fr.array, a)
    b = bytearray(SF.encode_frame(SF.Frame(trajectory_id=SF.ARRAY_ONLY, array=a)))
    with pytest.raises(SF.StateFileError):
        SF.decode_frame(bytes(b[:-3]))
    with pytest.raises(SF.StateFileError):
        SF.split_frames(b"12abc")
    assert SF.parse_philox(SF.format_philox(7, 9, 11)) == (7, 9, 11)
    assert SF.parse_philox("1 2 3") is None


# ---------------------------------------------------------------- the runs the files belong to

def _setup(system, kind):
    p = O.params(epsRel=1e-6, epsAbs=1e-5, rng="pcg64", seed=1001, stream=1)
    if system.startswith("1particle1mode_ndt_conf"):
        spec = P1M1_CONF[int(system[-1])]
        if kind == "master":
            return (O.System(spec.replace("eta=10 picture=IP", "eta=10 picture=UIP")), p,
                    {"dc": P1M1_MASTER_MODES["dc"]})
        return O.System(spec), p, {"dc": P1M1_MODES["dc"]}
    if system == "PTLA_CPPQED":
        return O.System(PTLA), p, GOLDEN_MODES
    if system == "1particle1mode":
        return O.System(P1M1_MASTER if kind == "master" else P1M1), p, (P1M1_MASTER_MODES if kind == "master"
                                                                         else P1M1_MODES)
    return O.System(RING_MASTER if kind == "master" else RING), p, (RING_MASTER_MODES if kind == "master"
                                                                   else P1M1_MODES)


def _continuing_run(run, frame_index):
    """The run a continuing invocation performs: same end time (Dt, dc) or the remaining
    NDt intervals (timeToReach = t + NDt*Dt, Trajectory.h:502-503)."""
    if run.mode == 2:
        return O.Run(run.mode, run.dc, run.Dt, run.NDt - SDF * frame_index, run.T)
    return run


def _assert_traj_equal(a, b, tol):
    np.testing.assert_allclose(a.psi, b.psi, rtol=0, atol=tol)
    for f in ("t", "dtDid", "dtTry"):
        assert abs(getattr(a, f) - getattr(b, f)) <= tol * max(1.0, abs(getattr(b, f))), f
    assert (a.ode_derivs, a.ode_steps, a.ode_failed) == (b.ode_derivs, b.ode_steps, b.ode_failed)
    assert (a.mcwf_steps, a.tolerance_overshot) == (b.mcwf_steps, b.tolerance_overshot)
    # A dpLimit overshoot is counted when totalRate*dtTry > dpLimit (MCWF_Trajectory.h:309-311); right
    # after dtTry was set to dpLimit/totalRate both sides agree to rounding, so where the oracle is not
    # bit-identical to the reference (the interaction-picture systems) the count may differ by one
    # such tie without any effect on the dynamics (dtTry = min(dtTry, dpLimit/totalRate) either way).
    if tol <= 1e-12:
        assert a.overshot == b.overshot
    else:
        assert abs(a.overshot - b.overshot) <= 2
    assert a.rng == b.rng
    assert [c for _, c in a.jumps] == [c for _, c in b.jumps]
    np.testing.assert_allclose([x for x, _ in a.jumps], [x for x, _ in b.jumps], rtol=tol, atol=0)
    for f in ("dp_max_overshoot", "dp_tolerance_max_overshoot", "norm_max_deviation"):
        assert abs(getattr(a, f) - getattr(b, f)) <= 1e-6 * max(1e-3, abs(getattr(b, f))), f


RUNS = ([(s, m) for s in ("PTLA_CPPQED", "1particle1mode", "Ring") for m in ("dc", "dt", "ndt")]
        + [(f"1particle1mode_ndt_conf{c}", "dc") for c in (2, 3, 4)])


def _stem(system, mode, kind):
    if system.startswith("1particle1mode_ndt_conf"):
        return f"run_{system}_{kind}"
    return f"run_{system}_{mode}_{kind}"


@pytest.mark.parametrize("system,mode", RUNS)
def test_single_continues_from_reference_state(system, mode):
    """MCWF_Trajectory archives (MCWF_Trajectory.h:134-147): the oracle continued from every
    saved instant reaches the reference's final archive -- state, time, dtDid, dtTry, both
    loggers, jump list and the pcg64 state -- and streams the reference's remaining rows."""
    s, p, modes = _setup(system, "single")
    frames = SF.read(os.path.join(HERE, "golden", "state", _stem(system, mode, "single") + ".state"))
    g = load_golden(_stem(system, mode, "single"))
    last = frames[-1].trajs[0]
    assert last.t == pytest.approx(g[-1, 0], rel=1e-5)
    tol = 1e-12 if system == "PTLA_CPPQED" else 1e-9
    for k in range(len(frames) - 1):
        rec = frames[k].trajs[0]
        rows, out = O.mcwf_resume(s, rec, p, _continuing_run(modes[mode], k))
        _assert_traj_equal(out, last, tol)
        if mode != "dc":  # dc streams every dc-th step counted from the restart
            rest = g[np.searchsorted(g[:, 0], rec.t * (1 + 1e-12) + 1e-300, side="right"):]
            assert rows.shape == rest.shape
            tcol = np.full(g.shape[1], 1e-8)
            tcol[-1] = 1e-6
            assert not (np.abs(rows - rest) > 1e-5 * np.abs(rest) + tcol).any()


@pytest.mark.parametrize("system,mode", RUNS)
def test_ensemble_continues_from_reference_state(system, mode):
    """Ensemble archives hold every trajectory's stateIO in order (StochasticTrajectory.h:238);
    each continued member reaches its record in the final archive."""
    s, p, modes = _setup(system, "ensemble")
    frames = SF.read(os.path.join(HERE, "golden", "state", _stem(system, mode, "ensemble") + ".state"))
    assert all(len(f.trajs) == 3 for f in frames)
    run = modes[mode] if mode != "dc" else (GOLDEN_MODES["dt"] if system == "PTLA_CPPQED"
                                            else O.run_spec("Dt", Dt=0.1, T=0.1))
    tol = 1e-12 if system == "PTLA_CPPQED" else 1e-9
    k = len(frames) // 2
    for j in range(3):
        _, out = O.mcwf_resume(s, frames[k].trajs[j], p, _continuing_run(run, k))
        _assert_traj_equal(out, frames[-1].trajs[j], tol)


@pytest.mark.parametrize("system,mode", RUNS)
def test_master_continues_from_reference_state(system, mode):
    """Master archives (Master.h:89-102): continued from each saved instant the oracle reaches
    the final density operator, time step and ODE logger, and streams the remaining rows."""
    s, p, modes = _setup(system, "master")
    frames = SF.read(os.path.join(HERE, "golden", "state", _stem(system, mode, "master") + ".state"))
    g = load_golden(_stem(system, mode, "master"))
    last = frames[-1].master
    tol = 1e-12 if system == "PTLA_CPPQED" else 1e-9
    for k in range(len(frames) - 1):
        rec = frames[k].master
        rows, out = O.master_resume(s, rec, p, _continuing_run(modes[mode], k))
        np.testing.assert_allclose(out.rho, last.rho, rtol=0, atol=tol)
        assert abs(out.t - last.t) <= tol * max(1, abs(last.t))
        assert abs(out.dtTry - last.dtTry) <= tol * abs(last.dtTry)
        assert (out.ode_derivs, out.ode_steps, out.ode_failed) == (last.ode_derivs, last.ode_steps, last.ode_failed)
        if mode != "dc":
            rest = g[np.searchsorted(g[:, 0], rec.t * (1 + 1e-12) + 1e-300, side="right"):]
            assert rows.shape == rest.shape
            assert not (np.abs(rows - rest) > 1e-5 * np.abs(rest) + 1e-6).any()


@pytest.mark.parametrize("system,kind", [(s, k) for s in ("PTLA_CPPQED", "1particle1mode", "Ring")
                                         for k in ("single", "master")])
def test_saved_states_reproduce_streamed_rows(system, kind):
    """Each archived state, put through the averages, gives the row streamed at its time."""
    s, _, _ = _setup(system, kind)
    for mode in ("dt", "ndt"):
        frames = SF.read(os.path.join(HERE, "golden", "state", f"run_{system}_{mode}_{kind}.state"))
        g = load_golden(f"run_{system}_{mode}_{kind}")
        for fr in frames:
            rec = fr.trajs[0] if kind == "single" else fr.master
            i = int(np.argmin(np.abs(g[:, 0] - rec.t)))
            assert abs(g[i, 0] - rec.t) <= 1e-5 * max(abs(rec.t), 1e-12) + 1e-12
            if kind == "single":
                av = s.averages(rec.psi.reshape(-1))
                np.testing.assert_allclose(av, g[i, 2:], rtol=1e-5, atol=1e-6)
            else:
                rho = rec.matrix()
                assert abs(np.trace(rho) - 1) < 1e-12
                np.testing.assert_allclose(rho, rho.conj().T, atol=1e-14)
