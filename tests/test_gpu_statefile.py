"""Checkpoint and resume on the GPU through the state-file format of the reference
(trajectory::run with a .state file, CPPQEDutils/Trajectory.h:484-500; MCWF_Trajectory.h:134-147;
Master.h:89-102).

* A device ensemble saved at a stream point and loaded into a fresh ensemble continues bit for
  bit as the uninterrupted ensemble does (the reference's Testing/continue check: a run split in
  two equals the whole run).
* The reference's own Master archives (tests/golden/state) continued on the device reproduce the
  reference's remaining streamed rows and its final density operator.
"""
import os

import numpy as np
import pytest

import cppqed_amd as Q
from cppqed_amd import statefile as SF
from sysdefs import JC60, P1M1, PTLA, ground, p1m1_initial
from test_oracle import load_golden

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _records_equal(a, b):
    assert len(a) == len(b)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x.psi, y.psi)
        assert (x.t, x.dtDid, x.dtTry, x.rng) == (y.t, y.dtDid, y.dtTry, y.rng)
        assert (x.ode_derivs, x.ode_steps, x.ode_failed) == (y.ode_derivs, y.ode_steps, y.ode_failed)
        assert (x.mcwf_steps, x.overshot, x.tolerance_overshot) == (y.mcwf_steps, y.overshot, y.tolerance_overshot)
        assert x.jumps == y.jumps


@pytest.mark.parametrize("name,spec,psi0,ntraj,prec", [
    pytest.param("jc60", JC60, ground(60), 256, Q.QT_FP64, id="jc60"),
    pytest.param("jc60_fp32", JC60, ground(60), 256, Q.QT_FP32_DERIV, id="jc60_fp32"),
    # interaction picture: t0 = t on load
    pytest.param("p1m1_ip", P1M1, p1m1_initial(), 64, Q.QT_FP64, id="p1m1_ip"),
    # one trajectory: an MCWF_Trajectory archive
    pytest.param("single", JC60, ground(60), 1, Q.QT_FP64, id="single"),
])
def test_ensemble_split_run_equals_whole_run(tmp_path, name, spec, psi0, ntraj, prec):
    print(f"[gpu] checkpoint/resume {name}", flush=True)
    s = Q.System(spec)
    iv = np.full(12, 0.05)
    pars = Q.mcwf_pars(seed=99, jump_log_cap=64, precision=prec)
    whole = Q.Ensemble(s, ntraj, psi0, pars)
    whole.advance(iv[:5])
    path = str(tmp_path / "run.state")
    whole.save_state(path)
    saved = whole.records()
    rows_whole = whole.advance(iv[5:])

    part = Q.Ensemble(s, ntraj, psi0, pars)
    fr = part.load_state(path)
    assert fr.trajectory_id == (SF.MCWF_ID if ntraj == 1 else SF.ENSEMBLE_ID)
    _records_equal(part.records(), saved)
    rows_part = part.advance(iv[5:])
    np.testing.assert_array_equal(rows_part, rows_whole)
    _records_equal(part.records(), whole.records())
    per, _ = part.counters()
    assert (per[:, Q.C_TRIES] == per[:, Q.C_ODE] + per[:, Q.C_FAILED]).all()
    assert per[:, Q.C_MCWF].sum() > 0


def test_ensemble_load_checks_streams(tmp_path):
    s = Q.System(JC60)
    a = Q.Ensemble(s, 8, ground(60), Q.mcwf_pars(seed=5, jump_log_cap=16))
    a.advance([0.1, 0.1])
    path = str(tmp_path / "a.state")
    a.save_state(path)
    b = Q.Ensemble(s, 8, ground(60), Q.mcwf_pars(seed=6, jump_log_cap=16))
    with pytest.raises(SF.StateFileError):
        b.load_state(path)                     # another seed: not this ensemble's streams
    b.load_state(path, rng="restart")
    st, psi, _, _ = b.get_states()
    st_a, psi_a, _, _ = a.get_states()
    np.testing.assert_array_equal(psi, psi_a)
    assert all(x.ordinal == 0 and x.t == y.t for x, y in zip(st, st_a))
    c = Q.Ensemble(s, 4, ground(60), Q.mcwf_pars(seed=5, jump_log_cap=16))
    with pytest.raises(SF.StateFileError):
        c.load_state(path)                     # trajectory count differs


def test_reference_ensemble_archive_loads(tmp_path):
    """The reference's pcg64 ensemble archive (PTLA, 3 trajectories): state, times and jump lists
    reach the device unchanged; continuing its pcg64 streams is refused."""
    fr = SF.read(os.path.join(HERE, "golden", "state", "run_PTLA_CPPQED_dt_ensemble.state"))
    path = str(tmp_path / "ref.state")
    SF.write(path, fr[2:3])
    ens = Q.Ensemble(Q.System(PTLA), 3, ground(2), Q.mcwf_pars(epsAbs=1e-5, jump_log_cap=16))
    with pytest.raises(SF.StateFileError):
        ens.load_state(path)
    ens.load_state(path, rng="restart")
    for got, ref in zip(ens.records(), fr[2].trajs):
        np.testing.assert_array_equal(got.psi, ref.psi)
        assert (got.t, got.dtDid, got.dtTry, got.jumps) == (ref.t, ref.dtDid, ref.dtTry, ref.jumps)
        assert (got.ode_derivs, got.ode_steps, got.mcwf_steps, got.overshot) == \
            (ref.ode_derivs, ref.ode_steps, ref.mcwf_steps, ref.overshot)


@pytest.mark.parametrize("mode", ["dt", "ndt"])
def test_master_continues_reference_archive(tmp_path, mode):
    """run_PTLA_CPPQED_{dt,ndt}_master.state: every saved instant, continued on the device's
    Master sweep, streams the reference's remaining rows and ends on its final archive."""
    frames = SF.read(os.path.join(HERE, "golden", "state", f"run_PTLA_CPPQED_{mode}_master.state"))
    g = load_golden(f"run_PTLA_CPPQED_{mode}_master")
    last = frames[-1].master
    s = Q.System(PTLA)
    for k in range(len(frames) - 1):
        print(f"[gpu] master resume PTLA {mode} frame {k}", flush=True)
        path = str(tmp_path / f"m{k}.state")
        SF.write(path, frames[k:k + 1], append=False)
        m = Q.MasterSweep([s], np.diag([1.0, 0.0]), epsRel=1e-6, epsAbs=1e-5)
        m.load_state(path)
        # the continuing run's intervals: Dt mode to T = 1, NDt mode the remaining intervals of 0.1
        t, iv = frames[k].master.t, []
        if mode == "dt":
            while t < 1.0:
                iv.append(min(0.1, 1.0 - t))
                t += iv[-1]
        else:
            iv = [0.1] * (10 - 3 * k)
        rows, _ = m.advance(iv)
        rest = g[len(g) - len(iv):]
        np.testing.assert_allclose(rows[0], rest, rtol=1e-5, atol=1e-7)
        st, rho = m.get_point(0)
        np.testing.assert_allclose(rho, last.matrix(), rtol=0, atol=1e-10)
        assert abs(st.t - last.t) <= 1e-12
        out = str(tmp_path / f"o{k}.state")
        m.save_state(out)
        back = SF.read_last(out).master
        np.testing.assert_array_equal(back.matrix(), rho)
        assert back.ode_steps == st.counters[Q.C_ODE]
