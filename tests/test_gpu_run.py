"""The host run loop over a device ensemble (trajectory::run, CPPQEDutils/Trajectory.h:462-610):
state files every sdf stream points, continuing from the last archive, autostop."""
import numpy as np
import pytest

import cppqed_amd as Q
from cppqed_amd import statefile as SF
from cppqed_amd.trajectory import AutostopHandler, run
from sysdefs import JC60, PLM20, ground

pytestmark = pytest.mark.gpu


def test_run_state_dumps_and_continuation(tmp_path):
    print("[gpu] run loop: state dumps and continuation", flush=True)
    s = Q.System(JC60)
    pars = Q.mcwf_pars(seed=11, jump_log_cap=32)
    path = str(tmp_path / "jc.d.state")
    rows = run(Q.Ensemble(s, 128, ground(60), pars), T=2.0, Dt=0.1, sdf=3, state_file=path)
    assert rows.shape == (21, 2 + s.n_avr)
    np.testing.assert_allclose(rows[:, 0], np.arange(21) * 0.1, rtol=0, atol=1e-12)
    frames = SF.read(path)
    # stream points 0, 3, ..., 18, then the end (20): Trajectory.h:569-576, 599
    assert [round(f.trajs[0].t, 9) for f in frames] == [round(0.1 * c, 9) for c in (0, 3, 6, 9, 12, 15, 18, 20)]
    # continue from the archive at t = 0.9: the remaining rows are the uninterrupted run's
    part = str(tmp_path / "part.d.state")
    SF.write(part, frames[:4])
    rest = run(Q.Ensemble(s, 128, ground(60), pars), T=2.0, Dt=0.1, state_file=part)
    np.testing.assert_array_equal(rest, rows[10:])


def test_run_autostop_truncates(tmp_path):
    print("[gpu] run loop: autostop", flush=True)
    s = Q.System(PLM20)
    pars = Q.mcwf_pars(seed=3)
    full = run(Q.Ensemble(s, 4096, ground(20), pars), T=8.0, Dt=0.1)
    rows = run(Q.Ensemble(s, 4096, ground(20), pars), T=8.0, Dt=0.1, autostop=(2e-2, 1e-12, 4))
    assert 6 < len(rows) < len(full)
    np.testing.assert_array_equal(rows, full[:len(rows)])
    h = AutostopHandler(2e-2, 1e-12, 4)
    stops = [h(r[2:]) for r in full]
    assert stops.index(True) == len(rows) - 1


def test_run_writes_the_reference_text_format(tmp_path):
    """trajectory_file: rows in the reference's text format, appended by a continuing run."""
    print("[gpu] run loop: trajectory file", flush=True)
    s = Q.System(JC60)
    pars = Q.mcwf_pars(seed=17, jump_log_cap=16)
    path = str(tmp_path / "jc.d")
    first = run(Q.Ensemble(s, 64, ground(60), pars), T=1.0, Dt=0.1, trajectory_file=path)
    rest = run(Q.Ensemble(s, 64, ground(60), pars), T=2.0, Dt=0.1, trajectory_file=path)
    text = open(path).read()
    assert "# Continuing from time 1 up to time 2" in text
    data = np.array([[float(x) for x in ln.split()] for ln in text.splitlines() if ln and not ln.startswith("#")])
    # ten steps of 0.1 end 1.1e-16 short of 1: the first run streams that tail interval too, as
    # the reference's golden files show ("1  1.11022e-16")
    assert data.shape == (len(first) + len(rest), 2 + s.n_avr) and len(rest) == 10
    np.testing.assert_allclose(data, np.vstack([first, rest]), rtol=1e-5, atol=1e-12)
