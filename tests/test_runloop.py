"""The run loop around the path: the reference's autostop criterion, AutostopHandlerGeneric
(CPPQEDutils/Trajectory.h:179-222),
in the oracle and in the host run loop (CPU only).

The known answers below follow from the handler's definition by hand: the first call only
starts the running average, the next `repetition` calls fill the queue, and only then is a
streamed array compared (max over components of |avg - v| / (|avg| + |v|) < epsilon), with
entries below epsilonAbs replaced by epsilonAbs.
"""
import numpy as np
import pytest

import oracle_lib as O
from cppqed_amd.trajectory import AutostopHandler
from sysdefs import PLM20, ground

IMPLS = {"oracle": O.Autostop, "host": AutostopHandler}


def first_stop(handler, seq):
    for i, v in enumerate(seq):
        if handler(np.atleast_1d(np.asarray(v, dtype=np.float64))):
            return i
    return None


@pytest.mark.parametrize("impl", IMPLS)
def test_constant_stream_stops_after_first_call_and_full_queue(impl):
    # call 0 starts the average, calls 1-3 fill the queue (repetition 3), call 4 compares
    assert first_stop(IMPLS[impl](1e-6, 1e-12, 3), [[2.0, -1.0]] * 10) == 4


@pytest.mark.parametrize("impl", IMPLS)
def test_small_entries_are_clamped(impl):
    # +-1e-15 all read as 1e-12: identical arrays, stop at call 4; unclamped the ratio stays 1
    seq = [[1e-15 * (-1) ** i] for i in range(10)]
    assert first_stop(IMPLS[impl](1e-6, 1e-12, 3), seq) == 4
    assert first_stop(IMPLS[impl](1e-6, 1e-18, 3), seq) is None


@pytest.mark.parametrize("impl", IMPLS)
def test_moving_average_known_answer(impl):
    # repetition 2, epsilon 0.1: call 0 (100) only starts; queue [1, 1] (avg 1); call 3 (2): 1/3,
    # avg -> 1.5; call 4 (2): 0.5/3.5 = 0.143, avg -> 2; call 5 (2): 0 -> stop
    assert first_stop(IMPLS[impl](0.1, 1e-12, 2), [100, 1, 1, 2, 2, 2, 2, 2]) == 5


@pytest.mark.parametrize("impl", IMPLS)
def test_every_component_must_converge(impl):
    # component 2 grows linearly: from call 4 on avg = i - 2, ratio 1/(i-1) < 0.05 first at i = 22
    seq = [[1.0, float(i)] for i in range(60)]
    assert first_stop(IMPLS[impl](0.05, 1e-12, 3), seq) == 22
    assert first_stop(IMPLS[impl](1e-3, 1e-12, 3), seq) is None


@pytest.mark.parametrize("impl", IMPLS)
def test_repetition_zero_never_stops(impl):
    assert first_stop(IMPLS[impl](1e-6, 1e-12, 0), [[1.0]] * 50) is None


def test_master_autostop_truncates_the_run():
    """trajectory::run with the handler (Trajectory.h:583-589): the run ends on the row where the
    criterion is met; up to there it is the run without the handler, and the final state is the
    state at that row."""
    s = O.System(PLM20)
    p = O.params(epsRel=1e-8, epsAbs=1e-12)
    rho0 = np.outer(ground(20), ground(20)).astype(complex)
    r = O.run_spec("Dt", Dt=0.1, T=30.0)
    full, _ = O.master_run(s, rho0, p, r)
    rows, rho = O.master_run_autostop(s, rho0, p, r, 1e-5, 1e-12, 5)
    assert 10 < len(rows) < len(full)
    np.testing.assert_array_equal(rows, full[:len(rows)])
    h = AutostopHandler(1e-5, 1e-12, 5)
    assert first_stop(h, [row[2:] for row in full]) == len(rows) - 1
    upto, rho_upto = O.master_run(s, rho0, p, O.run_spec("Dt", Dt=0.1, T=rows[-1, 0]))
    np.testing.assert_array_equal(upto, rows)
    np.testing.assert_array_equal(rho_upto, rho)


def test_row_format_reproduces_the_reference_files():
    """trajectory::stream's text (Trajectory.h:426-434, FormDouble): every data row of the golden
    files, re-streamed from its own values, is byte-identical (time and dtDid in 13 columns, a
    tab before each element's four averages, averages in 14 columns, 6 significant digits)."""
    import glob
    import os

    from cppqed_amd.trajectory import format_row
    files = glob.glob(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "run_*"))
    n = 0
    for f in files:
        for ln in open(f):
            if ln.startswith("#") or not ln.strip():
                continue
            v = [float(x) for x in ln.split()]
            assert format_row(v, groups=[4] * ((len(v) - 2) // 4)) == ln
            n += 1
    assert n > 400
