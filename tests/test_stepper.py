"""The oracle's ODE engine on the reference's own stepper tests (CPU only).

CPPQEDutils/Testing/EvolvedTest.cc, HarmonicOscillatorComplex.cc, HarmonicOscillatorRandomTimes.cc
and the GSL twins EvolvedTestGSL.cc, HarmonicOscillatorComplexGSL.cc drive a classical ODE through
the same ``ode_engine::Base::step`` and controlled RKCK54 stepper as the quantum trajectories,
and compare with closed-form solutions at the thresholds those files return on.  They are the
only checks the reference has of the GSL controller; for odeint's they add the reverse-direction
guard (``dtTry = -dtDid``, ODE.h:241), which advancing to random earlier times exercises.

The closed forms are textbook: exp(exp(Z t)) solves the Evolved equation (EvolvedTest.cc:4-15);
the driven damped oscillator y'' + 2 gamma y' + y = exp(i omega t) is the homogeneous solution of
the characteristic roots -gamma +- i sqrt(1 - gamma^2) plus the particular exp(i omega t) /
(1 - omega^2 + 2 i gamma omega), matched to y(0), y'(0).
"""
import numpy as np
import pytest

import oracle_lib as O

Z = -1 + 10j
OMEGA, GAMMA = 5.0, 0.1
Y0, DY0 = 10 + 5j, -1 + 1j


def rel_dev(a, b):
    """relativeDeviation, MathExtensions.h:50"""
    return np.abs(a - b) / (np.abs(a) + np.abs(b))


def oscillator(t):
    """Exact y(t), y'(t) of y'' + 2 gamma y' + y = exp(i omega t) from y(0) = Y0, y'(0) = DY0."""
    t = np.asarray(t, dtype=np.float64)
    c = 1 / (1 - OMEGA ** 2 + 2j * GAMMA * OMEGA)          # particular solution c exp(i omega t)
    sp, sm = -GAMMA + 1j * np.sqrt(1 - GAMMA ** 2), -GAMMA - 1j * np.sqrt(1 - GAMMA ** 2)
    # A + B = Y0 - c, sp A + sm B = DY0 - i omega c
    a = ((DY0 - 1j * OMEGA * c) - sm * (Y0 - c)) / (sp - sm)
    b = (Y0 - c) - a
    y = a * np.exp(sp * t) + b * np.exp(sm * t) + c * np.exp(1j * OMEGA * t)
    dy = sp * a * np.exp(sp * t) + sm * b * np.exp(sm * t) + 1j * OMEGA * c * np.exp(1j * OMEGA * t)
    return y, dy


def test_oscillator_closed_form_is_a_solution():
    """The closed form itself: initial values, and the equation by central differences."""
    y, dy = oscillator(0.0)
    assert abs(y - Y0) < 1e-13 and abs(dy - DY0) < 1e-13
    t, h = np.linspace(0.3, 60, 50), 1e-4
    ddy = (oscillator(t + h)[1] - oscillator(t - h)[1]) / (2 * h)
    y, dy = oscillator(t)
    np.testing.assert_allclose(ddy, np.exp(1j * OMEGA * t) - 2 * GAMMA * dy - y, atol=1e-6)


@pytest.mark.parametrize("controller,threshold", [("odeint", 2e-8), ("gsl", 5e-8)])
def test_evolved(controller, threshold):
    """EvolvedTest.cc:32-49 / EvolvedTestGSL.cc:32-55: Dt = 0.01 to T = 5, dtInit = 1/(10|Z|),
    user tolerances (1e-6, 1e-18); mean relative deviation of y from exp(exp(Z t))."""
    rows, st = O.simulated_run("evolved", controller, dtInit=1 / (10 * abs(Z)), T=5.0, Dt=0.01)
    # 501 stream points; the sum of the 0.01 intervals falls short of 5 by rounding, so the run
    # loop adds a last interval of that remainder (Trajectory.h:560-567) and a 502nd row
    assert len(rows) == 502 and rows[-1, 0].real == 5.0 and rows[-1, 1].real < 1e-12
    np.testing.assert_allclose(rows[:501, 0].real, np.linspace(0, 5, 501), rtol=0, atol=1e-12)
    dev = rel_dev(rows[:, 2], np.exp(np.exp(Z * rows[:, 0].real))).mean()
    assert dev < threshold, dev
    if controller == "gsl":
        assert st[2] == 0 and st[0] == 7 * st[3]     # never rejects; dydt_out is a seventh call
    else:
        assert st[0] == 6 * st[3] and st[3] == st[1] + st[2]


@pytest.mark.parametrize("controller", ["odeint", "gsl"])
def test_driven_damped_oscillator(controller):
    """HarmonicOscillatorComplex.cc:19-61 / HarmonicOscillatorComplexGSL.cc: run(S, 70., 1, ...) has an
    integer stream frequency, i.e. dc mode: a row after every ODE step to T = 70; dtInit =
    0.1/max(1, omega, gamma); mean relative deviations below 2e-6 (y) and 8e-7 (y').  The GSL case
    fixes hadjust's order at 5 (with 4 both means exceed the thresholds, as EvolvedTestGSL's does)."""
    rows, st = O.simulated_run("oscillator", controller, dtInit=0.1 / max(1.0, OMEGA, GAMMA), T=70.0, dc=1)
    assert len(rows) == st[1] + 1 and rows[-1, 0].real == 70.0
    t = rows[:, 0].real
    y, dy = oscillator(t)
    amp, ampd = rel_dev(rows[:, 2], y).mean(), rel_dev(rows[:, 3], dy).mean()
    assert amp < 2e-6 and ampd < 8e-7, (amp, ampd)


def test_mt19937_sample_times():
    """std::mt19937 + uniform_real_distribution as libstdc++ draws them: the raw stream is numpy's
    legacy MT19937 with the same integer seed (first output for seed 5489 = 3499211612, the C++
    standard's generator), combined two 32-bit words per double (generate_canonical<double, 53>)."""
    raw = np.random.RandomState(1001).randint(0, 2 ** 32, size=20, dtype=np.uint64).astype(np.float64)
    u = (raw[0::2] + raw[1::2] * 2.0 ** 32) / 2.0 ** 64
    np.testing.assert_array_equal(O.mt19937_uniform(1001, 10, 0.0, 70.0), 0.0 + 70.0 * u)
    assert np.random.RandomState(5489).randint(0, 2 ** 32, dtype=np.uint64) == 3499211612


def test_oscillator_random_times():
    """HarmonicOscillatorRandomTimes.cc:17-57: advanceTo 1000 uniform times in [0, 70) from
    std::mt19937{1001}; about half of them lie behind the current time, so the engine reverses
    direction (ODE.h:241).  Mean relative deviations below 3e-4 (y) and 4e-4 (y')."""
    times = O.mt19937_uniform(1001, 1000, 0.0, 70.0)
    assert (np.diff(times) < 0).sum() > 400
    rows, st = O.simulated_run("oscillator", "odeint", dtInit=0.1 / max(1.0, OMEGA, GAMMA), times=times)
    assert len(rows) == 1000
    # advanceTo reaches t + (time - t), which may differ from time in the last bit (Trajectory.h:64-69)
    np.testing.assert_allclose(rows[:, 0].real, times, rtol=0, atol=1e-13)
    y, dy = oscillator(times)
    amp, ampd = rel_dev(rows[:, 2], y).mean(), rel_dev(rows[:, 3], dy).mean()
    assert amp < 3e-4 and ampd < 4e-4, (amp, ampd)
    assert (rows[:, 1].real < 0).sum() > 400          # steps taken backwards (dtDid < 0)
