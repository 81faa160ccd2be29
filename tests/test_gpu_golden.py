"""The reference's own MCWF runs replayed on the device (HIP batched kernel, FP64).

Each golden file of Testing/expected (copied as data into tests/golden/) was produced by the
reference with pcg64 seed 1001, stream 1 for a single trajectory and streams 1, 2, 3 for the
ensembles of three (randomutils::Pars / incrementForNextStream, CPPQEDutils/Random.h:137-161),
uniforms through libstdc++'s generate_canonical.  The oracle restates that generator (pinned
against numpy's PCG64); its uniforms are handed to the device ensemble as host-supplied jump
draws (qt_mcwf_set_draws), and the device's streamed rows are compared with the golden rows at
the reference's own tolerance (np.allclose defaults, Testing/testdriver.py:395-400).  The dc-mode
singles (a row every dc MCWF steps) run through qt_mcwf_advance_dc; an ensemble's dc mode is its
Dt mode (an Ensemble has no step(), Trajectory.h:317-323)."""
import os

import numpy as np
import pytest

import cppqed_amd as Q
import oracle_lib as O
from sysdefs import P1M1, P1M1_CONF, PTLA, RING, ground, p1m1_initial, ring_initial

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
N_DRAWS = 60000   # uniforms per trajectory: more MCWF steps than any golden run takes


def load_golden(name):
    rows = []
    with open(os.path.join(GOLDEN, name)) as f:
        for ln in f:
            if ln.startswith("#") or not ln.strip():
                continue
            rows.append([float(x) for x in ln.split()])
    return np.array(rows)


def device_run(spec, psi0, n_traj, intervals=None, dc=None, T=None):
    s = Q.System(spec)
    e = Q.Ensemble(s, n_traj, psi0, Q.mcwf_pars(epsRel=1e-6, epsAbs=1e-5, jump_log_cap=4096))
    draws = np.stack([O.pcg64_uniform(1001, 1 + k, N_DRAWS)[0] for k in range(n_traj)])
    e.set_draws(draws)
    first = e.advance([0.0])   # the stream point at t = 0 (no step: no draw)
    if dc:
        return np.concatenate([first, e.advance_dc(T, dc, max_rows=7)])   # several calls per run
    return np.concatenate([first, e.advance(intervals)])


def compare(rows, g, dev_col=None):
    assert rows.shape == g.shape, (rows.shape, g.shape)
    tol = np.full(g.shape[1], 1e-8)
    if dev_col is not None:
        tol[dev_col] = 1e-6   # the particle's DEV(X): root of a cancellation for a one-grid-point state
    bad = np.abs(rows - g) > 1e-5 * np.abs(g) + tol
    assert not bad.any(), (np.argwhere(bad)[:5], rows[bad][:5], g[bad][:5])


PTLA_IV = {"dt": Q.run_intervals(T=1.0, Dt=0.1), "ndt": Q.run_intervals(Dt=0.1, NDt=10)}
P1M1_IV = {"dt": Q.run_intervals(T=0.1, Dt=0.01), "ndt": Q.run_intervals(Dt=0.01, NDt=10)}


@pytest.mark.parametrize("mode", ["dc", "dt", "ndt"])
@pytest.mark.parametrize("kind", ["single", "ensemble"])
def test_ptla_golden_on_device(mode, kind):
    """run_PTLA_CPPQED_{dc,dt,ndt}_{single,ensemble}: jumps, the controller's dt sequence (dtDid
    column) and the RNG consumption of the reference, on the device (--dc 10 for dc)."""
    print(f"[gpu] golden PTLA {mode} {kind}", flush=True)
    if kind == "single" and mode == "dc":
        rows = device_run(PTLA, ground(2), 1, dc=10, T=1.0)
    else:
        rows = device_run(PTLA, ground(2), 1 if kind == "single" else 3, PTLA_IV["dt" if mode == "dc" else mode])
    compare(rows, load_golden(f"run_PTLA_CPPQED_{mode}_{kind}"))


P1M1_IV["dc"] = Q.run_intervals(T=0.1, Dt=0.1)   # an ensemble's dc mode: its Dt mode


@pytest.mark.parametrize("mode", ["dc", "dt", "ndt"])
@pytest.mark.parametrize("kind", ["single", "ensemble"])
def test_1particle1mode_golden_on_device(mode, kind):
    """run_1particle1mode_{dc,dt,ndt}_{single,ensemble}: binary system in the interaction picture
    (pumped particle, particle-cavity interference term, Exact k^2 phases, FFT-based <X>)."""
    print(f"[gpu] golden 1particle1mode {mode} {kind}", flush=True)
    if kind == "single" and mode == "dc":
        rows = device_run(P1M1, p1m1_initial(), 1, dc=1, T=0.1)
    else:
        rows = device_run(P1M1, p1m1_initial(), 1 if kind == "single" else 3, P1M1_IV[mode])
    g = load_golden(f"run_1particle1mode_{mode}_{kind}")
    compare(rows, g, dev_col=g.shape[1] - 1)


@pytest.mark.parametrize("conf", [2, 3, 4])
@pytest.mark.parametrize("kind", ["single", "ensemble"])
def test_1particle1mode_conf_golden_on_device(conf, kind):
    """run_1particle1mode_ndt_conf{2,3,4}_{single,ensemble} (1particle1mode --dc 1 --T 0.1: a row
    per MCWF step for the single, the ensemble's Dt mode with Dt = 0.1): ParticleAlongCavity's
    dispersive and interference terms."""
    print(f"[gpu] golden 1particle1mode conf{conf} {kind}", flush=True)
    if kind == "single":
        rows = device_run(P1M1_CONF[conf], p1m1_initial(), 1, dc=1, T=0.1)
    else:
        rows = device_run(P1M1_CONF[conf], p1m1_initial(), 3, Q.run_intervals(T=0.1, Dt=0.1))
    g = load_golden(f"run_1particle1mode_ndt_conf{conf}_{kind}")
    compare(rows, g, dev_col=g.shape[1] - 1)


@pytest.mark.parametrize("mode", ["dc", "dt", "ndt"])
@pytest.mark.parametrize("kind", ["single", "ensemble"])
def test_ring_golden_on_device(mode, kind):
    """run_Ring_{dc,dt,ndt}_{single,ensemble}: a composite with a three-body act on permuted axes."""
    print(f"[gpu] golden Ring {mode} {kind}", flush=True)
    if kind == "single" and mode == "dc":
        rows = device_run(RING, ring_initial(), 1, dc=1, T=0.1)
    else:
        rows = device_run(RING, ring_initial(), 1 if kind == "single" else 3, P1M1_IV[mode])
    compare(rows, load_golden(f"run_Ring_{mode}_{kind}"), dev_col=5)


def test_host_draws_run_out():
    """A trajectory that needs more host-supplied draws than it was given fails the advance
    instead of drawing from another generator; a reset replays the same supplied draws."""
    print("[gpu] host draws run out", flush=True)
    s = Q.System(PTLA)
    e = Q.Ensemble(s, 2, ground(2), Q.mcwf_pars(epsRel=1e-6, epsAbs=1e-5))
    e.set_draws(np.stack([O.pcg64_uniform(1001, 1 + k, 50)[0] for k in range(2)]))
    with pytest.raises(Q.QtError, match="more jump draws"):
        e.advance(PTLA_IV["dt"])
    e.reset()
    e.set_draws(np.stack([O.pcg64_uniform(1001, 1 + k, N_DRAWS)[0] for k in range(2)]))
    a = e.advance(PTLA_IV["dt"])
    e.reset()
    np.testing.assert_array_equal(e.advance(PTLA_IV["dt"]), a)
    e.set_draws(None)   # back to the counter RNG
    e.reset()
    b = e.advance(PTLA_IV["dt"])
    assert not np.array_equal(a, b)


def test_dc_mode_refusals():
    """dc mode drives one trajectory; t_end before the trajectory's time is refused; at t_end no row."""
    e = Q.Ensemble(Q.System(PTLA), 2, ground(2))
    with pytest.raises(Q.QtError, match="single trajectory"):
        e.advance_dc(1.0, 5)
    one = Q.Ensemble(Q.System(PTLA), 1, ground(2))
    assert one.advance_dc(0.0, 5).shape == (0, 2 + one.system.n_avr)
    one.advance([0.2])
    with pytest.raises(Q.QtError, match="before"):
        one.advance_dc(0.1, 5)
