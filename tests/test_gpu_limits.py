"""Limits and refusals of the device path: empty and out-of-range inputs, the largest compiled
sizes, and parameters the reference rejects, each refused with a QT_E* code and a message
(the reference's exceptions, Tridiagonal.cc:54-58, MCWF_Trajectory.h:114-117, ODE.h) instead of
running; and the smallest legal cases (one trajectory, a zero-length interval) running."""
import numpy as np
import pytest

import cppqed_amd as Q
import oracle_lib as O
from sysdefs import JC60, PTLA, ground

pytestmark = pytest.mark.gpu


def _modes(n, cutoff, pumped=True):
    lines = [f"free mode cutoff={cutoff} delta=-1 kappa=1 eta={1 if pumped else 0}" for _ in range(n)]
    lines += [f"ia bs {i} {i + 1} g=0.5" for i in range(n - 1)] + ["system composite"]
    return "\n".join(lines)


def test_refusals():
    print("[gpu] refusals", flush=True)
    s = Q.System(JC60)
    psi = ground(60)
    with pytest.raises(Q.QtError):
        Q.Ensemble(s, 0, psi)                                   # empty ensemble
    with pytest.raises(Q.QtError):
        Q.Ensemble(s, 4, psi, Q.mcwf_pars(epsRel=0.0))          # tolerances must be positive
    with pytest.raises(Q.QtError):
        Q.Ensemble(s, 4, psi, Q.mcwf_pars(dpLimit=0.0))
    with pytest.raises(Q.QtError):
        Q.Ensemble(s, 4, psi, Q.mcwf_pars(overshootTol=0.5))
    with pytest.raises(Q.QtError):
        Q.Ensemble(s, 4, psi, Q.mcwf_pars(layout=(1, 2)))     # a layout that is not compiled in
    e = Q.Ensemble(s, 4, psi, Q.mcwf_pars(seed=3))
    for bad in ([], [-0.1], [np.nan], [np.inf]):
        with pytest.raises(Q.QtError):
            e.advance(bad)                                     # no intervals / negative / non-finite
    with pytest.raises(Q.QtError):
        e.state(4)                                             # trajectory index out of range
    with pytest.raises(Q.QtError):
        e.state(-1)
    # 13 qubits: dimension 8192 exceeds the largest layout (4 rows x 16 wavefronts x 64 lanes)
    big = Q.System("\n".join(["free qbit deltaA=1 gamma=0.1 eta=0.1"] * 13 + ["system composite"]))
    assert big.dim == 8192
    with pytest.raises(Q.QtError):
        Q.Ensemble(big, 2, ground(big.dim))
    # nine pumped modes: every row of H_nH/i holds 1 + 2*9 + 2*8 entries, more than 16
    wide = Q.System(_modes(9, 2))
    with pytest.raises(Q.QtError):
        Q.Ensemble(wide, 2, ground(wide.dim))


def test_smallest_cases_run():
    """One trajectory; a zero-length interval streams the current averages; the largest dimension
    the layouts cover (4096) runs."""
    print("[gpu] smallest cases", flush=True)
    s, so = Q.System(PTLA), O.System(PTLA)
    e = Q.Ensemble(s, 1, ground(2), Q.mcwf_pars(seed=11, jump_log_cap=64))
    res = O.mcwf_run(so, ground(2), O.params(rng="philox", seed=11, stream=0), O.run_spec("NDt", Dt=0.1, NDt=5))
    rows0 = e.advance([0.0])
    assert rows0[0, 0] == 0.0
    np.testing.assert_allclose(rows0[0, 2:], res["rows"][0, 2:], rtol=0, atol=1e-15)
    rows = e.advance(np.full(5, 0.1))
    psi, _ = e.state(0)
    assert np.linalg.norm(psi - res["psi"]) <= 1e-9
    np.testing.assert_allclose(rows[:, 1:], res["rows"][1:, 1:], rtol=1e-9, atol=1e-12)
    big = Q.System(_modes(4, 8))
    assert big.dim == 4096
    eb = Q.Ensemble(big, 2, ground(big.dim), Q.mcwf_pars(seed=5, precision=Q.QT_FP32_DERIV))
    rb = eb.advance([0.05, 0.05])
    assert np.isfinite(rb).all() and rb[-1, 0] == pytest.approx(0.1)
    for k in range(2):
        p, _ = eb.state(k)
        assert abs(np.linalg.norm(p) - 1) < 1e-12


@pytest.mark.parametrize("name", ["pcav1024", "modes4096"])
def test_large_states_with_offset_gathers_match_oracle(name):
    """The largest compiled layouts (4 rows x 4 and 16 wavefronts, psi in LDS) with the
    constant-offset gathers, FP64, each trajectory against the oracle."""
    from sysdefs import MIXED, basis
    spec, idx = next((sp, i) for n, sp, i in MIXED if n == name)
    print(f"[gpu] large state {name} vs oracle", flush=True)
    s, so = Q.System(spec), O.System(spec)
    assert s.info.h_offsets > 0
    psi0 = basis(s.dim, idx)
    seed, ntraj, iv = 31, 6, np.full(20, 0.1)
    e = Q.Ensemble(s, ntraj, psi0, Q.mcwf_pars(seed=seed, jump_log_cap=256))
    e.advance(iv)
    per, _ = e.counters()
    for k in range(ntraj):
        res = O.mcwf_run(so, psi0, O.params(rng="philox", seed=seed, stream=k), O.run_spec("NDt", Dt=0.1, NDt=len(iv)))
        psi, _ = e.state(k)
        jt, jch, nj = e.jumplog(k)
        assert nj == len(res["jump_t"]), (k, nj, len(res["jump_t"]))
        np.testing.assert_array_equal(jch, res["jump_ch"])
        np.testing.assert_allclose(jt, res["jump_t"], rtol=1e-10, atol=1e-12)
        assert np.linalg.norm(psi - res["psi"]) <= 1e-9 * np.linalg.norm(res["psi"]), k
        for c in (0, 1, 2, 4, 5, 7):
            assert per[k][c] == res["counters"][c], (k, c, per[k][c], res["counters"][c])


@pytest.mark.parametrize("cutoff", [1024, 1000])
def test_offset_gathers_in_a_team_match_oracle(cutoff):
    """A pumped lossy mode in a 4-wavefront team (4 rows per lane, 1024 lanes' worth of rows):
    at cutoff 1024 the state fills the team and the slot-major gather runs; at 1000 it does not
    and the general constant-offset gather does.  FP64, each trajectory against the oracle."""
    spec = f"free mode cutoff={cutoff} delta=-1 kappa=1 eta=2"
    print(f"[gpu] offset gathers in a team, mode cutoff {cutoff}", flush=True)
    s, so = Q.System(spec), O.System(spec)
    assert s.dim == cutoff and s.info.h_offsets > 0
    psi0 = ground(s.dim)
    seed, ntraj, iv = 41, 4, np.full(10, 0.1)
    e = Q.Ensemble(s, ntraj, psi0, Q.mcwf_pars(seed=seed, jump_log_cap=256, layout=(4, 4)))
    e.advance(iv)
    per, _ = e.counters()
    for k in range(ntraj):
        res = O.mcwf_run(so, psi0, O.params(rng="philox", seed=seed, stream=k), O.run_spec("NDt", Dt=0.1, NDt=len(iv)))
        psi, _ = e.state(k)
        jt, jch, nj = e.jumplog(k)
        assert nj == len(res["jump_t"]), (k, nj, len(res["jump_t"]))
        np.testing.assert_array_equal(jch, res["jump_ch"])
        assert np.linalg.norm(psi - res["psi"]) <= 1e-9 * np.linalg.norm(res["psi"]), k
        for c in (0, 1, 2, 4, 5, 7):
            assert per[k][c] == res["counters"][c], (k, c, per[k][c], res["counters"][c])


def test_non_finite_state_is_reported():
    """The reference's numeric-failure exceptions.  A non-finite initial state is refused when the
    ensemble is made (its rates: LiouvilleanAveragedCommon.h:82-93 via the MCWF_Trajectory
    constructor, MCWF_Trajectory.h:114-117).  A non-finite state loaded into one trajectory (a
    host-forced checkpoint) makes the next advance fail with QT_ENUMERIC "Infinite detected in
    averages" (the rates of that trajectory's first step) -- the kernel itself finishes: with
    non-finite rates no jump fires and the controller only grows dt.  A reset clears the failure,
    and the ensemble then runs as new."""
    print("[gpu] non-finite state", flush=True)
    s = Q.System(JC60)
    bad = ground(60)
    bad[5] = np.nan   # a photon-number state the mode's jump operator reads
    with pytest.raises(Q.QtError) as ex:
        Q.Ensemble(s, 4, bad)
    assert ex.value.code == Q.QT_ENUMERIC and "Infinite detected in averages" in str(ex.value)
    e = Q.Ensemble(s, 8, ground(60), Q.mcwf_pars(seed=3))
    ref = e.advance(np.full(5, 0.1))
    e.reset()
    st, psi, jt, jc = e.get_states()
    psi[3, 7] = np.inf
    e.set_states(st, psi, jt, jc)
    with pytest.raises(Q.QtError) as ex:
        e.advance(np.full(5, 0.1))
    assert ex.value.code == Q.QT_ENUMERIC and "Infinite detected in averages" in str(ex.value)
    e.reset()
    np.testing.assert_array_equal(e.advance(np.full(5, 0.1)), ref)
