"""GPU parity of the mixed batch: ensembles of four different systems (dims 1024, 4096, 64, 60)
advanced by one batched launch (work-group teams in one persistent kernel, quarter and wavefront
teams in a second beside it), against the CPU oracle (FP64) and against each member run alone
(FP32 derivatives)."""
import numpy as np
import pytest

import cppqed_amd as Q
import oracle_lib as O
from sysdefs import MIXED, basis

pytestmark = pytest.mark.gpu

SEED = 20240611


def _members(counts, prec, jcap=256):
    out = []
    for i, ((name, spec, idx), n) in enumerate(zip(MIXED, counts)):
        s = Q.System(spec)
        out.append(Q.Ensemble(s, int(n), basis(s.dim, idx), Q.mcwf_pars(seed=SEED + i, precision=prec, jump_log_cap=jcap)))
    return out


def test_mixed_fp64_trajectories_match_oracle():
    counts = [5, 3, 12, 12]
    print(f"[gpu] mixed batch fp64 parity counts={counts}", flush=True)
    iv = np.full(10, 0.1)
    mem = _members(counts, Q.QT_FP64)
    b = Q.Batch(mem)
    v, team = b.layout()
    assert team == [2, 0, 1, 1], team
    rows = b.advance(iv)
    for i, ((name, spec, idx), e) in enumerate(zip(MIXED, mem)):
        so = O.System(spec)
        psi0 = basis(so.dim, idx)
        per, _ = e.counters()
        for k in range(e.n):
            res = O.mcwf_run(so, psi0, O.params(rng="philox", seed=SEED + i, stream=k), O.run_spec("NDt", Dt=0.1, NDt=10))
            psi_g, _ = e.state(k)
            jt, jch, nj = e.jumplog(k)
            assert nj == len(res["jump_t"]), (name, k)
            np.testing.assert_array_equal(jch, res["jump_ch"])
            np.testing.assert_allclose(jt, res["jump_t"], rtol=1e-10, atol=1e-12)
            err = np.linalg.norm(psi_g - res["psi"]) / np.linalg.norm(res["psi"])
            assert err <= 1e-9, (name, k, err)
            for c in (0, 1, 2, 4, 5, 7):
                assert per[k][c] == res["counters"][c], (name, k, c)
        ro, _ = O.ensemble_run(so, psi0, O.params(rng="philox", seed=SEED + i, stream=0), O.run_spec("NDt", Dt=0.1, NDt=10), e.n)
        np.testing.assert_allclose(rows[i], ro[1:], rtol=1e-9, atol=1e-9, err_msg=name)


def _check_against_oracle(entries, mem, rows, nint, seed0=SEED):
    """Every trajectory of every member against the oracle (FP64), as in the test above."""
    for i, ((name, spec, idx), e) in enumerate(zip(entries, mem)):
        so = O.System(spec)
        psi0 = basis(so.dim, idx)
        per, _ = e.counters()
        for k in range(e.n):
            res = O.mcwf_run(so, psi0, O.params(rng="philox", seed=seed0 + i, stream=k), O.run_spec("NDt", Dt=0.1, NDt=nint))
            psi_g, _ = e.state(k)
            jt, jch, nj = e.jumplog(k)
            assert nj == len(res["jump_t"]), (name, k)
            np.testing.assert_array_equal(jch, res["jump_ch"])
            err = np.linalg.norm(psi_g - res["psi"]) / np.linalg.norm(res["psi"])
            assert err <= 1e-9, (name, k, err)
            for c in (0, 1, 2, 4, 5, 7):
                assert per[k][c] == res["counters"][c], (name, k, c)
        ro, _ = O.ensemble_run(so, psi0, O.params(rng="philox", seed=seed0 + i, stream=0), O.run_spec("NDt", Dt=0.1, NDt=nint), e.n)
        np.testing.assert_allclose(rows[i], ro[1:], rtol=1e-9, atol=1e-9, err_msg=name)


def test_mixed_single_kernel_batches_match_oracle():
    """The batched launch is two kernels (work-group teams; quarter and wavefront teams on the
    batch's side stream).  A batch with every member on work groups runs only the first, one
    without a work-group member only the second; both match the oracle trajectory by trajectory."""
    iv = np.full(6, 0.1)
    counts = [3, 2, 5, 5]
    print(f"[gpu] mixed batch, every member on work groups, counts={counts}", flush=True)
    mem = _members(counts, Q.QT_FP64)
    b = Q.Batch(mem, team=[0, 0, 0, 0])
    assert b.layout()[1] == [0, 0, 0, 0]
    _check_against_oracle(MIXED, mem, b.advance(iv), len(iv))
    sub = [MIXED[0], MIXED[2], MIXED[3]]
    print(f"[gpu] mixed batch without a work-group member: {[e[0] for e in sub]}", flush=True)
    mem2 = []
    for i, ((name, spec, idx), n) in enumerate(zip(sub, [4, 6, 6])):
        s = Q.System(spec)
        mem2.append(Q.Ensemble(s, n, basis(s.dim, idx), Q.mcwf_pars(seed=SEED + i, precision=Q.QT_FP64, jump_log_cap=256)))
    b2 = Q.Batch(mem2)
    assert 0 not in b2.layout()[1], b2.layout()
    _check_against_oracle(sub, mem2, b2.advance(iv), len(iv))


def test_mixed_equals_members_alone():
    """Each member's rows from the mixed launch equal the member run alone with the same engine."""
    counts = [64, 32, 256, 256]
    print(f"[gpu] mixed batch vs alone counts={counts}", flush=True)
    iv = np.full(10, 0.1)
    mem = _members(counts, Q.QT_FP32_DERIV)
    rows = Q.Batch(mem).advance(iv)
    layouts = [(4, 4), (4, 16), (1, 1), (1, 1)]   # the engines of team kinds 2, 0, 1, 1
    for i, ((name, spec, idx), n) in enumerate(zip(MIXED, counts)):
        s = Q.System(spec)
        alone = Q.Ensemble(s, n, basis(s.dim, idx), Q.mcwf_pars(seed=SEED + i, precision=Q.QT_FP32_DERIV, layout=layouts[i]))
        r = alone.advance(iv)
        diff = np.abs(r - rows[i]).max()
        print(f"  {name}: max |alone - mixed| = {diff:.3g}", flush=True)
        np.testing.assert_allclose(rows[i], r, rtol=1e-13, atol=1e-14, err_msg=name)
        _, t1 = alone.counters()
        _, t2 = mem[i].counters()
        np.testing.assert_array_equal(t1, t2)


def test_mixed_autotune_keeps_results():
    counts = [16, 8, 64, 64]
    print(f"[gpu] mixed batch autotune counts={counts}", flush=True)
    iv = np.full(5, 0.1)
    ref = Q.Batch(_members(counts, Q.QT_FP64)).advance(iv)
    mem = _members(counts, Q.QT_FP64)
    b = Q.Batch(mem)
    v, ms = b.autotune(probe_span=0.2, probe_traj=256)
    print(f"  autotune picked variant {v} of {Q.batch_variants()}, probe ms {ms}", flush=True)
    assert 0 <= v < len(Q.batch_variants()) and len(ms) >= 1
    rows = b.advance(iv)
    for a, r in zip(rows, ref):
        np.testing.assert_allclose(a, r, rtol=1e-9, atol=1e-12)


def test_mixed_gsl_and_interaction_picture_members_match_oracle():
    """Members that differ in controller and picture share one launch: GSL's controller
    (ODE_GSL.h:44-84) on work-group, quarter and wavefront teams, interaction-picture systems
    (Exact::actWithU after each step, MCWF_Trajectory.h:269-272) beside Schroedinger ones."""
    from sysdefs import JC60, MODES4096, P1M1, PCAV1024, RING, ground, p1m1_initial, ring_initial
    cases = [("pcav1024", PCAV1024, basis(1024, 32), 3, "gsl"),
             ("modes4096", MODES4096, basis(4096, 0), 2, "gsl"),
             ("jc60", JC60, ground(60), 8, "odeint"),
             ("p1m1_ip", P1M1, p1m1_initial(), 8, "odeint"),
             ("ring_ip", RING, ring_initial(), 6, "gsl")]
    iv = np.full(5, 0.1)
    mem = [Q.Ensemble(Q.System(spec), n, psi, Q.mcwf_pars(seed=SEED + i, jump_log_cap=256, controller=ctl))
           for i, (_, spec, psi, n, ctl) in enumerate(cases)]
    b = Q.Batch(mem)
    v, team = b.layout()
    print(f"[gpu] mixed batch gsl + interaction picture: variant {v} teams {team}", flush=True)
    assert {0, 2} <= set(team) and 1 in team
    rows = b.advance(iv)
    for i, ((name, spec, psi0, n, ctl), e) in enumerate(zip(cases, mem)):
        so = O.System(spec)
        per, _ = e.counters()
        for k in range(n):
            p = O.params(rng="philox", seed=SEED + i, stream=k, controller=ctl)
            res = O.mcwf_run(so, psi0, p, O.run_spec("NDt", Dt=0.1, NDt=len(iv)))
            psi_g, _ = e.state(k)
            jt, jch, nj = e.jumplog(k)
            assert nj == len(res["jump_t"]), (name, k)
            np.testing.assert_array_equal(jch, res["jump_ch"])
            np.testing.assert_allclose(jt, res["jump_t"], rtol=1e-10, atol=1e-12)
            err = np.linalg.norm(psi_g - res["psi"]) / np.linalg.norm(res["psi"])
            assert err <= 1e-9, (name, k, err)
            for c in (0, 1, 2, 4, 5, 7) + ((3,) if ctl == "gsl" else ()):
                assert per[k][c] == res["counters"][c], (name, k, c)
        ro, _ = O.ensemble_run(so, psi0, O.params(rng="philox", seed=SEED + i, stream=0, controller=ctl),
                               O.run_spec("NDt", Dt=0.1, NDt=len(iv)), n)
        np.testing.assert_allclose(rows[i], ro[1:], rtol=1e-9, atol=1e-9, err_msg=name)
