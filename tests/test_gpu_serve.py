"""GPU tests of the serving loop: a request's result is the same alone and in a mix of
requests arriving while others run; admission respects the memory budget; results match the
oracle."""
import numpy as np
import pytest

import cppqed_amd as Q
import oracle_lib as O
from sysdefs import JC60, MIXED, PLM20, basis, coherent, ground

pytestmark = pytest.mark.gpu


def _drain(srv, rids, max_steps=10000):
    for _ in range(max_steps):
        a, w = srv.step()
        if a == 0 and w == 0:
            break
    return {r: srv.result(r) for r in rids}


@pytest.mark.parametrize("prec", [Q.QT_FP64, Q.QT_FP32_DERIV])
def test_request_same_alone_and_in_mix(prec):
    print(f"[gpu] serve alone vs mix precision={prec}", flush=True)
    pars = Q.mcwf_pars(precision=prec)
    jc = Q.System(JC60)
    others = [(Q.System(spec), idx) for _, spec, idx in MIXED]
    plm = Q.System(PLM20)
    # alone
    srv = Q.Server(chunk_intervals=5, pars=pars)
    rid = srv.submit(jc, 256, ground(60), T=2.0, seed=11)
    alone = _drain(srv, [rid])[rid][0]
    # in a mix: requests arrive before, with and after it, between steps
    srv = Q.Server(chunk_intervals=5, pars=pars)
    ids = [srv.submit(plm, 512, coherent(0.5, 20), T=1.5, seed=3)]
    ids.append(srv.submit(others[1][0], 48, basis(4096, others[1][1]), T=1.0, seed=4))
    srv.step()
    rid = srv.submit(jc, 256, ground(60), T=2.0, seed=11)
    srv.step()
    ids.append(srv.submit(others[0][0], 128, basis(1024, others[0][1]), T=1.2, seed=5))
    ids.append(srv.submit(others[2][0], 300, basis(64, others[2][1]), T=0.7, seed=6))
    res = _drain(srv, ids + [rid])
    mixed = res[rid][0]
    assert mixed.shape == alone.shape == (20, 2 + jc.n_avr)
    np.testing.assert_array_equal(mixed, alone)
    # and the same as the ensemble driven directly
    ens = Q.Ensemble(jc, 256, ground(60), Q.mcwf_pars(precision=prec, seed=11))
    direct = np.concatenate([ens.advance(np.full(5, 0.1)) for _ in range(4)])
    np.testing.assert_allclose(mixed, direct, rtol=1e-13, atol=1e-15)
    for r in ids:
        assert res[r] is not None and res[r][1] > 0


def test_admission_by_memory():
    print("[gpu] serve admission by memory", flush=True)
    jc = Q.System(JC60)
    probe = Q.Server(chunk_intervals=4)
    probe.submit(jc, 2048, ground(60), T=0.8, seed=0)
    probe.step()
    one = probe.stats()["bytes_in_use"]
    assert one > 2048 * 60 * 16
    budget = int(2.5 * one)            # two such requests fit at a time
    srv = Q.Server(hbm_budget=budget, chunk_intervals=4)
    ids = [srv.submit(jc, 2048, ground(60), T=0.8, seed=s) for s in range(5)]
    peak = 0
    for _ in range(1000):
        a, w = srv.step()
        assert srv.stats()["bytes_in_use"] <= budget
        peak = max(peak, a)
        if a == 0 and w == 0:
            break
    assert peak == 2, peak
    waits = sorted(srv.result(r)[2] for r in ids)
    assert waits[-1] > waits[0]
    with pytest.raises(Q.QtError):
        srv.submit(jc, 1 << 16, ground(60), T=1.0)   # larger than the whole budget


def test_served_results_match_oracle():
    print("[gpu] serve fp64 vs oracle", flush=True)
    srv = Q.Server(chunk_intervals=3, pars=Q.mcwf_pars(precision=Q.QT_FP64))
    cases = [(PLM20, coherent(0.5, 20), 40, 1.0, 21), (JC60, ground(60), 24, 0.9, 22)]
    ids = [srv.submit(Q.System(spec), n, psi, T=T, seed=seed) for spec, psi, n, T, seed in cases]
    res = _drain(srv, ids)
    for (spec, psi, n, T, seed), r in zip(cases, ids):
        ro, _ = O.ensemble_run(O.System(spec), psi, O.params(rng="philox", seed=seed, stream=0),
                               O.run_spec("Dt", Dt=0.1, T=T), n)
        np.testing.assert_allclose(res[r][0], ro[1:], rtol=1e-9, atol=1e-9)


def test_sweep_requests_same_alone_and_in_mix():
    """Parameter-sweep solves (Master-equation points) as requests: the rows are the same alone,
    among ensemble requests arriving before and after, and as the sweep driven directly."""
    from sysdefs import jc_master_spec
    print("[gpu] serve sweep alone vs mix", flush=True)
    pts = [Q.System(jc_master_spec(40, eta=float(e), g=10.0)) for e in (0.5, 4.0, 9.0)]
    rho0 = np.zeros((80, 80), complex)
    rho0[0, 0] = 1
    T = 1.0
    srv = Q.Server(chunk_intervals=4)
    rid = srv.submit_sweep(pts, rho0, T=T, epsRel=1e-8, epsAbs=1e-8)
    alone = _drain(srv, [rid])[rid][0]
    iv = Q.run_intervals(T=T, Dt=0.1)
    assert alone.shape == (len(pts) * len(iv), 2 + pts[0].n_avr)
    jc = Q.System(JC60)
    srv = Q.Server(chunk_intervals=4, pars=Q.mcwf_pars(precision=Q.QT_FP32_DERIV))
    ids = [srv.submit(jc, 512, ground(60), T=1.5, seed=3)]
    srv.step()
    rid = srv.submit_sweep(pts, rho0, T=T, epsRel=1e-8, epsAbs=1e-8)
    srv.step()
    ids.append(srv.submit(jc, 256, ground(60), T=0.8, seed=4))
    res = _drain(srv, ids + [rid])
    np.testing.assert_array_equal(res[rid][0], alone)
    assert res[rid][3] > 0 and all(res[r] is not None for r in ids)
    m = Q.MasterSweep(pts, rho0, epsRel=1e-8, epsAbs=1e-8)
    rows, _ = m.advance(iv)
    np.testing.assert_array_equal(alone.reshape(len(pts), len(iv), -1), rows)


def test_several_sweeps_in_flight():
    """Three sweep requests in flight at once (each on its own stream with its own staging),
    different point counts and lengths so that last chunks are shorter (chunks of 4 intervals:
    5 = 4 + 1, 7 = 4 + 3, 13 = 4 + 4 + 4 + 1), beside ensemble requests: each sweep's rows equal
    the same sweep driven directly."""
    from sysdefs import jc_master_spec
    print("[gpu] serve: three sweeps in flight", flush=True)
    rho0 = np.zeros((80, 80), complex)
    rho0[0, 0] = 1
    cases = [([0.5], 0.5), ([2.0, 7.5], 0.7), ([1.0, 4.0, 12.0], 1.3)]
    jc = Q.System(JC60)
    srv = Q.Server(chunk_intervals=4, pars=Q.mcwf_pars(precision=Q.QT_FP32_DERIV))
    ens_ids = [srv.submit(jc, 300, ground(60), T=1.1, seed=5)]
    sweep_ids = []
    for etas, T in cases:
        pts = [Q.System(jc_master_spec(40, eta=e, g=10.0)) for e in etas]
        sweep_ids.append((srv.submit_sweep(pts, rho0, T=T, epsRel=1e-8, epsAbs=1e-8), pts, T))
    ens_ids.append(srv.submit(jc, 200, ground(60), T=0.6, seed=6))
    res = _drain(srv, ens_ids + [r for r, _, _ in sweep_ids])
    assert all(res[r] is not None for r in ens_ids)
    for rid, pts, T in sweep_ids:
        iv = Q.run_intervals(T=T, Dt=0.1)
        m = Q.MasterSweep(pts, rho0, epsRel=1e-8, epsAbs=1e-8)
        rows, _ = m.advance(iv)
        np.testing.assert_array_equal(res[rid][0].reshape(len(pts), len(iv), -1), rows)


def test_sweep_request_refusals():
    """Sweep points must share dimension and jumps, have a unitary interaction picture and fit the
    requested Master kernel (the MFMA one: d <= 96); refused at submission, so the loop itself
    never fails on them."""
    from sysdefs import jc_master_spec
    srv = Q.Server(chunk_intervals=4)
    a = Q.System(jc_master_spec(40, eta=1.0, g=10.0))
    b = Q.System(jc_master_spec(30, eta=1.0, g=10.0))
    rho = np.zeros((80, 80), complex)
    rho[0, 0] = 1
    with pytest.raises(Q.QtError):
        srv.submit_sweep([a, b], rho, T=1.0)                 # dimensions differ
    big = Q.System(jc_master_spec(60, eta=1.0, g=10.0))
    rho_big = np.zeros((120, 120), complex)
    rho_big[0, 0] = 1
    with pytest.raises(Q.QtError, match="up to 96"):
        srv.submit_sweep([big], rho_big, T=1.0, kernel="mfma")   # beyond the MFMA kernel's 96
    ip = Q.System("free mode cutoff=20 delta=-1 kappa=1 eta=2 picture=IP")
    rho20 = np.zeros((20, 20), complex)
    rho20[0, 0] = 1
    with pytest.raises(Q.QtError):
        srv.submit_sweep([ip], rho20, T=1.0)                 # interaction picture
    # the parameters qt_master_create refuses are refused at submission too
    with pytest.raises(Q.QtError):
        srv.submit_sweep([a], rho, T=1.0, epsRel=0.0)
    with pytest.raises(Q.QtError):
        srv.submit_sweep([a], rho, T=1.0, autostop=(1e-3, 1e-6, -1))
    with pytest.raises(Q.QtError):
        srv.submit_sweep([a], rho, T=1.0, autostop=(0.0, 1e-6, 5))
    with pytest.raises(ValueError):
        srv.submit_sweep([a], rho20, T=1.0)                  # rho0 of the wrong shape
    with pytest.raises(ValueError):
        srv.submit(a, 16, ground(20), T=1.0)                 # psi0 of the wrong size
    assert srv.step() == (0, 0)


def test_failed_requests_are_retired_alone():
    """A request that fails is retired with the reference's error and the others are unaffected:
    one refused at admission (a non-finite initial rate: LiouvilleanAveragedCommon.h:82-93 throws
    when the MCWF_Trajectory constructor computes the rates, MCWF_Trajectory.h:114-117), one that
    fails inside the shared mixed launch (a system without jumps, so nothing is checked before the
    run, whose streamed averages come out non-finite: "Infinite detected in averages").  The
    neighbour's rows equal its rows alone, bit for bit, and the loop keeps serving."""
    print("[gpu] serve failure containment", flush=True)
    jc = Q.System(JC60)
    nojump = Q.System("free mode cutoff=8 delta=-1 kappa=0 eta=1")
    bad_psi = ground(60).copy()
    bad_psi[1] = np.nan
    srv = Q.Server(chunk_intervals=5)
    rid = srv.submit(jc, 128, ground(60), T=1.0, seed=9)
    alone = _drain(srv, [rid])[rid][0]
    srv = Q.Server(chunk_intervals=5)
    ids = [srv.submit(jc, 64, bad_psi, T=1.0, seed=8)]
    nan8 = ground(8).copy()
    nan8[2] = np.inf
    ids.append(srv.submit(nojump, 32, nan8, T=1.0, seed=7))
    rid = srv.submit(jc, 128, ground(60), T=1.0, seed=9)
    for _ in range(100):
        a, w = srv.step()
        if a == 0 and w == 0:
            break
    for r in ids:
        with pytest.raises(Q.QtError) as ex:
            srv.result(r)
        assert ex.value.code == Q.QT_ENUMERIC
        assert "Infinite detected in averages" in str(ex.value)
    np.testing.assert_array_equal(srv.result(rid)[0], alone)
    st = srv.stats()
    assert st["failed"] == 2 and st["finished"] == 1 and st["bytes_in_use"] == 0
    # the loop goes on serving
    rid2 = srv.submit(jc, 128, ground(60), T=1.0, seed=9)
    np.testing.assert_array_equal(_drain(srv, [rid2])[rid2][0], alone)
    for r in ids + [rid, rid2]:
        srv.release(r)


@pytest.mark.parametrize("ctl", ["odeint", "gsl"])
def test_served_interaction_picture_and_gsl_match_oracle(ctl):
    """Requests of interaction-picture systems (Exact::actWithU after each step, MCWF_Trajectory.h:
    269-272) and a server running GSL's controller (ODE_GSL.h:44-84) match the oracle's ensembles."""
    from sysdefs import P1M1, RING, p1m1_initial, ring_initial
    print(f"[gpu] serve interaction picture, controller {ctl}", flush=True)
    srv = Q.Server(chunk_intervals=3, pars=Q.mcwf_pars(precision=Q.QT_FP64, controller=ctl))
    cases = [(P1M1, p1m1_initial(), 16, 0.5, 31), (PLM20.replace("kappa", "picture=IP kappa"), coherent(0.5, 20), 24, 0.6, 32),
             (RING, ring_initial(), 8, 0.4, 33), (JC60, ground(60), 12, 0.5, 34)]
    ids = [srv.submit(Q.System(spec), n, psi, T=T, seed=seed) for spec, psi, n, T, seed in cases]
    res = _drain(srv, ids)
    for (spec, psi, n, T, seed), r in zip(cases, ids):
        ro, _ = O.ensemble_run(O.System(spec), psi, O.params(rng="philox", seed=seed, stream=0, controller=ctl),
                               O.run_spec("Dt", Dt=0.1, T=T), n)
        np.testing.assert_allclose(res[r][0], ro[1:], rtol=1e-9, atol=1e-9)
