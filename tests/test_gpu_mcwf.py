"""GPU parity of the MCWF hot path against the CPU oracle (same counter-based draws)."""
import numpy as np
import pytest

import cppqed_amd as Q
import oracle_lib as O
from sysdefs import JC60, MODES4096, P1M1, PCAV1024, PLM20, PTLA, basis, coherent, ground, p1m1_initial

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


@pytest.mark.parametrize("spec", [JC60, PLM20, PTLA])
@pytest.mark.parametrize("prec,tol", [(Q.QT_FP64, 1e-14), (Q.QT_FP32_DERIV, 2e-6)])
def test_apply_heff_matches_oracle(spec, prec, tol):
    sq, so = Q.System(spec), O.System(spec)
    rng = np.random.default_rng(7)
    n = 33
    psi = rng.normal(size=(n, sq.dim)) + 1j * rng.normal(size=(n, sq.dim))
    base = rng.normal(size=(n, sq.dim)) + 1j * rng.normal(size=(n, sq.dim))
    dpsi = _dev(base.copy())
    x = _dev(psi)
    sq.apply_heff(x.data_ptr(), dpsi.data_ptr(), n=n, precision=prec)
    torch.cuda.synchronize()
    got = dpsi.cpu().numpy()
    for i in range(n):
        ref = so.apply_heff(psi[i], dpsidt=base[i])
        scale = np.abs(ref).max()
        assert np.abs(got[i] - ref).max() <= tol * scale, (i, np.abs(got[i] - ref).max(), scale)


def _oracle_traj(so, psi0, k, seed, intervals, **kw):
    p = O.params(rng="philox", seed=seed, stream=k, **kw)
    r = O.run_spec("NDt", Dt=intervals[0], NDt=len(intervals))
    return O.mcwf_run(so, psi0, p, r)


def _assert_traj_parity(res_o, psi_g, jt_g, jch_g, nj_g, cnt_g=None):
    assert nj_g == len(res_o["jump_t"]), (nj_g, len(res_o["jump_t"]))
    np.testing.assert_array_equal(jch_g, res_o["jump_ch"])
    np.testing.assert_allclose(jt_g, res_o["jump_t"], rtol=1e-10, atol=1e-12)
    err = np.linalg.norm(psi_g - res_o["psi"]) / np.linalg.norm(res_o["psi"])
    assert err <= 1e-9, err
    if cnt_g is not None:
        # MCWF steps, ODE steps, failed tries, tries, step-backs, jumps
        for c in (0, 1, 2, 4, 5, 7):
            assert cnt_g[c] == res_o["counters"][c], (c, cnt_g[c], res_o["counters"][c])


@pytest.mark.parametrize("glanes", [0, 64])
@pytest.mark.parametrize("name,spec,psi0,T,ntraj", [
    ("jc60", JC60, ground(60), 3.0, 48),
    ("plm20", PLM20, coherent(0.5, 20), 2.0, 32),
    ("ptla", PTLA, ground(2), 1.0, 32),
])
def test_batched_fp64_trajectories_match_oracle(name, spec, psi0, T, ntraj, glanes):
    print(f"[gpu] batched fp64 parity {name} ntraj={ntraj} T={T} group_lanes={glanes}", flush=True)
    sq, so = Q.System(spec), O.System(spec)
    seed = 20240601
    Dt = 0.1 if name != "ptla" else 0.05
    iv = np.full(int(round(T / Dt)), Dt)
    ens = Q.Ensemble(sq, ntraj, psi0, Q.mcwf_pars(seed=seed, traj0=0, jump_log_cap=256, group_lanes=glanes))
    rows = ens.advance(iv)
    per, tot = ens.counters()
    njumps = 0
    for k in range(ntraj):
        res_o = _oracle_traj(so, psi0, k, seed, iv)
        psi_g, t_g = ens.state(k)
        jt, jch, nj = ens.jumplog(k)
        _assert_traj_parity(res_o, psi_g, jt, jch, nj, per[k])
        njumps += nj
    assert tot[7] == njumps
    # ensemble rows against the oracle's ensemble
    rows_o, tries_o = O.ensemble_run(so, psi0, O.params(rng="philox", seed=seed, stream=0),
                                     O.run_spec("NDt", Dt=Dt, NDt=len(iv)), ntraj)
    np.testing.assert_allclose(rows[:, 0], rows_o[1:, 0], rtol=0, atol=1e-12)
    np.testing.assert_allclose(rows[:, 2:], rows_o[1:, 2:], rtol=1e-9, atol=1e-9)
    assert tot[4] == tries_o


@pytest.mark.parametrize("name,spec,psi0,n_iv,ctl", [
    ("jc60", JC60, ground(60), 10, "odeint"),
    ("jc60_gsl", JC60, ground(60), 10, "gsl"),
    # interaction picture: the post-step kernel applies Exact::actWithU before the renormalisation
    ("p1m1_ip", P1M1, p1m1_initial(), 3, "odeint"),
    ("p1m1_ip_gsl", P1M1, p1m1_initial(), 3, "gsl"),
    ("plm20_ip", PLM20.replace("kappa", "picture=IP kappa"), coherent(0.5, 20), 10, "odeint"),
])
def test_plain_form_matches_oracle(name, spec, psi0, n_iv, ctl):
    print(f"[gpu] plain form parity {name}", flush=True)
    sq, so = Q.System(spec), O.System(spec)
    seed = 77
    iv = np.full(n_iv, 0.1)
    for k in range(3):
        res_g = Q.mcwf_plain_run(sq, psi0, iv, Q.mcwf_pars(seed=seed, controller=ctl), traj=k)
        res_o = _oracle_traj(so, psi0, k, seed, iv, controller=ctl)
        _assert_traj_parity(res_o, res_g["psi"], res_g["jump_t"], res_g["jump_ch"], int(res_g["counters"][7]),
                            res_g["counters"])
        if ctl == "gsl":
            assert res_g["counters"][3] == res_o["counters"][3]


def test_fp32_derivative_observable_error():
    """FP32 derivative evaluations under FP64 control: ensemble <a^dagger a> error vs the FP64 run."""
    print("[gpu] fp32-derivative observable error jc60", flush=True)
    sq = Q.System(JC60)
    iv = np.full(20, 0.1)
    n = 1024
    r64 = Q.Ensemble(sq, n, ground(60), Q.mcwf_pars(seed=5)).advance(iv)
    r32 = Q.Ensemble(sq, n, ground(60), Q.mcwf_pars(seed=5, precision=Q.QT_FP32_DERIV)).advance(iv)
    # column 6 = <n> of the mode (after time, dtDid, 4 qubit columns)
    err = np.abs(r32[:, 6] - r64[:, 6]).max()
    assert err < 1e-3, err


def test_layouts_agree():
    """Every compiled layout that covers the dimension gives the same trajectories."""
    print("[gpu] layout agreement jc60", flush=True)
    sq = Q.System(JC60)
    iv = np.full(5, 0.1)
    ref = None
    variants = [dict(layout=(1, 1)), dict(layout=(2, 1)), dict(layout=(4, 1)), dict(layout=(1, 4)),
                dict(group_lanes=16), dict(group_lanes=32), dict(group_lanes=64)]
    for var in variants:
        print(f"[gpu] layout {var}", flush=True)
        ens = Q.Ensemble(sq, 16, ground(60), Q.mcwf_pars(seed=9, **var))
        rows = ens.advance(iv)
        states = np.array([ens.state(k)[0] for k in range(16)])
        if ref is None:
            ref = (rows, states)
        else:
            np.testing.assert_allclose(rows[:, 2:], ref[0][:, 2:], rtol=1e-11, atol=1e-12)
            assert np.abs(states - ref[1]).max() < 1e-10


def test_graph_replay_matches_direct_launches():
    """Replays of the captured advance (jump draws, queue reset, batched kernel) give the same rows
    and counters as direct launches, with eager work and synchronisation between replays; a new
    interval count re-captures; after a reset the replay reproduces the first advance."""
    print("[gpu] graph replay vs direct", flush=True)
    sq = Q.System(JC60)
    pars = Q.mcwf_pars(seed=9, precision=Q.QT_FP32_DERIV)
    a = Q.Ensemble(sq, 512, ground(60), pars)
    b = Q.Ensemble(sq, 512, ground(60), pars)
    a.set_graph(True)
    iv = np.full(5, 0.1)
    first = None
    for rep in range(4):
        ra, rb = a.advance(iv), b.advance(iv)
        first = ra if first is None else first
        np.testing.assert_array_equal(ra, rb)
        pa, _ = a.counters()
        pb, _ = b.counters()
        np.testing.assert_array_equal(pa, pb)
    np.testing.assert_array_equal(a.advance(np.full(3, 0.1)), b.advance(np.full(3, 0.1)))
    assert a.set_graph(None) == (2, 5)
    a.reset()
    np.testing.assert_array_equal(a.advance(iv)[:, 1:], first[:, 1:])
    assert a.set_graph(None) == (3, 6)


@pytest.mark.parametrize("name,spec,psi0,T,ntraj", [
    ("jc60", JC60, ground(60), 2.0, 32),
    ("plm20", PLM20, coherent(0.5, 20), 2.0, 24),
    ("ptla", PTLA, ground(2), 1.0, 24),
    # interaction picture: the Exact propagation after each accepted step
    ("p1m1_ip", P1M1, p1m1_initial(), 0.5, 16),
    ("plm20_ip", PLM20.replace("kappa", "picture=IP kappa"), coherent(0.5, 20), 2.0, 16),
])
def test_gsl_controller_trajectories_match_oracle(name, spec, psi0, T, ntraj):
    """ODE_EngineGSL (ControlledErrorStepperGSL, ODE_GSL.h:44-84): every step accepted, the next dt
    from gsl_odeiv2_control_hadjust, seven derivative evaluations per step (dydt_out)."""
    print(f"[gpu] gsl controller parity {name} ntraj={ntraj}", flush=True)
    sq, so = Q.System(spec), O.System(spec)
    seed = 4711
    Dt = 0.1 if name != "ptla" else 0.05
    iv = np.full(int(round(T / Dt)), Dt)
    ens = Q.Ensemble(sq, ntraj, psi0, Q.mcwf_pars(seed=seed, jump_log_cap=256, controller="gsl"))
    ens.advance(iv)
    per, tot = ens.counters()
    assert tot[2] == 0                       # no rejected tries
    assert tot[3] == 7 * tot[4]              # k1, five stages, dydt_out
    for k in range(ntraj):
        res_o = _oracle_traj(so, psi0, k, seed, iv, controller="gsl")
        psi_g, _ = ens.state(k)
        jt, jch, nj = ens.jumplog(k)
        _assert_traj_parity(res_o, psi_g, jt, jch, nj, per[k])
        assert per[k][3] == res_o["counters"][3]


@pytest.mark.parametrize("name,spec,psi0,T,ntraj,layout", [
    # multi-wavefront teams with the state in LDS (LdsEngine): a small system forced onto four
    # wavefronts, in both pictures, and the two large members of the mixed batch
    ("jc60_4w", JC60, ground(60), 1.5, 12, (1, 4)),
    ("p1m1_ip_4w", P1M1, p1m1_initial(), 0.3, 8, (1, 4)),
    ("pcav1024", PCAV1024, basis(1024, 32), 0.4, 4, (0, 0)),
    ("modes4096", MODES4096, basis(4096, 0), 0.3, 3, (0, 0)),
])
def test_gsl_controller_lds_teams_match_oracle(name, spec, psi0, T, ntraj, layout):
    """ODE_EngineGSL (ODE_GSL.h:44-84) in the multi-wavefront engine: the GSL error ratio needs the
    derivative of the new state, formed in LDS by the whole team."""
    print(f"[gpu] gsl controller lds team {name} ntraj={ntraj}", flush=True)
    sq, so = Q.System(spec), O.System(spec)
    seed = 815
    iv = np.full(int(round(T / 0.1)), 0.1)
    ens = Q.Ensemble(sq, ntraj, psi0, Q.mcwf_pars(seed=seed, jump_log_cap=256, controller="gsl", layout=layout))
    ens.advance(iv)
    per, tot = ens.counters()
    assert tot[2] == 0 and tot[3] == 7 * tot[4]
    for k in range(ntraj):
        res_o = _oracle_traj(so, psi0, k, seed, iv, controller="gsl")
        psi_g, _ = ens.state(k)
        jt, jch, nj = ens.jumplog(k)
        _assert_traj_parity(res_o, psi_g, jt, jch, nj, per[k])
        assert per[k][3] == res_o["counters"][3]


def test_gsl_controller_is_refused_where_not_built():
    sq = Q.System(JC60)
    with pytest.raises(Q.QtError, match="odeint controller only"):
        Q.Ensemble(sq, 4, ground(60), Q.mcwf_pars(controller="gsl", group_lanes=16))


@pytest.mark.parametrize("prec", [Q.QT_FP64, Q.QT_FP32_DERIV])
def test_gsl_controller_pumped_mode_closed_form(prec):
    """GSL controller, both derivative precisions: PLM-20 from a coherent state stays coherent,
    <n> = |alpha(t)|^2 with alpha(t) = alpha_ss + (alpha_0 - alpha_ss) exp((-kappa + i Delta) t)."""
    print(f"[gpu] gsl closed form prec={prec}", flush=True)
    s = Q.System(PLM20)
    iv = Q.run_intervals(T=5.0, Dt=0.1)
    rows = Q.Ensemble(s, 64, coherent(0.5, 20), Q.mcwf_pars(seed=2, precision=prec, controller="gsl")).advance(iv)
    a0, D, eta, kappa = 0.5, -1.0, 2.0, 1.0
    ass = eta / (kappa - 1j * D)
    alpha = ass + (a0 - ass) * np.exp((-kappa + 1j * D) * rows[:, 0])
    err = np.abs(rows[:, 2] - np.abs(alpha) ** 2).max()
    print(f"  max |<n> - |alpha|^2| = {err:.2e}", flush=True)
    assert err < (1e-6 if prec == Q.QT_FP64 else 2e-5)
