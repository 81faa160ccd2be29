"""Interaction picture on the GPU (Exact U, time-dependent H entries, the stale picture origin
after a step-back of MCWF_Trajectory.h:305): parity with the oracle per trajectory, the
addContribution drop-in at t - t0, and the pumped lossy mode's coherent closed form."""
import numpy as np
import pytest

import cppqed_amd as Q
import oracle_lib as O
from sysdefs import (P1M1, P1M1_CONF, PCAV_OTHER_K, PCAV_OTHER_K_IP, RING, basis, coherent, ground, p1m1_initial,
                     ring_initial)

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

JC_IP = ("free qbit deltaA=0.5 gamma=0.2 eta=0.3 picture=IP\n"
         "free mode cutoff=30 delta=-1 kappa=1 eta=1 picture=IP\n"
         "ia jc 0 1 g=3")
JC_UIP = ("free qbit deltaA=0.5 gamma=0.2 eta=0.3 picture=UIP\n"
          "free mode cutoff=12 delta=-1 kappa=1 eta=1 picture=UIP\n"
          "ia jc 0 1 g=3")
PLM_IP = "free mode cutoff=20 delta=-1 kappa=1 eta=2 picture=IP"


@pytest.mark.parametrize("spec", [JC_IP, JC_UIP, PLM_IP])
@pytest.mark.parametrize("prec,tol", [(Q.QT_FP64, 1e-13), (Q.QT_FP32_DERIV, 2e-6)])
def test_apply_heff_interaction_picture(spec, prec, tol):
    sq, so = Q.System(spec), O.System(spec)
    rng = np.random.default_rng(3)
    n = 17
    psi = rng.normal(size=(n, sq.dim)) + 1j * rng.normal(size=(n, sq.dim))
    x = torch.from_numpy(psi).to("cuda:0")
    for t, t0 in [(0.37, 0.1), (1.25, 1.0)]:
        dpsi = torch.zeros_like(x)
        sq.apply_heff(x.data_ptr(), dpsi.data_ptr(), n=n, precision=prec, t=t, t0=t0)
        torch.cuda.synchronize()
        got = dpsi.cpu().numpy()
        so2 = O.System(spec)  # fresh Tridiagonals: the oracle's propagation is cumulative
        for i in range(n):
            ref = so2.apply_heff(psi[i], t=t, t0=t0)
            assert np.abs(got[i] - ref).max() <= tol * np.abs(ref).max(), (i, t, t0)


@pytest.mark.parametrize("layout", [(0, 0), (1, 4)])
@pytest.mark.parametrize("name,spec,psi0,T,ntraj,ovt", [
    ("jc_ip", JC_IP, ground(60), 2.0, 24, 10.0),
    ("jc_ip_stepbacks", JC_IP, ground(60), 2.0, 16, 1.0),
    ("jc_uip", JC_UIP, ground(24), 2.0, 16, 10.0),
    ("plm_ip", PLM_IP, coherent(0.5, 20), 2.0, 16, 10.0),
    # the 1particle1mode system whose golden runs pin the oracle (tests/test_oracle.py)
    ("p1m1", P1M1, p1m1_initial(), 0.5, 24, 10.0),
    ("ring", RING, ring_initial(), 0.5, 16, 10.0),
    # ParticleAlongCavity with the interference terms (1particle1mode --1p1mconf 2, 3)
    ("p1m1_conf2", P1M1_CONF[2], p1m1_initial(), 0.5, 16, 10.0),
    ("p1m1_conf3", P1M1_CONF[3], p1m1_initial(), 0.5, 16, 10.0),
    # a pumped particle of another wavenumber (the chained product term), both pictures
    ("pcav_other_k", PCAV_OTHER_K, basis(24, 4), 1.0, 16, 10.0),
    ("pcav_other_k_ip", PCAV_OTHER_K_IP, basis(24, 4), 1.0, 16, 10.0),
])
def test_ip_trajectories_match_oracle(name, spec, psi0, T, ntraj, ovt, layout):
    """With overshootTol = 1 dpLimit step-backs are frequent, and the redone step runs from the
    stale picture origin as the reference's does (MCWF_Trajectory.h:305)."""
    print(f"[gpu] interaction picture parity {name} layout={layout}", flush=True)
    sq, so = Q.System(spec), O.System(spec)
    seed = 424242
    iv = np.full(int(round(T / 0.1)), 0.1)
    ens = Q.Ensemble(sq, ntraj, psi0, Q.mcwf_pars(seed=seed, jump_log_cap=256, layout=layout, overshootTol=ovt))
    rows = ens.advance(iv)
    per, tot = ens.counters()
    for k in range(ntraj):
        res = O.mcwf_run(so, psi0, O.params(rng="philox", seed=seed, stream=k, overshootTol=ovt),
                         O.run_spec("NDt", Dt=0.1, NDt=len(iv)))
        psi_g, _ = ens.state(k)
        jt, jch, nj = ens.jumplog(k)
        assert nj == len(res["jump_t"]), (k, nj, len(res["jump_t"]))
        np.testing.assert_array_equal(jch, res["jump_ch"])
        np.testing.assert_allclose(jt, res["jump_t"], rtol=1e-10, atol=1e-12)
        err = np.linalg.norm(psi_g - res["psi"]) / np.linalg.norm(res["psi"])
        assert err <= 1e-9, (k, err)
        for c in (0, 1, 2, 4, 5, 7):
            assert per[k][c] == res["counters"][c], (k, c, per[k][c], res["counters"][c])
    ro, _ = O.ensemble_run(so, psi0, O.params(rng="philox", seed=seed, stream=0, overshootTol=ovt),
                           O.run_spec("NDt", Dt=0.1, NDt=len(iv)), ntraj)
    np.testing.assert_allclose(rows, ro[1:], rtol=1e-9, atol=1e-9)
    if name == "jc_ip_stepbacks":
        assert tot[5] > 0, tot  # step-backs happened: the stale picture origin was exercised


def test_plm20_interaction_picture_closed_form():
    """PLM-20 in the interaction picture, 1000 trajectories to T=10: <n> = |alpha(t)|^2 as in the
    Schroedinger picture (the occupation is picture independent; every trajectory stays coherent)."""
    print("[gpu] PLM-20 IP 1000 trajectories closed form", flush=True)
    s = Q.System(PLM_IP)
    iv = Q.run_intervals(T=10.0, Dt=0.1)
    rows = Q.Ensemble(s, 1000, coherent(0.5, 20), Q.mcwf_pars(seed=0, precision=Q.QT_FP32_DERIV)).advance(iv)
    a0, D, eta, kappa = 0.5, -1.0, 2.0, 1.0
    ass = eta / (kappa - 1j * D)
    alpha = ass + (a0 - ass) * np.exp((-kappa + 1j * D) * rows[:, 0])
    err = np.abs(rows[:, 2] - np.abs(alpha) ** 2).max()
    print(f"  max |<n> - |alpha|^2| = {err:.2e}", flush=True)
    assert err < 2e-5
