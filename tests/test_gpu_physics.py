"""Ensemble physics at the workload sizes, on the GPU: PLM-20 with 1000 trajectories (keys
0..999) against the coherent-state closed form and the Master solution; driven JC-60 with 16384
trajectories, FP64 and FP32-derivative, against the FP64 Master solution (3 sigma of the
ensemble mean) and against each other (the FP32 variant's observable error)."""
import numpy as np
import pytest

import cppqed_amd as Q
from sysdefs import JC60, PLM20, coherent, ground

pytestmark = pytest.mark.gpu


def test_plm20_1000_trajectories_closed_form_and_master():
    print("[gpu] PLM-20 1000 trajectories T=10 vs closed form and Master", flush=True)
    s = Q.System(PLM20)
    iv = Q.run_intervals(T=10.0, Dt=0.1)
    psi0 = coherent(0.5, 20)
    rows = Q.Ensemble(s, 1000, psi0, Q.mcwf_pars(seed=0, precision=Q.QT_FP64)).advance(iv)
    a0, D, eta, kappa = 0.5, -1.0, 2.0, 1.0
    ass = eta / (kappa - 1j * D)
    alpha = ass + (a0 - ass) * np.exp((-kappa + 1j * D) * rows[:, 0])
    # every trajectory stays coherent (jumps do not change a coherent state): no sampling error
    err_cf = np.abs(rows[:, 2] - np.abs(alpha) ** 2).max()
    mrows, _ = Q.MasterSweep([s], np.outer(psi0, psi0.conj())).advance(iv)
    err_m = np.abs(rows[:, 2] - mrows[0, :, 2]).max()
    print(f"  max |<n>_ens - |alpha|^2| = {err_cf:.2e}, max |<n>_ens - <n>_master| = {err_m:.2e}", flush=True)
    assert err_cf < 2e-5 and err_m < 2e-5
    np.testing.assert_allclose(rows[:, 4], alpha.real, atol=2e-5)
    np.testing.assert_allclose(rows[:, 5], alpha.imag, atol=2e-5)


def rows_until(t, tmax):
    return int(np.searchsorted(t, tmax + 1e-9))


def test_jc60_16k_ensemble_vs_master_and_fp32_error():
    print("[gpu] JC-60 16384 trajectories T=20, FP64 and FP32-derivative, vs Master", flush=True)
    s = Q.System(JC60)
    n = 16384
    iv = Q.run_intervals(T=20.0, Dt=0.1)
    r64 = Q.Ensemble(s, n, ground(60), Q.mcwf_pars(seed=1001, precision=Q.QT_FP64)).advance(iv)
    r32 = Q.Ensemble(s, n, ground(60), Q.mcwf_pars(seed=1001, precision=Q.QT_FP32_DERIV)).advance(iv)
    rho0 = np.zeros((60, 60), complex)
    rho0[0, 0] = 1
    mr, _ = Q.MasterSweep([s], rho0, epsRel=1e-8).advance(iv)
    # columns: time, dtDid, qubit (P0, P1, Re s, Im s), mode (<n>, VAR n, Re a, Im a)
    nM = mr[0, :, 6]
    sig = np.sqrt(r64[:, 7] / n)
    pick = [int(round(t / 0.1)) - 1 for t in np.arange(2.0, 20.01, 2.0)]
    z64 = np.abs(r64[pick, 6] - nM[pick]) / sig[pick]
    z32 = np.abs(r32[pick, 6] - nM[pick]) / sig[pick]
    d32 = np.abs(r32[:, 6] - r64[:, 6])
    dP = np.abs(r32[:, 3] - r64[:, 3])
    print(f"  <n>_master(T) = {nM[-1]:.5f}, sigma of the mean = {sig[-1]:.2e}", flush=True)
    print(f"  |FP64 ens - Master| / sigma at t=2..20: {np.round(z64, 2).tolist()}", flush=True)
    print(f"  |FP32 ens - Master| / sigma at t=2..20: {np.round(z32, 2).tolist()}", flush=True)
    print(f"  FP32-derivative vs FP64 ensemble (same draws): max |d<n>| = {d32.max():.2e}, "
          f"mean |d<n>| = {d32.mean():.2e}, max |dP0| = {dP.max():.2e}", flush=True)
    assert z64.max() < 3.0 and z32.max() < 3.0
    # with the same draws the FP32-derivative trajectories follow the FP64 ones to rounding level
    # until the first jump decision flips (t ~ 2.5 for these keys); after that flipped trajectories
    # are independent samples and the ensembles differ by sampling noise only
    early = rows_until(r64[:, 0], 2.0)
    assert d32[:early].max() < 1e-7, d32[:early].max()
    assert (d32 <= 3 * np.sqrt(2) * np.sqrt(r64[:, 7] / n) + 1e-7).all()
