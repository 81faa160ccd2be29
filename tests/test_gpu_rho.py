"""The ensemble's density operator on request (qt_mcwf_rho; AverageTrajectories<MCWF_Trajectory>,
MCWF_Trajectory.h:406-420, with StateVector::addTo, StateVector.h:173-180) against numpy's
(1/N) sum_k psi_k psi_k^dagger of the same device states, and against the trajectories' own
averages, which are linear in rho: mean_k <psi_k|O|psi_k> = Tr(O rho)."""
import numpy as np
import pytest

import cppqed_amd as Q
from sysdefs import JC60, MODES4096, basis, ground

pytestmark = pytest.mark.gpu


def _numpy_rho(psi):
    return np.einsum("ki,kj->ij", psi, psi.conj()) / len(psi)


@pytest.mark.parametrize("n", [1, 17, 600, 5000])
def test_rho_matches_numpy_and_averages(n):
    print(f"[gpu] rho: JC-60 {n} trajectories", flush=True)
    s = Q.System(JC60)
    ens = Q.Ensemble(s, n, ground(60), Q.mcwf_pars(seed=11, precision=Q.QT_FP64))
    ens.advance(np.full(10, 0.1))
    rho = ens.rho()
    _, psi, _, _ = ens.get_states()
    # n-term FP64 sums in two orders: differences within the summation bound n eps
    assert np.abs(rho - _numpy_rho(psi)).max() <= 4 * max(n, 4) * 2.2e-16
    assert abs(np.trace(rho) - 1) <= 1e-12
    assert np.abs(rho - rho.conj().T).max() <= 1e-15
    assert np.array_equal(ens.rho(), rho)  # a fixed-order sum: the same bits on every call
    # raw averages of the last stream point, trajectory by trajectory
    avg = ens.traj_averages().mean(axis=0)
    _, imag, _, _ = s.rules()
    for q in range(s.n_avr):
        tr = np.trace(s.dense(1 + s.n_jumps + q) @ rho)
        want = tr.imag if imag[q] else tr.real
        assert abs(want - avg[q]) <= 1e-12 * max(1.0, abs(avg[q])), (q, want, avg[q])


def test_rho_large_dimension():
    print("[gpu] rho: four modes dim 4096, 40 trajectories", flush=True)
    s = Q.System(MODES4096)
    ens = Q.Ensemble(s, 40, basis(4096, 0), Q.mcwf_pars(seed=3, precision=Q.QT_FP64))
    ens.advance(np.full(5, 0.1))
    rho = ens.rho()
    _, psi, _, _ = ens.get_states()
    assert np.abs(rho - _numpy_rho(psi)).max() <= 1e-14
    assert abs(np.trace(rho) - 1) <= 1e-12
