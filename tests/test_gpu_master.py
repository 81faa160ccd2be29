"""GPU parity of the Master sweep -- the FP64-MFMA kernel (d <= 96) and the sparse kernel (any d up
to 4096) -- against the CPU oracle and the reference's golden output."""
import numpy as np
import pytest

import cppqed_amd as Q
import oracle_lib as O
from sysdefs import PTLA, jc_master_spec
from test_oracle import load_golden

pytestmark = pytest.mark.gpu


KERNELS = ["mfma", "sparse"]


def _golden_master(spec, rho0, mode, Dt, T, NDt, dc, kernel="mfma"):
    """rows of the device Master (one point) in the run loop's dc, Dt or NDt mode, from t = 0"""
    m = Q.MasterSweep([Q.System(spec)], rho0, epsRel=1e-6, epsAbs=1e-5, kernel=kernel)
    first, _ = m.advance([0.0])
    if mode == "dc":
        rest = m.advance_dc(T, dc, max_rows=5)
    else:
        rest = m.advance(Q.run_intervals(T=T, Dt=Dt, NDt=NDt if mode == "ndt" else 0))[0][0]
    return np.concatenate([first[0], rest])


def _compare(rows, g, dev_col=None):
    assert rows.shape == g.shape, (rows.shape, g.shape)
    tol = np.full(g.shape[1], 1e-8)
    if dev_col is not None:
        tol[dev_col] = 1e-6   # the particle's DEV(X): root of a cancellation for a one-grid-point state
    bad = np.abs(rows - g) > 1e-5 * np.abs(g) + tol
    assert not bad.any(), (np.argwhere(bad)[:5], rows[bad][:5], g[bad][:5])


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("mode", ["dc", "dt", "ndt"])
def test_master_ptla_golden(mode, kernel):
    """Testing/expected/run_PTLA_CPPQED_{dc,dt,ndt}_master replayed by the device Master (d=2, one tile)."""
    print(f"[gpu] master ptla golden {mode} {kernel}", flush=True)
    rows = _golden_master(PTLA, np.diag([1.0, 0.0]).astype(complex), mode, 0.1, 1.0, 10, 10, kernel)
    _compare(rows, load_golden(f"run_PTLA_CPPQED_{mode}_master"))


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("mode", ["dc", "dt", "ndt"])
def test_master_1particle1mode_golden(mode, kernel):
    """run_1particle1mode_{dc,dt,ndt}_master: the unitary interaction picture on the device Master --
    H(tau) = E H E^-1 around the MFMA product and rho <- U rho U^dagger after every step
    (Master.h:186-193) -- for the pumped particle, the particle-cavity term and the lossy mode."""
    from sysdefs import P1M1_MASTER, p1m1_initial
    print(f"[gpu] master 1particle1mode golden {mode} {kernel}", flush=True)
    psi = p1m1_initial()
    rows = _golden_master(P1M1_MASTER, np.outer(psi, psi.conj()), mode, 0.01, 0.03, 3, 1, kernel)
    g = load_golden(f"run_1particle1mode_{mode}_master")
    _compare(rows, g, dev_col=g.shape[1] - 1)


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("conf", [2, 3, 4])
def test_master_1particle1mode_conf_golden(conf, kernel):
    """run_1particle1mode_ndt_conf{2,3,4}_master (--dc 1 --T 0.03 with the Master): ParticleAlongCavity."""
    from sysdefs import P1M1_CONF, p1m1_initial
    print(f"[gpu] master 1particle1mode conf{conf} golden {kernel}", flush=True)
    psi = p1m1_initial()
    spec = P1M1_CONF[conf].replace("eta=10 picture=IP", "eta=10 picture=UIP")
    rows = _golden_master(spec, np.outer(psi, psi.conj()), "dc", 0.01, 0.03, 3, 1, kernel)
    g = load_golden(f"run_1particle1mode_ndt_conf{conf}_master")
    _compare(rows, g, dev_col=g.shape[1] - 1)


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("mode", ["dc", "dt", "ndt"])
def test_master_ring_golden(mode, kernel):
    """run_Ring_{dc,dt,ndt}_master: a composite of three Exact frees (d = 72) in the unitary
    interaction picture on the device Master."""
    from sysdefs import RING_MASTER, ring_initial
    print(f"[gpu] master Ring golden {mode} {kernel}", flush=True)
    psi = ring_initial()
    rows = _golden_master(RING_MASTER, np.outer(psi, psi.conj()), mode, 0.0001, 0.0002, 2, 1, kernel)
    _compare(rows, load_golden(f"run_Ring_{mode}_master"), dev_col=5)


def test_master_picture_refusals():
    """The Master takes the unitary interaction picture only (applicableInMaster, Master.h:73)."""
    ip = Q.System("free mode cutoff=10 delta=-1 kappa=1 eta=2 picture=IP")
    rho = np.zeros((10, 10), complex)
    rho[0, 0] = 1
    with pytest.raises(Q.QtError, match="SystemNotApplicable"):
        Q.MasterSweep([ip], rho)
    with pytest.raises(Q.QtError, match="SystemNotApplicable"):
        Q.MasterSweep([ip], rho, kernel="sparse")


@pytest.mark.parametrize("cutoff,etas,T,kernel", [
    (4, [0.6, 2.0], 1.0, "mfma"), (8, [1.0, 3.0, 5.0], 0.5, "mfma"), (40, [0.5, 7.0], 0.2, "mfma"),
    (8, [1.0, 3.0, 5.0], 0.5, "sparse"), (40, [0.5, 7.0], 0.2, "sparse"),
    # beyond the MFMA kernel's 96: the sparse kernel is the automatic choice
    (60, [0.5, 7.0], 0.2, "auto"), (100, [2.0, 9.0, 15.0], 0.1, "auto")])
def test_master_sweep_matches_oracle(cutoff, etas, T, kernel):
    print(f"[gpu] master sweep cutoff={cutoff} etas={etas} kernel={kernel}", flush=True)
    specs = [jc_master_spec(cutoff, eta=e, g=10.0).replace("gamma=0", "gamma=0.3") for e in etas]
    sq = [Q.System(sp) for sp in specs]
    d = sq[0].dim
    rho0 = np.zeros((d, d), complex)
    rho0[0, 0] = 1
    iv = np.full(int(round(T / 0.05)), 0.05)
    m = Q.MasterSweep(sq, rho0, kernel=kernel)
    rows, dn = m.advance(iv)
    _, _, _, cnt = m.status()
    for k, sp in enumerate(specs):
        so = O.System(sp)
        rows_o, rho_o = O.master_run(so, rho0, O.params(), O.run_spec("NDt", Dt=0.05, NDt=len(iv)))
        rho_g = m.rho(k)
        err = np.linalg.norm(rho_g - rho_o) / np.linalg.norm(rho_o)
        assert err < 1e-9, (k, err)
        np.testing.assert_allclose(rows[k][:, 2:], rows_o[1:, 2:], rtol=1e-8, atol=1e-10)
        assert np.allclose(rho_g, rho_g.conj().T, atol=0)
        assert abs(np.trace(rho_g) - 1) < 1e-13


@pytest.mark.parametrize("picture", ["", " picture=UIP"])
@pytest.mark.parametrize("cutoff", [4, 12, 20, 30, 40, 48])
def test_master_tile_counts_match_oracle(cutoff, picture):
    """The MFMA kernel at each tile count NB = 1..6 (d = 8, 24, 40, 60, 80, 96), in the Schroedinger
    and the unitary interaction picture, from a density operator with coherences everywhere: the
    RK slots are kept as lower 16 x 16 tiles and mirrored into the upper ones, so every lower-tile
    index and every mirrored entry takes part; rho must match the oracle's and stay Hermitian."""
    print(f"[gpu] master tile counts cutoff={cutoff} picture={picture.strip() or 'S'}", flush=True)
    spec = (f"free qbit deltaA=0.5 gamma=0.2 eta=0.3{picture}\n"
            f"free mode cutoff={cutoff} delta=-1 kappa=1 eta=1.5{picture}\n"
            "ia jc 0 1 g=3")
    sq = Q.System(spec)
    d = sq.dim
    rng = np.random.default_rng(cutoff)
    psi = rng.normal(size=d) + 1j * rng.normal(size=d)
    psi /= np.linalg.norm(psi)
    rho0 = np.outer(psi, psi.conj())
    iv = np.full(4, 0.05)
    m = Q.MasterSweep([sq], rho0, kernel="mfma")
    rows, _ = m.advance(iv)
    rows_o, rho_o = O.master_run(O.System(spec), rho0, O.params(), O.run_spec("NDt", Dt=0.05, NDt=len(iv)))
    rho_g = m.rho(0)
    err = np.linalg.norm(rho_g - rho_o) / np.linalg.norm(rho_o)
    assert err < 1e-9, err
    np.testing.assert_allclose(rows[0][:, 2:], rows_o[1:, 2:], rtol=1e-8, atol=1e-10)
    assert np.allclose(rho_g, rho_g.conj().T, atol=0)
    assert abs(np.trace(rho_g) - 1) < 1e-13


def test_master_sweep_autostop_matches_oracle():
    """The reference's autostop criterion (AutostopHandlerGeneric, Trajectory.h:179-222) per sweep
    point: every point stops at the oracle's stop row, on the oracle's density operator, and the
    point's rows up to there are the oracle's."""
    print("[gpu] master sweep autostop", flush=True)
    kappas = [0.5, 1.0, 2.0, 3.0]  # the relaxation rate sets the stop time (oracle: 20.7, 10.8, 5.7, 4.1)
    specs = [f"free mode cutoff=20 delta=-1 kappa={k} eta=2" for k in kappas]
    rho0 = np.zeros((20, 20), complex)
    rho0[0, 0] = 1
    auto = (1e-5, 1e-12, 5)
    m = Q.MasterSweep([Q.System(sp) for sp in specs], rho0, epsRel=1e-8, epsAbs=1e-12, autostop=auto)
    iv = np.full(300, 0.1)
    rows, _ = m.advance(iv)
    done, tss, _, _ = m.status()
    assert done.all()
    for k, sp in enumerate(specs):
        ro, rho_o = O.master_run_autostop(O.System(sp), rho0, O.params(epsRel=1e-8, epsAbs=1e-12),
                                          O.run_spec("Dt", Dt=0.1, T=30.0), *auto)
        assert abs(tss[k] - ro[-1, 0]) < 1e-9, (k, tss[k], ro[-1, 0])
        n = len(ro) - 1  # the oracle's rows include t = 0
        # the device's FP64 MFMA sums in another order, so the step-size controller's accept/grow
        # decisions may tip differently near its thresholds: the averages agree to the ODE
        # tolerance, the last step size (dtDid, column 1) need not
        np.testing.assert_allclose(rows[k, :n, 0], ro[1:, 0], rtol=0, atol=1e-12)
        np.testing.assert_allclose(rows[k, :n, 2:], ro[1:, 2:], rtol=1e-6, atol=1e-9)
        assert (rows[k, :n, 1] > 0).all()
        # both integrate to the absolute ODE tolerance 1e-8 (epsRel after the swap of ODE.h:205)
        # along their own step sequences
        assert np.abs(m.rho(k) - rho_o).max() < 1e-7
    assert len(set(np.round(tss, 6))) > 1  # the points stop at different times


def test_master_kernels_agree():
    """The two kernels on one 16-point sweep at d = 80: the same averages, norms and states to the
    ODE tolerance (they sum in different orders; the step-size controller may tip)."""
    print("[gpu] master kernels agree", flush=True)
    etas = np.linspace(0.5, 12.0, 16)
    sq = [Q.System(jc_master_spec(40, eta=float(e), g=10.0)) for e in etas]
    rho0 = np.zeros((80, 80), complex)
    rho0[0, 0] = 1
    iv = np.full(20, 0.1)
    out = {}
    for kernel in KERNELS:
        m = Q.MasterSweep(sq, rho0, epsRel=1e-10, epsAbs=1e-10, kernel=kernel)
        rows, dn = m.advance(iv)
        out[kernel] = (rows, dn, m.status()[0], [m.rho(k) for k in range(len(sq))])
    (ra, da, sa, pa), (rb, db, sb, pb) = out["mfma"], out["sparse"]
    np.testing.assert_allclose(ra[..., 0], rb[..., 0], rtol=0, atol=1e-12)
    np.testing.assert_allclose(ra[..., 2:], rb[..., 2:], rtol=1e-7, atol=1e-9)
    np.testing.assert_allclose(da, db, rtol=1e-5, atol=1e-9)
    np.testing.assert_array_equal(sa, sb)
    for a, b in zip(pa, pb):
        assert np.abs(a - b).max() < 1e-8
