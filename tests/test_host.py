"""Host-side product code on the CPU: the lowered ELL tables against the oracle, the C ABI
symbol table, error behaviour and the run-loop intervals (no GPU calls)."""
import ctypes
import os
import re

import numpy as np
import pytest

import cppqed_amd as Q
import oracle_lib as O
from sysdefs import JC60, PLM20, PTLA

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SPECS = {
    "ptla": PTLA,
    "plm20": PLM20,
    "jc60": JC60,
    "jc_ip_flavour_sch_qbit": "free qbit deltaA=-2 gamma=0.5 eta=0.3\nfree mode cutoff=5 delta=1 kappa=0.2 eta=0.7\nia jc 0 1 g=0.4 gI=0.2",
    "lossy_mode": "free mode cutoff=7 delta=-2 kappa=0.5 eta=0",
    "pumped_mode": "free mode cutoff=7 delta=-2 kappa=0 eta=1.5",
    "pcav_small": "free mode cutoff=4 delta=-0.5 kappa=1 eta=1\nfree particle fin=3 omrec=0.3\nia pcav 0 1 u0=-1 kcav=1 mf=cos",
    "pcav_sin": "free mode cutoff=3 delta=1 kappa=0.5 eta=0.5\nfree particle fin=4 omrec=1\nia pcav 0 1 u0=2 kcav=1 mf=sin",
    "pcav_sch_conf3": "free mode cutoff=3 delta=-9.5 kappa=10 eta=10\nfree particle fin=3 omrec=1 vclass=-10 kpart=1 mf=sin\n"
                      "ia pcav 0 1 u0=-1 kcav=1 mf=sin",
    "pcav_sch_conf2": "free mode cutoff=3 delta=-9.5 kappa=10 eta=10\nfree particle fin=3 omrec=1\n"
                      "ia pcav 0 1 u0=-1 kcav=1 mf=cos vclass=-10",
    # a pumped particle of another wavenumber than the cavity's: the chained product term
    "pcav_pumped_other_k": "free mode cutoff=3 delta=-1 kappa=1 eta=0.5\n"
                           "free particle fin=4 omrec=0.7 vclass=-2 kpart=1 mf=sin\nia pcav 0 1 u0=-1.5 kcav=2 mf=cos",
    "pcav_pumped_other_k_plus": "free mode cutoff=4 delta=0.5 kappa=0.5 eta=1\n"
                                "free particle fin=4 omrec=0.3 vclass=1.2 kpart=3 mf=plus\nia pcav 0 1 u0=0.8 kcav=1 mf=minus",
    "pumped_particle_sch": "free mode cutoff=3 delta=-1 kappa=1 eta=1\nfree particle fin=3 omrec=1 vclass=-10 kpart=1 mf=sin\n"
                           "ia pcav_orth 0 1 u0=-1",
    "bs_chain": "free mode cutoff=3 delta=-1 kappa=1 eta=1\nfree mode cutoff=4 delta=0.5 kappa=0.5 eta=0\n"
                "free mode cutoff=2 delta=0 kappa=1 eta=0\nia bs 0 1 g=0.5\nia bs 1 2 g=0.25 gI=0.5\nsystem composite",
    "zz_triangle": "free qbit deltaA=0 gamma=0.2 eta=0.7\nfree qbit deltaA=0.3 gamma=0.1 eta=0.7\n"
                   "free qbit deltaA=0 gamma=0.2 eta=0\nia zz 0 1 J=-0.5\nia zz 1 2 J=0.75\nia zz 0 2 J=0.25\nsystem composite",
}


def process(spec, raw):
    """Averaged::process of the frees, in spec order: VAR of n (Mode.cc:198-201), VAR(P) and
    DEV(X) of the particle (Particle.cc:163-167)."""
    out = raw.copy()
    frees = [l.split()[1] for l in spec.splitlines() if l.startswith("free")]
    for f, kind in enumerate(frees):
        b = 4 * f
        if kind in ("mode", "particle"):
            out[b + 1] -= out[b] ** 2
        if kind == "particle":
            out[b + 3] = np.sqrt(out[b + 3] - out[b + 2] ** 2)
    return out


@pytest.mark.parametrize("name", sorted(SPECS))
def test_tables_match_oracle(name):
    spec = SPECS[name]
    sq, so = Q.System(spec), O.System(spec)
    assert sq.dim == so.dim and sq.n_jumps == so.njumps and sq.n_avr == so.navr
    assert sq.highest_frequency == so.highest_frequency
    if "other_k" in name:
        # the chained term: the oracle applies firstH, then secondH to the result (as the reference);
        # the table holds their product, summed in another order -- equal to a few ulp
        H = so.dense_heff()
        np.testing.assert_allclose(sq.dense(0), H, rtol=0, atol=8 * 2.2e-16 * np.abs(H).max())
    else:
        np.testing.assert_array_equal(sq.dense(0), so.dense_heff())
    rng = np.random.default_rng(3)
    psi = rng.normal(size=sq.dim) + 1j * rng.normal(size=sq.dim)
    psi /= np.linalg.norm(psi)
    for m in range(sq.n_jumps):
        np.testing.assert_allclose(sq.dense(1 + m) @ psi, so.act_jump(psi, m), rtol=0, atol=1e-15)
    raw = np.array([np.vdot(psi, sq.dense(1 + sq.n_jumps + q) @ psi) for q in range(sq.n_avr)])
    # observable q streams Re or Im of <psi|O_q|psi>: the 4th of a mode's or qubit's block is Im<a>
    frees = [l.split()[1] for l in spec.splitlines() if l.startswith("free")]
    vals = np.array([raw[q].imag if q % 4 == 3 and frees[q // 4] != "particle" else raw[q].real
                     for q in range(sq.n_avr)])
    np.testing.assert_allclose(process(spec, vals), so.averages(psi), rtol=1e-12, atol=1e-12)


IP_SPECS = {
    "jc_ip": "free qbit deltaA=0.5 gamma=0.2 eta=0.3 picture=IP\nfree mode cutoff=6 delta=-1 kappa=1 eta=1 picture=IP\n"
             "ia jc 0 1 g=3",
    "plm_ip": "free mode cutoff=8 delta=-1 kappa=1 eta=2 picture=IP",
    "jc_uip": "free qbit deltaA=0.5 gamma=0.2 eta=0.3 picture=UIP\nfree mode cutoff=5 delta=-1 kappa=1 eta=1 picture=UIP\n"
              "ia jc 0 1 g=3 gI=0.5",
    "bs_mixed_pictures": "free mode cutoff=3 delta=-1 kappa=1 eta=1 picture=IP\nfree mode cutoff=4 delta=0.5 kappa=0.5 eta=0.2\n"
                         "ia bs 0 1 g=0.5\nsystem composite",
    "pcav_orth_ip": "free mode cutoff=3 delta=-10 kappa=10 eta=10 picture=IP\n"
                    "free particle fin=3 omrec=1 vclass=-10 kpart=1 mf=sin picture=IP\nia pcav_orth 0 1 u0=-1",
    "pcav_orth_cos": "free mode cutoff=4 delta=-2 kappa=1 eta=2 picture=IP\n"
                     "free particle fin=4 omrec=0.5 vclass=3 kpart=1 mf=cos picture=IP\nia pcav_orth 0 1 u0=2",
    "pcav_disp_ip": "free mode cutoff=4 delta=-0.5 kappa=1 eta=1 picture=IP\nfree particle fin=3 omrec=0.3 picture=IP\n"
                    "ia pcav 0 1 u0=-1 kcav=1 mf=cos",
    "p1m1_conf2": __import__("sysdefs").P1M1_CONF[2],
    "p1m1_conf3": __import__("sysdefs").P1M1_CONF[3],
    "p1m1_conf4": __import__("sysdefs").P1M1_CONF[4],
    "pcav_pumped_cos": "free mode cutoff=4 delta=-2 kappa=1 eta=2 picture=IP\n"
                       "free particle fin=4 omrec=0.5 vclass=3 kpart=2 mf=cos picture=IP\n"
                       "ia pcav 0 1 u0=2 kcav=2 mf=sin",
    "pcav_pumped_other_k_ip": "free mode cutoff=3 delta=-1 kappa=1 eta=0.5 picture=IP\n"
                              "free particle fin=4 omrec=0.7 vclass=-2 kpart=1 mf=sin picture=IP\n"
                              "ia pcav 0 1 u0=-1.5 kcav=2 mf=cos",
    "pcav_plus_vclass": "free mode cutoff=4 delta=-2 kappa=1 eta=2 picture=IP\n"
                        "free particle fin=4 omrec=0.5 picture=IP\nia pcav 0 1 u0=2 kcav=1 mf=plus vclass=1.5",
    "ring": __import__("sysdefs").RING,
    "ring_sin_cos": __import__("sysdefs").RING.replace("mf0=plus mf1=minus", "mf0=sin mf1=cos"),
}


@pytest.mark.parametrize("name", sorted(IP_SPECS))
def test_interaction_picture_tables_match_oracle(name):
    """H_nH/i(tau) assembled from the lowered entries and their frequencies, val * exp(freq tau), equals
    the oracle's propagated Tridiagonals (Tridiagonal.tcc:153-163) applied at t - t0 = tau."""
    spec = IP_SPECS[name]
    sq, so = Q.System(spec), O.System(spec)
    assert sq.info.time_dependent
    cols, vals, freqs = sq.table(0)
    rows = np.repeat(np.arange(sq.dim), cols.shape[1])
    rng = np.random.default_rng(11)
    psi = rng.normal(size=sq.dim) + 1j * rng.normal(size=sq.dim)
    for tau in (0.0, 0.13, 0.71, 0.05):
        H = np.zeros((sq.dim, sq.dim), np.complex128)
        np.add.at(H, (rows, cols.ravel()), (vals * np.exp(freqs * tau)).ravel())
        ref = so.apply_heff(psi, t=tau, t0=0.0)
        np.testing.assert_allclose(H @ psi, ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    assert sq.highest_frequency == so.highest_frequency


def test_particle_position_operator_is_the_fft_population():
    """<X> of the dense momentum-basis operator equals sum_j x_j |psi_x(j)|^2 with psi_x from numpy's
    inverse FFT (the GSL backward transform up to 1/N), halves swapped (LazyDensityOperatorFFT.cc:16-40);
    a particle with fin=3 has x_j = j*pi/4 - pi."""
    spec = "free particle fin=3 omrec=0.3"
    sq = Q.System(spec)
    rng = np.random.default_rng(5)
    psi = rng.normal(size=8) + 1j * rng.normal(size=8)
    psi /= np.linalg.norm(psi)
    px = np.fft.ifft(psi) * np.sqrt(8)
    px = np.concatenate([px[4:], px[:4]])
    x = np.arange(8) * np.pi / 4 - np.pi
    pop = np.abs(px) ** 2
    X, X2 = sq.dense(3), sq.dense(4)
    assert np.vdot(psi, X @ psi).real == pytest.approx((x * pop).sum(), abs=1e-13)
    assert np.vdot(psi, X2 @ psi).real == pytest.approx((x * x * pop).sum(), abs=1e-13)
    np.testing.assert_allclose(X, X.conj().T, atol=1e-15)
    # kinetic term: H/i = -i omrec k^2, k = n - 4
    np.testing.assert_allclose(np.diag(sq.dense(0)), -1j * 0.3 * (np.arange(8) - 4.0) ** 2, atol=0)


def test_mixed_builds_lower():
    """The four builds of the mixed batch lower to the expected dimensions, widths and channels."""
    from sysdefs import MIXED, mixed_counts
    want = {"pcav1024": (1024, 1, 5), "modes4096": (4096, 4, 9), "qubits64": (64, 6, 7), "jc60": (60, 2, 4)}
    for name, spec, idx in MIXED:
        s = Q.System(spec)
        d, nj, w = want[name]
        assert (s.dim, s.n_jumps, s.info.h_width) == (d, nj, w), name
        assert 0 <= idx < s.dim
    c = mixed_counts()
    assert c.sum() == 8192 and c.min() > 1900


def test_jc_ell_width_and_nnz():
    s = Q.System(JC60)
    assert s.info.h_width == 4
    cols, vals, _ = s.table(0)
    # every row has its diagonal first
    assert np.all(cols[:, 0] == np.arange(60))


def test_bad_specs_raise():
    for bad in ["", "free blob x=1", "free mode cutoff=3\nia jc 0 1 g=1", "free mode cutoff=3 picture=XY",
                "free qbit deltaA=1 gamma=1 eta=1\nfree qbit deltaA=1 gamma=1 eta=1\nia jc 0 0 g=1",
                # the chained pcav term with only the particle in the interaction picture (its
                # entries would rotate per intermediate momentum)
                "free mode cutoff=3 delta=-1 kappa=1 eta=0.5\nfree particle fin=4 omrec=0.7 vclass=-2 kpart=1 "
                "mf=sin picture=IP\nia pcav 0 1 u0=-1.5 kcav=2 mf=cos"]:
        with pytest.raises(Q.QtError):
            Q.System(bad)


def test_header_symbols_exported():
    """Every function include/qtraj.h declares is exported by libqtraj.so."""
    text = open(os.path.join(ROOT, "include", "qtraj.h")).read()
    names = set(re.findall(r"\b(qt_[a-z0-9_]+)\s*\(", text))
    lib = ctypes.CDLL(Q.LIBPATH)
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    assert len(names) >= 15


def _dense(cols, vals):
    d = cols.shape[0]
    m = np.zeros((d, d), complex)
    np.add.at(m, (np.repeat(np.arange(d), cols.shape[1]), cols.ravel()), vals.ravel())
    return m


@pytest.mark.parametrize("name,n_off", [("pcav1024", 5), ("modes4096", 9), ("plm20", 3), ("jc60", 0),
                                        ("qubits64", 0), ("jc_ip", 0)])
def test_constant_offset_form(name, n_off):
    """The device's re-slotted H (constant column offset per slot) is the same matrix as the ELL
    table, every slot's entries sit at one offset, and its empty places point at the clamped
    column; systems with more distinct offsets than the row width, or with picture
    frequencies, keep the ELL table (h_offsets = 0)."""
    from sysdefs import MIXED, PLM20
    specs = {n: sp for n, sp, _ in MIXED}
    specs.update(plm20=PLM20, jc60=JC60,
                 jc_ip="free qbit deltaA=0.5 gamma=0.2 eta=0.3 picture=IP\n"
                       "free mode cutoff=30 delta=-1 kappa=1 eta=1 picture=IP\nia jc 0 1 g=3")
    s = Q.System(specs[name])
    assert s.info.h_offsets == n_off
    if not n_off:
        with pytest.raises(Q.QtError):
            s.table(-1)
        return
    cols, vals, _ = s.table(0)
    oc, ov, _ = s.table(-1)
    assert oc.shape == cols.shape
    np.testing.assert_array_equal(_dense(oc, ov), _dense(cols, vals))
    rows = np.arange(s.dim)
    for k in range(oc.shape[1]):
        nz = ov[:, k] != 0
        offs = np.unique(oc[nz, k] - rows[nz])
        assert len(offs) <= 1, (k, offs)
        off = offs[0] if len(offs) else 0
        np.testing.assert_array_equal(oc[:, k], np.clip(rows + off, 0, s.dim - 1))


def test_null_handles_are_refused():
    """A null handle gives QT_EINVAL and a message, never a crash (no GPU call is reached)."""
    lib = ctypes.CDLL(Q.LIBPATH)
    lib.qt_last_error.restype = ctypes.c_char_p
    d = ctypes.c_double()
    i = ctypes.c_int()
    ll = (ctypes.c_longlong * 8)()
    calls = [
        ("qt_sys_info", (None, None)),
        ("qt_sys_table", (None, 0, None, None, None, None)),
        ("qt_mcwf_state", (None, 0, None, None)),
        ("qt_mcwf_jumplog", (None, 0, None, None, 0, ctypes.byref(i))),
        ("qt_mcwf_counters", (None, None, ll)),
        ("qt_mcwf_traj_averages", (None, None)),
        ("qt_mcwf_rho", (None, None)),
        ("qt_mcwf_kernel_ms", (None, ctypes.byref(d), ctypes.byref(i))),
        ("qt_master_kernel_ms", (None, ctypes.byref(d), ctypes.byref(i))),
        ("qt_batch_kernel_ms", (None, ctypes.byref(d), ctypes.byref(i))),
        ("qt_server_release", (None, ctypes.c_longlong(1))),
        ("qt_server_stats", (None, None)),
    ]
    for name, args in calls:
        assert getattr(lib, name)(*args) == 1, name  # QT_EINVAL


def test_run_intervals_mirror_trajectory_run():
    """Dt mode: advance(min(Dt, T - t)) while t < T -- 0.1 summed ten times falls short of 1
    and leaves a 1.1e-16 interval, as the reference's golden files show (Trajectory.h:562-567)."""
    iv = Q.run_intervals(T=1.0, Dt=0.1)
    assert len(iv) == 11 and iv[-1] == pytest.approx(1.1102230246251565e-16, rel=1e-6)
    assert len(Q.run_intervals(NDt=10, Dt=0.1)) == 10
    assert len(Q.run_intervals(T=20.0, Dt=0.1)) == 200 or len(Q.run_intervals(T=20.0, Dt=0.1)) == 201


def test_bench_refuses_several_gpus():
    """bench.py measures one GPU; --gpus N > 1 exits with a message instead of a number."""
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr and not r.stdout.strip()
