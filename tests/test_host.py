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
}


@pytest.mark.parametrize("name", sorted(SPECS))
def test_tables_match_oracle(name):
    spec = SPECS[name]
    sq, so = Q.System(spec), O.System(spec)
    assert sq.dim == so.dim and sq.n_jumps == so.njumps and sq.n_avr == so.navr
    assert sq.highest_frequency == so.highest_frequency
    np.testing.assert_array_equal(sq.dense(0), so.dense_heff())
    rng = np.random.default_rng(3)
    psi = rng.normal(size=sq.dim) + 1j * rng.normal(size=sq.dim)
    psi /= np.linalg.norm(psi)
    for m in range(sq.n_jumps):
        np.testing.assert_allclose(sq.dense(1 + m) @ psi, so.act_jump(psi, m), rtol=0, atol=1e-15)
    raw = np.array([np.vdot(psi, sq.dense(1 + sq.n_jumps + q) @ psi) for q in range(sq.n_avr)])
    # observable q streams Re or Im of <psi|O_q|psi>; every 4th block is (a, a, Re, Im) per free
    vals = np.array([raw[q].imag if q % 4 == 3 else raw[q].real for q in range(sq.n_avr)])
    proc = vals.copy()
    for f in range(sq.n_avr // 4):
        if "mode" in spec.splitlines()[f]:
            proc[4 * f + 1] -= proc[4 * f] ** 2
    np.testing.assert_allclose(proc, so.averages(psi), rtol=1e-12, atol=1e-12)


def test_jc_ell_width_and_nnz():
    s = Q.System(JC60)
    assert s.info.h_width == 4
    cols, vals, _ = s.table(0)
    # every row has its diagonal first
    assert np.all(cols[:, 0] == np.arange(60))


def test_bad_specs_raise():
    for bad in ["", "free blob x=1", "free mode cutoff=3\nia jc 0 1 g=1", "free mode cutoff=3 picture=XY",
                "free qbit deltaA=1 gamma=1 eta=1\nfree qbit deltaA=1 gamma=1 eta=1\nia jc 0 0 g=1"]:
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


def test_run_intervals_mirror_trajectory_run():
    """Dt mode: advance(min(Dt, T - t)) while t < T -- 0.1 summed ten times falls short of 1
    and leaves a 1.1e-16 interval, as the reference's golden files show (Trajectory.h:562-567)."""
    iv = Q.run_intervals(T=1.0, Dt=0.1)
    assert len(iv) == 11 and iv[-1] == pytest.approx(1.1102230246251565e-16, rel=1e-6)
    assert len(Q.run_intervals(NDt=10, Dt=0.1)) == 10
    assert len(Q.run_intervals(T=20.0, Dt=0.1)) == 200 or len(Q.run_intervals(T=20.0, Dt=0.1)) == 201
