"""The allocation-free advance (qt_mcwf_advance_noalloc) against qt_mcwf_advance: the same rows,
states and counters, on two calls in a row (the second on a smaller ensemble with fewer stream
intervals, so stale scratch contents would show), and the size limit refused with an error."""
import numpy as np
import pytest

import cppqed_amd as Q
from sysdefs import JC60, PLM20, coherent, ground

pytestmark = pytest.mark.gpu


def _pair(spec, n, psi0, pars):
    s = Q.System(spec)
    return Q.Ensemble(s, n, psi0, pars), Q.Ensemble(s, n, psi0, pars)


def _same(a, b):
    ca, cb = a.counters()[0], b.counters()[0]
    np.testing.assert_array_equal(ca, cb)
    for i in (0, a.n // 2, a.n - 1):
        pa, ta = a.state(i)
        pb, tb = b.state(i)
        np.testing.assert_array_equal(pa, pb)
        assert ta == tb


def test_noalloc_equals_advance():
    print("[gpu] noalloc: JC-60 512 traj x 30 intervals, then PLM-20 64 traj x 7 intervals", flush=True)
    a, b = _pair(JC60, 512, ground(60), Q.mcwf_pars(seed=5, precision=Q.QT_FP32_DERIV, jump_log_cap=64))
    iv = np.full(30, 0.1)
    ra = a.advance(iv)
    rb = b.advance_noalloc(iv)
    np.testing.assert_array_equal(ra, rb)
    _same(a, b)
    # a second call continues the first ensemble (its state is resident) ...
    np.testing.assert_array_equal(a.advance(iv[:4]), b.advance_noalloc(iv[:4]))
    # ... and a smaller input on another ensemble reuses the scratch the first calls left behind
    c, d = _pair(PLM20, 64, coherent(0.5, 20), Q.mcwf_pars(seed=9))
    iv2 = np.full(7, 0.25)
    rc, rd = c.advance(iv2), d.advance_noalloc(iv2)
    assert rd.shape == (7, 2 + c.system.n_avr)
    np.testing.assert_array_equal(rc, rd)
    _same(c, d)


def test_noalloc_refuses_beyond_its_scratch():
    print("[gpu] noalloc size limit", flush=True)
    e = Q.Ensemble(Q.System(JC60), 4, ground(60), Q.mcwf_pars(seed=1))
    n_avr = e.system.n_avr
    too_many = (8 << 20) // (256 * (n_avr + 1)) + 1
    with pytest.raises(Q.QtError, match="scratch"):
        e.advance_noalloc(np.full(too_many, 1e-3))
    # the ensemble is untouched by the refused call and still runs
    rows = e.advance_noalloc(np.full(3, 0.1))
    assert rows.shape == (3, 2 + n_avr) and rows[-1, 0] == pytest.approx(0.3)
