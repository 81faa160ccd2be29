"""qt_sys_build_terms, the generic boundary: systems handed over as Tridiagonal operator data
(quantumoperator::Tridiagonal, CPPQEDcore/quantumoperator/Tridiagonal.h:71-287) lower to the same
tables as the spec catalogue, bit for bit, for every element of the catalogue in both pictures;
the Tridiagonal algebra (direct product, Hermitian conjugate, frequencies) against dense
matrices; malformed terms are refused.  CPU only."""
import ctypes as C

import numpy as np
import pytest

import cppqed_amd as Q
from cppqed_amd import elements as E
from sysdefs import JC60, MIXED, P1M1, P1M1_CONF, PLM20, PTLA, RING
from test_host import IP_SPECS, SPECS


def _kv(toks):
    axes, kv = [], {}
    for t in toks:
        if "=" in t:
            k, v = t.split("=", 1)
            kv[k] = v
        else:
            axes.append(int(t))
    return axes, kv


def from_spec(spec):
    """the spec's elements as Tridiagonal data (the statements of qt_sys_build)"""
    frees, ias, forced = [], [], None
    for ln in spec.replace(";", "\n").splitlines():
        toks = ln.split()
        if not toks:
            continue
        what, kind = toks[0], toks[1]
        axes, kv = _kv(toks[2:])
        f = lambda k, d: float(kv.get(k, d))  # noqa: E731
        pic = kv.get("picture", "Sch")
        if what == "free":
            if kind == "mode":
                frees.append(E.mode(int(f("cutoff", 10)), f("delta", -10), f("kappa", 0),
                                    complex(f("eta", 0), f("etaI", 0)), pic))
            elif kind == "qbit":
                frees.append(E.qbit(f("deltaA", -10), f("gamma", 0), complex(f("eta", 0), f("etaI", 0)), pic))
            elif kind == "ptla":
                frees.append(E.pumped_two_level_atom_sch(f("gamma", 10), f("deltaA", -10), complex(f("eta", 0), f("etaI", 0))))
            elif kind == "particle":
                frees.append(E.particle(int(f("fin", 6)), f("omrec", 1.0), f("vclass", 0.0), int(f("kpart", 1)),
                                        kv.get("mf", "sin"), pic))
        elif what == "ia":
            fr = [frees[a] for a in axes]
            if kind == "jc":
                ias.append(E.jaynes_cummings(fr[0], fr[1], *axes, complex(f("g", 0), f("gI", 0))))
            elif kind == "bs":
                ias.append(E.coupled_modes(fr[0], fr[1], *axes, complex(f("g", 0), f("gI", 0))))
            elif kind == "zz":
                ias.append(E.zz(fr[0], fr[1], *axes, f("J", 0)))
            elif kind == "pcav":
                ias.append(E.particle_along_cavity(fr[0], fr[1], *axes, f("u0", 0), int(f("kcav", 1)), kv.get("mf", "cos"),
                                                   float(kv["vclass"]) if "vclass" in kv else None))
            elif kind == "pcav_orth":
                ias.append(E.particle_orthogonal_to_cavity(fr[0], fr[1], *axes, f("u0", -1.0)))
            elif kind == "ptm":
                ias.append(E.particle_two_modes(fr[0], fr[1], fr[2], *axes, f("u0", -1), f("u1", -1), f("phi", 0),
                                                int(f("k0", 1)), kv.get("mf0", "plus"), kv.get("mf1", "plus")))
        elif what == "system":
            forced = kind == "binary"
    return E.build_system(frees, ias, forced)


def _dense_freqs(s):
    cols, vals, freqs = s.table(0)
    F = np.zeros((s.dim, s.dim), np.complex128)
    for r in range(s.dim):
        for c, v, f in zip(cols[r], vals[r], freqs[r]):
            if v != 0:
                F[r, c] = f
    return F


ALL = dict(SPECS)
ALL.update(IP_SPECS)
ALL.update({"ptla_golden": PTLA, "plm20": PLM20, "p1m1": P1M1, "ring_golden": RING, "jc60": JC60,
            **{f"p1m1_conf{k}_golden": v for k, v in P1M1_CONF.items()},
            **{f"mixed_{n}": spec for n, spec, _ in MIXED if n != "modes4096"},
            "modes_small": "\n".join([f"free mode cutoff=4 delta=-1 kappa=1 eta={1 if i == 0 else 0}" for i in range(3)]
                                     + ["ia bs 0 1 g=0.5", "ia bs 1 2 g=0.5", "system composite"]),
            "jc_uip_qbit_ip_mode": "free qbit deltaA=0.5 gamma=0.2 eta=0.3 picture=UIP\n"
                                   "free mode cutoff=5 delta=-1 kappa=1 eta=1 picture=IP\nia jc 0 1 g=3"})


@pytest.mark.parametrize("name", sorted(ALL))
def test_terms_lower_like_the_spec(name):
    """Every table both routes produce is the same to the bit: H_nH/i's values and frequencies,
    every jump and average, the Exact exponents, the streaming rules and the system's data."""
    spec = ALL[name]
    a, b = Q.System(spec), from_spec(spec)
    assert (a.dim, a.dims, a.n_jumps, a.n_avr) == (b.dim, b.dims, b.n_jumps, b.n_avr)
    assert a.highest_frequency == b.highest_frequency
    assert a.info.time_dependent == b.info.time_dependent
    for w in range(1 + a.n_jumps + a.n_avr):
        np.testing.assert_array_equal(a.dense(w), b.dense(w), err_msg=f"table {w}")
    np.testing.assert_array_equal(_dense_freqs(a), _dense_freqs(b))
    for x, y in zip(a.rules(), b.rules()):
        np.testing.assert_array_equal(x, y)


def _dense_tri(t):
    """dense matrix of a Tridiagonal over its own dims at picture time 0, and of its frequencies"""
    n = int(np.prod(t.dims))
    M = np.zeros((n, n), complex)
    F = np.zeros((n, n), complex)
    strides = [int(np.prod(t.dims[a + 1:])) for a in range(t.rank)]
    for i, d in enumerate(t.diags):
        if d is None:
            continue
        codes = t.codes(i)
        for idx in np.ndindex(d.shape):
            r = sum((m + (t.k[a] if codes[a] == 1 else 0)) * strides[a] for a, m in enumerate(idx))
            c = sum((m + (t.k[a] if codes[a] == 2 else 0)) * strides[a] for a, m in enumerate(idx))
            M[r, c] += d[idx]
            if t.freqs[i] is not None:
                F[r, c] = t.freqs[i][idx]
    return M, F


def test_tridiagonal_algebra_against_dense_matrices():
    """Direct product = kron, dagger = conjugate transpose (transposeIndex), the frequencies of
    furnishWithFreqs are main(row) - main(col) (Tridiagonal.cc:25-37) and add under the direct
    product (Tridiagonal.tcc:27-65); the unary constructor's consistency check."""
    rng = np.random.default_rng(5)
    a = E.Tridiagonal.unary(zero=rng.normal(size=4) + 1j * rng.normal(size=4), k=1,
                            minus=rng.normal(size=3) + 1j, plus=rng.normal(size=3) - 2j)
    b = E.Tridiagonal.unary(k=2, minus=[1 + 1j, 2.0, 3j], plus=[0.5, -1j, 2.0])
    Ma, _ = _dense_tri(a)
    Mb, _ = _dense_tri(b)
    np.testing.assert_array_equal(_dense_tri(a * b)[0], np.kron(Ma, Mb))
    np.testing.assert_array_equal(_dense_tri((a * b).dagger())[0], np.kron(Ma, Mb).conj().T)
    np.testing.assert_array_equal(_dense_tri(a + a.dagger())[0], Ma + Ma.conj().T)
    main_a = np.array([0.3 + 1j, -0.2 + 0.5j, 1.5j, 0.7])
    main_b = np.array([1j, 2 + 0.1j, -0.4j, 0.25, 3j])
    fa = E.aop(4).furnish_with_freqs(main_a)
    fb = E.aop(5).dagger().furnish_with_freqs(main_b)
    M, F = _dense_tri(fa * fb)
    for (r, c) in zip(*np.nonzero(M)):
        ra, rb, ca, cb = r // 5, r % 5, c // 5, c % 5
        assert F[r, c] == (main_a[ra] - main_a[ca]) + (main_b[rb] - main_b[cb])
    with pytest.raises(ValueError):
        E.Tridiagonal.unary(zero=[1, 2, 3], k=0, minus=[1, 2])
    with pytest.raises(ValueError):
        E.Tridiagonal.unary(zero=[1, 2, 3], k=1, minus=[1])


def test_a_new_element_built_from_tridiagonals():
    """An element the spec catalogue does not have -- a Kerr mode, H/i = -i chi n(n-1) plus a
    drive, and a cross-Kerr coupling n (x) n -- builds through the terms and its tables are the
    textbook matrices."""
    chi, cross, eta = 0.3, 0.15, 0.8 + 0.2j
    dim = 6
    n = np.arange(dim, dtype=float)
    kerr = E._diag(dim, [complex(0.0, -chi * k * (k - 1)) for k in n])
    drive = E.aop(dim).dagger().scaled(eta)
    h = E._with_k(kerr, 1) + drive + drive.dagger().scaled(-1.0)
    m1 = E.Free(dim, [h], [[E.aop(dim).scaled(1.0)]], [[E._diag(dim, n)]], [0], [-1], [0], None, 1.0)
    m2 = E.mode(3, -1.0, 0.5, 0.0)
    xk = E.Interaction((0, 1), [(E._diag(dim, n) * E._diag(3, [0.0, 1.0, 2.0])).scaled(complex(0.0, -cross))])
    s = E.build_system([m1, m2], [xk])
    a = np.diag(np.sqrt(np.arange(1, dim)), 1)
    a2 = np.diag(np.sqrt(np.arange(1, 3)), 1)
    H1 = -1j * chi * np.diag(n * (n - 1)) + eta * a.T - np.conj(eta) * a
    H2 = np.diag(-(0.5 + 1j) * np.arange(3))
    want = np.kron(H1, np.eye(3)) + np.kron(np.eye(dim), H2) - 1j * cross * np.kron(np.diag(n), np.diag(np.arange(3.0)))
    np.testing.assert_allclose(s.dense(0), want, rtol=0, atol=1e-15)
    np.testing.assert_allclose(s.dense(1), np.kron(a, np.eye(3)), rtol=0, atol=0)
    assert s.n_jumps == 2 and s.n_avr == 5


def _raw_desc(**over):
    """a minimal one-free description, with fields overridden, for the refusal checks"""
    d = np.ascontiguousarray([0.0, 0.0, 1.0, 0.0])
    dp = (C.c_void_p * 3)(d.ctypes.data, None, None)
    axes = np.array([0], np.int32)
    k = np.array([0], np.int32)
    tri = E._Tri(1, axes.ctypes.data, k.ctypes.data, C.cast(dp, C.c_void_p).value, None)
    for key, v in over.pop("tri", {}).items():
        setattr(tri, key, v)
    ta = (E._Tri * 1)(tri)
    fd = (E._FreeDesc * 1)(E._FreeDesc(over.pop("dim", 2), None))
    desc = E._SysDesc(1, C.cast(fd, C.c_void_p).value, E._Op(1, C.cast(ta, C.c_void_p).value, 0, None),
                      0, None, 0, None, None, None, None, 1.0)
    for key, v in over.items():
        setattr(desc, key, v)
    return desc, (d, dp, axes, k, ta, fd)


def test_malformed_terms_are_refused():
    l = Q.lib()
    l.qt_sys_build_terms.argtypes = [C.POINTER(E._SysDesc), C.POINTER(C.c_void_p)]
    h = C.c_void_p()
    desc, keep = _raw_desc()
    assert l.qt_sys_build_terms(C.byref(desc), C.byref(h)) == 0
    Q.System.adopt(h)
    bad_axes = np.array([3], np.int32)
    neg_k = np.array([-1], np.int32)
    for over in ({"n_frees": 0}, {"dim": 0}, {"tri": {"axes": bad_axes.ctypes.data}}, {"tri": {"rank": 2}},
                 {"tri": {"k": neg_k.ctypes.data}}, {"highest_frequency": float("nan")}):
        desc, keep = _raw_desc(**over)
        h = C.c_void_p()
        assert l.qt_sys_build_terms(C.byref(desc), C.byref(h)) == Q.QT_EINVAL, over
    # an off-diagonal with offset 0: the reference's "Tridiagonal consistency error"
    d = np.ascontiguousarray([1.0, 0.0, 1.0, 0.0])
    dp = (C.c_void_p * 3)(None, d.ctypes.data, None)
    desc, keep = _raw_desc(tri={"diags": C.cast(dp, C.c_void_p).value})
    assert l.qt_sys_build_terms(C.byref(desc), C.byref(h)) == Q.QT_EINVAL
    assert "Tridiagonal consistency error" in l.qt_last_error().decode()
