"""The elements of the mixed-batch builds pinned against textbook operators (CPU only).

Each system's H_nH/i and jump operators are assembled here from kron products of the textbook
single-mode / qubit / particle operators, with the definitions the reference states for them:
  mode          H = -delta a+a + i (eta a+ - eta* a),      J = sqrt(2 kappa) a     (Mode.cc:56-86, 118-125)
  qubit         H = -deltaA s+s + i (eta* s - eta s+),     J = sqrt(2 gamma) s     (Qbit_.h:99-118: a two-level
                mode pumped with -eta; s = |0><1|, |1> excited; PumpedTwoLevelAtom.cc:42-49 writes the
                same 2x2 matrix out)
  particle      H = omrec k^2, k_n = n - N/2 in the momentum basis               (Particle.cc:29-48, 89-92)
  beam splitter H/i = conj(u) a1 a2+ - u a1+ a2                                  (CoupledModes.cc:36-41:
                tridiagMinusHC(conj(u) aop(m1) aop(m2).dagger()))
  dispersive    H = (u0/2) n (x) (+-cos(2 kcav x)), + for a cos, - for a sin mode function,
                cos(Kx) = (e^{iKx} + e^{-iKx})/2 with e^{iKx}|k> = |k+K>         (ParticleCavity.cc:26-32)
  pumped particle  H += (vClass/2) (+-cos(2 kPart x)), + for cos, - for sin (sin^2 = (1 - cos 2kx)/2
                with the constant dropped; none for e^{+-ikx})                  (Particle.cc:84-87)
  pumped particle along a cavity of another wavenumber (ParticleCavity.cc:102-140):
                H += etaeff (a (x) f_p^dagger f_c + a^dagger (x) f_p f_c^dagger), etaeff = sign(u0)
                sqrt(u0 vClass), f = cos(kx), sin(kx), e^{ikx}, e^{-ikx} as operators on the truncated
                momentum grid, multiplied in the order the reference applies them (secondH after
                firstH, secondHT after firstHT): at the grid's edges f_p f_c^dagger differs from
                (f_p^dagger f_c)^dagger, since truncated shifts do not commute there
  zz            H = J sz (x) sz with sz = [s+, s] = diag(-1, +1)  (no reference element: qbit::sigmazop is
                declared, never defined, Qbit_.h:68-70; this build defines the coupling so)
and H_nH/i = -i (H - (i/2) sum_m J_m+ J_m).

Both the oracle (oracle_lib, per-slice Tridiagonal algebra) and the product's lowered ELL tables
(qt_sys_table) are compared with these matrices, so a wrong conjugation, sign or axis order in
either fails here.  A mutated build (conjugated u, flipped J, sin for cos) must NOT match: the
mistake is planted in the oracle's input, on the CPU."""
import numpy as np
import pytest
import scipy.sparse as sp

import cppqed_amd as Q
import oracle_lib as O
from sysdefs import MODES4096, PCAV1024, QUBITS64, qubits64_spec


def _a(n):
    return sp.diags(np.sqrt(np.arange(1, n, dtype=float)), 1, shape=(n, n), dtype=complex, format="csr")


def _sig():
    return sp.csr_matrix(np.array([[0, 1], [0, 0]], complex))


def _kron_all(ops):
    out = ops[0]
    for o in ops[1:]:
        out = sp.kron(out, o, format="csr")
    return out


def _embed(dims, ops_by_axis):
    """kron over the axes (axis 0 most significant, the reference's row-major layout)"""
    return _kron_all([ops_by_axis.get(i, sp.identity(d, complex, format="csr")) for i, d in enumerate(dims)])


def _shift(n, K):
    """e^{iKx} in the momentum basis: |k> -> |k+K>, i.e. M[m+K, m] = 1"""
    return sp.diags(np.ones(n - K), -K, shape=(n, n), dtype=complex, format="csr")


def textbook(frees, ias):
    """(H_nH/i, [J_m]) as sparse matrices.  frees: list of dicts (kind + parameters); ias: list of
    (kind, axes, parameters).  Jumps are ordered frees first, in axis order (BinarySystem/Composite)."""
    dims = [f["dim"] for f in frees]
    D = int(np.prod(dims))
    H = sp.csr_matrix((D, D), dtype=complex)
    Js = []
    for i, f in enumerate(frees):
        if f["kind"] == "mode":
            a = _a(f["dim"])
            h = -f["delta"] * (a.getH() @ a) + 1j * (f["eta"] * a.getH() - np.conj(f["eta"]) * a)
            H = H + _embed(dims, {i: h})
            if f["kappa"]:
                Js.append(_embed(dims, {i: np.sqrt(2 * f["kappa"]) * a}))
        elif f["kind"] == "qbit":
            s = _sig()
            h = -f["deltaA"] * (s.getH() @ s) + 1j * (np.conj(f["eta"]) * s - f["eta"] * s.getH())
            H = H + _embed(dims, {i: h})
            if f["gamma"]:
                Js.append(_embed(dims, {i: np.sqrt(2 * f["gamma"]) * s}))
        elif f["kind"] == "particle":
            k = np.arange(f["dim"]) - f["dim"] // 2
            H = H + _embed(dims, {i: sp.diags(f["omrec"] * k.astype(float) ** 2, 0, dtype=complex, format="csr")})
            if f.get("vclass") and f["mf"] in ("sin", "cos"):  # |e^{+-ikx}|^2 = 1: no pump term
                K = 2 * f["kpart"]
                cos = 0.5 * (_shift(f["dim"], K) + _shift(f["dim"], K).getH())
                sgn = -1.0 if f["mf"] == "sin" else 1.0
                H = H + _embed(dims, {i: 0.5 * f["vclass"] * sgn * cos})
        else:
            raise ValueError(f["kind"])
    for kind, ax, p in ias:
        if kind == "bs":
            a1, a2 = _a(dims[ax[0]]), _a(dims[ax[1]])
            u = p["u"]
            hoi = (np.conj(u) * _embed(dims, {ax[0]: a1, ax[1]: a2.getH()})
                   - u * _embed(dims, {ax[0]: a1.getH(), ax[1]: a2}))
            H = H + 1j * hoi
        elif kind == "zz":
            s = _sig()
            sz = s.getH() @ s - s @ s.getH()
            H = H + p["J"] * _embed(dims, {ax[0]: sz, ax[1]: sz})
        elif kind == "disp":
            m, q = ax
            n = sp.diags(np.arange(dims[m], dtype=float), 0, dtype=complex, format="csr")
            K = 2 * p["kcav"]
            cos = 0.5 * (_shift(dims[q], K) + _shift(dims[q], K).getH())
            sgn = -1.0 if p["mf"] == "sin" else 1.0
            H = H + 0.5 * p["u0"] * sgn * _embed(dims, {m: n, q: cos})
        elif kind == "pcav_other_k":
            m, q = ax
            fp = frees[q]
            etaeff = np.sign(p["u0"]) * np.sqrt(p["u0"] * fp["vclass"])
            fc, fpart = _mode_function(dims[q], p["mf"], p["kcav"]), _mode_function(dims[q], fp["mf"], fp["kpart"])
            a = _a(dims[m])
            H = H + etaeff * (_embed(dims, {m: a, q: fpart.getH() @ fc}) + _embed(dims, {m: a.getH(), q: fpart @ fc.getH()}))
        else:
            raise ValueError(kind)
    Hnh = H - 0.5j * sum((J.getH() @ J for J in Js), sp.csr_matrix((D, D), dtype=complex))
    return (-1j * Hnh).tocsr(), Js


def _mode_function(n, mf, k):
    """cos(kx), sin(kx), e^{ikx}, e^{-ikx} on the momentum grid (e^{ikx} shifts |q> to |q+k>, truncated)"""
    e = _shift(n, k)
    if mf == "cos":
        return 0.5 * (e + e.getH())
    if mf == "sin":
        return (e - e.getH()) / 2j
    return e if mf == "plus" else e.getH()


def _table_sparse(sq, which):
    cols, vals, _ = sq.table(which)
    d = sq.dim
    rows = np.repeat(np.arange(d), cols.shape[1])
    return sp.csr_matrix((vals.ravel(), (rows, cols.ravel())), shape=(d, d))


def _max_dev(A, B):
    return abs(A - B).max() if (A - B).nnz else 0.0


def check_against_textbook(spec, frees, ias, n_probe=3):
    """Oracle and product against the textbook matrices: H_nH/i applied to random vectors (oracle)
    and entry by entry (product table); every jump operator (product table) and J psi (oracle)."""
    Hs, Js = textbook(frees, ias)
    d = Hs.shape[0]
    so, sq = O.System(spec), Q.System(spec)
    assert so.dim == d and sq.dim == d
    assert sq.n_jumps == len(Js) and so.njumps == len(Js)
    scale = abs(Hs).max()
    rng = np.random.default_rng(17)
    for _ in range(n_probe):
        psi = rng.normal(size=d) + 1j * rng.normal(size=d)
        np.testing.assert_allclose(so.apply_heff(psi), Hs @ psi, rtol=0, atol=1e-12 * scale * np.sqrt(d))
        for m, J in enumerate(Js):
            np.testing.assert_allclose(so.act_jump(psi, m), J @ psi, rtol=0, atol=1e-12 * np.sqrt(d))
    assert _max_dev(_table_sparse(sq, 0), Hs) <= 1e-13 * scale
    for m, J in enumerate(Js):
        assert _max_dev(_table_sparse(sq, 1 + m), J) <= 1e-13


def _mode(dim, delta, kappa, eta):
    return {"kind": "mode", "dim": dim, "delta": delta, "kappa": kappa, "eta": eta}


def _qbit(deltaA, gamma, eta):
    return {"kind": "qbit", "dim": 2, "deltaA": deltaA, "gamma": gamma, "eta": eta}


# ---------------------------------------------------------------- beam splitter (config 2b)

BS_SMALL = ("free mode cutoff=3 delta=-1 kappa=1 eta=1\n"
            "free mode cutoff=4 delta=0.3 kappa=0.5 eta=0\n"
            "free mode cutoff=2 delta=0.7 kappa=0.2 eta=0.4\n"
            "ia bs 0 1 g=0.5 gI=0.3\nia bs 1 2 g=-0.7 gI=0.2\nsystem composite")
BS_SMALL_FREES = [_mode(3, -1, 1, 1), _mode(4, 0.3, 0.5, 0), _mode(2, 0.7, 0.2, 0.4)]
BS_SMALL_IAS = [("bs", (0, 1), {"u": 0.5 + 0.3j}), ("bs", (1, 2), {"u": -0.7 + 0.2j})]


def test_beam_splitter_chain_textbook():
    """Three modes of different cutoffs, two RWA beam splitters with complex u."""
    check_against_textbook(BS_SMALL, BS_SMALL_FREES, BS_SMALL_IAS)


def test_beam_splitter_binary_textbook():
    spec = "free mode cutoff=5 delta=-0.2 kappa=0.3 eta=0.6\nfree mode cutoff=3 delta=1.1 kappa=0.9 eta=0\nia bs 0 1 g=0.8 gI=-0.4"
    check_against_textbook(spec, [_mode(5, -0.2, 0.3, 0.6), _mode(3, 1.1, 0.9, 0)], [("bs", (0, 1), {"u": 0.8 - 0.4j})])


def test_modes4096_textbook():
    """Config 2b as committed: four modes at cutoff 8 (dim 4096), beam splitters g = 0.5 between
    neighbours, kappa = 1 and deltaC = -1 on every mode, the first pumped with eta = 1."""
    frees = [_mode(8, -1, 1, 1 if i == 0 else 0) for i in range(4)]
    ias = [("bs", (i, i + 1), {"u": 0.5}) for i in range(3)]
    check_against_textbook(MODES4096, frees, ias, n_probe=2)


def test_beam_splitter_mutation_is_caught():
    """A conjugated coupling (u -> u*) or a swapped pair of axes changes H: the textbook pin sees it."""
    Hs, _ = textbook(BS_SMALL_FREES, BS_SMALL_IAS)
    rng = np.random.default_rng(3)
    psi = rng.normal(size=Hs.shape[0]) + 1j * rng.normal(size=Hs.shape[0])
    for bad in (BS_SMALL.replace("gI=0.3", "gI=-0.3"), BS_SMALL.replace("ia bs 0 1", "ia bs 1 0")):
        try:
            so = O.System(bad)
        except ValueError:
            continue  # a swapped pair of unequal cutoffs may not even build
        assert np.abs(so.apply_heff(psi) - Hs @ psi).max() > 1e-3


# ---------------------------------------------------------------- ZZ + transverse field (config 2c)

def _qubits_textbook(spec_seed=20240611):
    rng = np.random.default_rng(spec_seed)
    Jij = {}
    for i in range(6):
        for j in range(i + 1, 6):
            Jij[(i, j)] = rng.uniform(-1, 1)
    frees = [_qbit(0, 0.2, 0.7)] * 6
    ias = [("zz", k, {"J": v}) for k, v in Jij.items()]
    return frees, ias


def test_qubits64_textbook():
    """Config 2c as committed: six qubits, ZZ couplings J_ij ~ U(-1, 1) (numpy PCG64 seed 20240611)
    on all 15 pairs, transverse drive eta = 0.7, decay gamma = 0.2, deltaA = 0."""
    frees, ias = _qubits_textbook()
    check_against_textbook(QUBITS64, frees, ias)
    # the spec carries the same couplings, in the same order
    assert qubits64_spec() == QUBITS64


def test_zz_pair_textbook():
    spec = "free qbit deltaA=0.4 gamma=0.1 eta=0.3\nfree qbit deltaA=-0.2 gamma=0.5 eta=0\nia zz 0 1 J=0.37"
    check_against_textbook(spec, [_qbit(0.4, 0.1, 0.3), _qbit(-0.2, 0.5, 0)], [("zz", (0, 1), {"J": 0.37})])


def test_zz_mutation_is_caught():
    spec = "free qbit deltaA=0.4 gamma=0.1 eta=0.3\nfree qbit deltaA=-0.2 gamma=0.5 eta=0\nia zz 0 1 J=0.37"
    Hs, _ = textbook([_qbit(0.4, 0.1, 0.3), _qbit(-0.2, 0.5, 0)], [("zz", (0, 1), {"J": 0.37})])
    psi = np.random.default_rng(4).normal(size=4) + 0j
    so = O.System(spec.replace("J=0.37", "J=-0.37"))
    assert np.abs(so.apply_heff(psi) - Hs @ psi).max() > 0.1
    # a qubit drive entering with +eta (a mode's sign) instead of -eta is caught as well
    Hm, _ = textbook([{**_qbit(0.4, 0.1, 0.3), "eta": -0.3}, _qbit(-0.2, 0.5, 0)], [("zz", (0, 1), {"J": 0.37})])
    assert np.abs(so.apply_heff(psi) - Hm @ psi).max() > 0.1


# ---------------------------------------------------------------- dispersive particle-cavity term (config 2a)

def _particle(fin, omrec):
    return {"kind": "particle", "dim": 1 << fin, "omrec": omrec}


@pytest.mark.parametrize("mf", ["cos", "sin"])
@pytest.mark.parametrize("kcav", [1, 2])
def test_dispersive_textbook(mf, kcav):
    spec = (f"free mode cutoff=4 delta=-0.5 kappa=1 eta=1\nfree particle fin=3 omrec=0.3\n"
            f"ia pcav 0 1 u0=-1.3 kcav={kcav} mf={mf}")
    check_against_textbook(spec, [_mode(4, -0.5, 1, 1), _particle(3, 0.3)],
                           [("disp", (0, 1), {"u0": -1.3, "kcav": kcav, "mf": mf})])


def test_pcav1024_textbook():
    """Config 2a as committed: mode cutoff 16 (deltaC = -0.5 after folding U0/2, kappa = 1, eta = 1)
    (x) particle with 64 momentum states (omrec = 0.1), cos^2 coupling U0 = -1 (unpumped particle:
    no interference term)."""
    check_against_textbook(PCAV1024, [_mode(16, -0.5, 1, 1), _particle(6, 0.1)],
                           [("disp", (0, 1), {"u0": -1.0, "kcav": 1, "mf": "cos"})], n_probe=2)


def test_dispersive_mutation_is_caught():
    """The sign flip between cos and sin mode functions (ParticleCavity.cc:30) is seen."""
    spec = "free mode cutoff=4 delta=-0.5 kappa=1 eta=1\nfree particle fin=3 omrec=0.3\nia pcav 0 1 u0=-1.3 kcav=1 mf=cos"
    Hs, _ = textbook([_mode(4, -0.5, 1, 1), _particle(3, 0.3)], [("disp", (0, 1), {"u0": -1.3, "kcav": 1, "mf": "cos"})])
    psi = np.random.default_rng(5).normal(size=32) + 0j
    so = O.System(spec.replace("mf=cos", "mf=sin"))
    assert np.abs(so.apply_heff(psi) - Hs @ psi).max() > 0.1


# ---------------------------------------------------------------- pumped particle along a cavity of another wavenumber

def _pumped_particle(fin, omrec, vclass, kpart, mf):
    return {"kind": "particle", "dim": 1 << fin, "omrec": omrec, "vclass": vclass, "kpart": kpart, "mf": mf}


@pytest.mark.parametrize("pmf,cmf,kpart,kcav", [("sin", "cos", 1, 2), ("cos", "sin", 2, 1), ("plus", "minus", 3, 1),
                                                ("sin", "plus", 1, 3)])
def test_pumped_particle_other_wavenumber_textbook(pmf, cmf, kpart, kcav):
    spec = (f"free mode cutoff=3 delta=-1 kappa=1 eta=0.5\n"
            f"free particle fin=4 omrec=0.7 vclass=-2 kpart={kpart} mf={pmf}\n"
            f"ia pcav 0 1 u0=-1.5 kcav={kcav} mf={cmf}")
    ias = [("pcav_other_k", (0, 1), {"u0": -1.5, "kcav": kcav, "mf": cmf})]
    if cmf in ("cos", "sin"):
        ias.append(("disp", (0, 1), {"u0": -1.5, "kcav": kcav, "mf": cmf}))
    check_against_textbook(spec, [_mode(3, -1, 1, 0.5), _pumped_particle(4, 0.7, -2, kpart, pmf)], ias)


def test_pumped_particle_other_wavenumber_mutation_is_caught():
    """The particle's mode function enters conjugated (secondH = mfNKX(particle)^dagger,
    ParticleCavity.cc:107): the oracle given the conjugate function (minus for plus) must not match."""
    frees = [_mode(3, -1, 1, 0.5), _pumped_particle(4, 0.7, -2, 3, "plus")]
    Hs, _ = textbook(frees, [("pcav_other_k", (0, 1), {"u0": -1.5, "kcav": 1, "mf": "cos"}),
                             ("disp", (0, 1), {"u0": -1.5, "kcav": 1, "mf": "cos"})])
    spec = ("free mode cutoff=3 delta=-1 kappa=1 eta=0.5\nfree particle fin=4 omrec=0.7 vclass=-2 kpart=3 mf=minus\n"
            "ia pcav 0 1 u0=-1.5 kcav=1 mf=cos")
    psi = np.random.default_rng(6).normal(size=48) + 0j
    assert np.abs(O.System(spec).apply_heff(psi) - Hs @ psi).max() > 0.1
