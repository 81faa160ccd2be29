"""Elements as Tridiagonal operator data, and systems built from them (qt_sys_build_terms).

`Tridiagonal` restates quantumoperator::Tridiagonal<R> (CPPQEDcore/quantumoperator/Tridiagonal.h:
71-287) on numpy arrays: the unary constructor with its frequencies, the direct product, scaling,
the Hermitian conjugate and the sum.  The element functions below give the catalogue's elements
in that form -- each free's H/i, jumps and averages, each interaction's H/i terms -- and
`build_system` hands frees and interactions to the device library as operator data, the way a
caller with its own elements (anything made of Tridiagonals) would.

Every complex product and sum runs element by element in Python's complex arithmetic (no fused
multiply-add), in the order the spec lowering of qt_sys_build uses, so the two routes lower the
catalogue to the same tables bit for bit.
"""
import cmath
import ctypes as C
import math

import numpy as np

from . import System, check, lib


def _cl(a):
    return [complex(x) for x in np.asarray(a, np.complex128).ravel()]


class Tridiagonal:
    """Tridiagonal<R>: dims and offsets K per axis; diags[i] / freqs[i] (i < 3^R, base-3 digits
    with axis 0 the most significant: 0 main, 1 |n+K><n|, 2 |n><n+K|) are arrays of extent dim or
    dim - K per axis, or None for an empty diagonal.  An entry at picture time tau is
    diag exp(freq tau) (Tridiagonal::propagate, Tridiagonal.tcc:154-163)."""

    def __init__(self, dims, k, diags, freqs=None):
        self.dims = tuple(int(d) for d in dims)
        self.k = tuple(int(x) for x in k)
        n = 3 ** len(self.dims)
        freqs = freqs if freqs is not None else [None] * n
        if len(diags) != n or len(freqs) != n:
            raise ValueError("a Tridiagonal of rank R has 3^R diagonals")
        self.diags = [None if d is None else np.asarray(d, np.complex128).reshape(self.extent(i)) for i, d in enumerate(diags)]
        self.freqs = [None if f is None else np.asarray(f, np.complex128).reshape(self.extent(i)) for i, f in enumerate(freqs)]

    @property
    def rank(self):
        return len(self.dims)

    def codes(self, i):
        c = []
        for _ in range(self.rank):
            c.append(i % 3)
            i //= 3
        return c[::-1]

    def extent(self, i):
        return tuple(d if c == 0 else max(0, d - k) for d, k, c in zip(self.dims, self.k, self.codes(i)))

    @classmethod
    def unary(cls, zero=(), k=0, minus=(), plus=(), to_freqs=False):
        """Tridiagonal<1>(zero, k, minus, plus, toFreqs) (Tridiagonal.cc:47-70); with to_freqs the
        main diagonal becomes the frequencies of furnishWithFreqs(-zero)."""
        zero, minus, plus = _cl(zero), _cl(minus), _cl(plus)
        dim = max(len(zero), len(minus) + k, len(plus) + k)
        if ((minus or plus) and not k) or (zero and minus and len(zero) - k != len(minus)) or \
                (zero and plus and len(zero) - k != len(plus)) or (minus and plus and len(minus) != len(plus)):
            raise ValueError("Tridiagonal consistency error")
        t = cls((dim,), (k,), [None if to_freqs or not zero else zero, minus or None, plus or None])
        if to_freqs:
            t = t.furnish_with_freqs([-z for z in zero])
        return t

    def furnish_with_freqs(self, main):
        """furnishWithFreqs (Tridiagonal.cc:25-37): freqs(1)(n) = main(n+K) - main(n), freqs(2) = -freqs(1)."""
        if self.rank != 1:
            raise ValueError("furnishWithFreqs works in the unary case only")
        main = _cl(main)
        k, dim = self.k[0], self.dims[0]
        t = Tridiagonal(self.dims, self.k, self.diags, self.freqs)
        if dim - k > 0:
            f1 = [main[n + k] - main[n] for n in range(dim - k)]
            t.freqs[1] = np.array(f1, np.complex128)
            t.freqs[2] = np.array([-f for f in f1], np.complex128)
        return t

    def __mul__(self, o):
        """The direct product Tridiagonal(t1, t2) (Tridiagonal.tcc:27-65): diagonal i*3^R2 + j is the
        outer product of diagonals i and j, its frequencies their outer sum (empty if either is)."""
        if not isinstance(o, Tridiagonal):
            return self.scaled(o)
        diags, freqs = [], []
        for i in range(3 ** self.rank):
            for j in range(3 ** o.rank):
                d1, d2, f1, f2 = self.diags[i], o.diags[j], self.freqs[i], o.freqs[j]
                diags.append(None if d1 is None or d2 is None else _outer(d1, d2, lambda a, b: a * b))
                freqs.append(None if f1 is None or f2 is None else _outer(f1, f2, lambda a, b: a + b))
        return Tridiagonal(self.dims + o.dims, self.k + o.k, diags, freqs)

    def scaled(self, c):
        """operator*= (Tridiagonal.tcc:68-74): the diagonals times c, frequencies untouched."""
        c = complex(c)
        return Tridiagonal(self.dims, self.k, [None if d is None else _map(d, lambda a: a * c) for d in self.diags],
                           self.freqs)

    def dagger(self):
        """hermitianConjugate (Tridiagonal.tcc:78-106): diagonal ind goes to transposeIndex(ind),
        conjugated; the frequencies stay where they are (they are their own transpose)."""
        n = 3 ** self.rank
        diags = [None] * n
        for ind in range(n):
            if self.diags[ind] is not None:
                t = sum(((3 - c) % 3) * 3 ** (self.rank - 1 - a) for a, c in enumerate(self.codes(ind)))
                diags[t] = _map(self.diags[ind], lambda a: a.conjugate())
        return Tridiagonal(self.dims, self.k, diags, self.freqs)

    def __add__(self, o):
        """operator+= (Tridiagonal.tcc:110-140): offsets must agree where both have off-diagonals."""
        if self.dims != o.dims:
            raise ValueError("Tridiagonal dimension mismatch")
        k = []
        for a, b in zip(self.k, o.k):
            if a and b and a != b:
                raise ValueError("Tridiagonal structure mismatch")
            k.append(a or b)
        t = Tridiagonal(self.dims, k, [None] * 3 ** self.rank)
        for i in range(3 ** self.rank):
            d1, d2 = self.diags[i], o.diags[i]
            t.diags[i] = d2 if d1 is None else d1 if d2 is None else _zip(d1, d2, lambda a, b: a + b)
            f1, f2 = self.freqs[i], o.freqs[i]
            if f1 is not None and f2 is not None and np.any(f1 != f2):
                raise ValueError("Tridiagonal structure mismatch")
            t.freqs[i] = f1 if f1 is not None else f2
        return t

    def __neg__(self):
        return self.scaled(-1.0)

    def __sub__(self, o):
        return self + (-o)


def _map(a, f):
    return np.array([f(x) for x in _cl(a)], np.complex128).reshape(a.shape)


def _zip(a, b, f):
    return np.array([f(x, y) for x, y in zip(_cl(a), _cl(b))], np.complex128).reshape(a.shape)


def _outer(a, b, f):
    bl = _cl(b)
    return np.array([f(x, y) for x in _cl(a) for y in bl], np.complex128).reshape(a.shape + b.shape)


class Sparse:
    """An operator that is no Tridiagonal (qt_coo): entries (row, col, value) over its dims, each
    rotating with its frequency in the interaction picture when freqs is given."""

    def __init__(self, dims, rows, cols, vals, freqs=None):
        self.dims = tuple(dims)
        self.rows = np.asarray(rows, np.int64)
        self.cols = np.asarray(cols, np.int64)
        self.vals = np.asarray(vals, np.complex128)
        self.freqs = None if freqs is None else np.asarray(freqs, np.complex128)


class Free:
    """A free element: dimension, H/i terms (list of Tridiagonal), jumps, averages (each a list of
    terms), the streaming rules of its averages, the Exact exponents (or None) and its highest
    frequency (QuantumSystem::highestFrequency)."""

    def __init__(self, dim, h=(), jumps=(), avr=(), imag=(), var_of=(), sqrt=(), exact=None, hf=0.0):
        self.dim, self.h, self.jumps, self.avr = dim, list(h), list(jumps), list(avr)
        self.imag, self.var_of, self.sqrt = list(imag), list(var_of), list(sqrt)
        self.exact = None if exact is None else np.asarray(exact, np.complex128)
        self.hf = hf


class Interaction:
    """H/i terms on the axes `axes` (Act<V...>'s V) and the interaction's highest frequency."""

    def __init__(self, axes, h, hf=0.0):
        self.axes, self.h, self.hf = tuple(axes), list(h), hf


# ---------------------------------------------------------------- element operators

def _nonzero(z):
    return z.real != 0.0 or z.imag != 0.0


def _with_freqs(t, free):
    """the free's Exact exponents as the operator's frequencies (a unary Tridiagonal of that free)"""
    return t if free is None or free.exact is None else t.furnish_with_freqs(free.exact)


def aop(dim, free=None):
    """a = sum_n sqrt(n+1) |n><n+1| (mode::aop, Mode.cc:277-282)"""
    return _with_freqs(Tridiagonal.unary(k=1, plus=[complex(math.sqrt(n + 1.0), 0.0) for n in range(dim - 1)]), free)


def _diag(dim, vals):
    return Tridiagonal((dim,), (0,), [np.array([complex(v) for v in vals]), None, None])


def _mode_h_over_i(z, eta, dim, free=None):
    """-z n + eta a^dagger - eta^* a (Mode.cc:56-86); None when both vanish"""
    t = None
    if _nonzero(z):
        t = _diag(dim, [complex(float(n), 0.0) * (-z) for n in range(dim)])
    if _nonzero(eta):
        ad = aop(dim, free).dagger().scaled(eta)
        p = ad + ad.dagger().scaled(-1.0)
        t = p if t is None else _with_k(t, 1) + p
    return t


def _with_k(t, k):
    return Tridiagonal(t.dims, (k,), t.diags, t.freqs)


def _exact_free(dim, exd):
    return Free(dim, exact=exd)


def mode(cutoff, delta=-10.0, kappa=0.0, eta=0.0, picture="Sch"):
    """mode::make(ParsPumpedLossy, picture) for nTh = 0 (Mode.tcc:40-60, 160-197)"""
    dim, eta = int(cutoff), complex(eta)
    z = complex(kappa, -delta)
    exd = None
    if picture == "Sch":
        h = _mode_h_over_i(z, eta, dim)
        if kappa and _nonzero(eta):
            hf = abs(z.conjugate() * float(dim))
        elif kappa:
            hf = max(abs(kappa * dim), abs(delta * dim))
        else:
            hf = abs(delta * dim)
        if _nonzero(eta):
            hf = max(hf, abs(eta * math.sqrt(float(dim))))
    else:
        if not kappa or not _nonzero(eta):
            raise ValueError("interaction-picture mode needs kappa and eta")
        zI = z if picture == "IP" else complex(0.0, -delta)
        exd = [zI * float(n) for n in range(dim)]
        ex = _exact_free(dim, exd)
        if picture == "IP":
            h, hf = _mode_h_over_i(0j, eta, dim, ex), abs(z.conjugate())
        else:
            h, hf = _mode_h_over_i(complex(kappa, 0.0), eta, dim, ex), max(abs(kappa * dim), abs(delta))
        hf = max(hf, abs(eta * math.sqrt(float(dim))))
    jumps = []
    if kappa:  # sqrt(2 kappa) a (Mode.cc:118-125)
        fact = math.sqrt(2.0 * kappa)
        jumps.append([Tridiagonal.unary(k=1, plus=[complex(fact * math.sqrt(n + 1.0), 0.0) for n in range(dim - 1)])])
    a = aop(dim)
    avr = [[_diag(dim, [float(n) for n in range(dim)])], [_diag(dim, [float(n) * n for n in range(dim)])], [a], [a]]
    return Free(dim, [h] if h is not None else [], jumps, avr, [0, 0, 0, 1], [-1, 0, -1, -1], [0, 0, 0, 0], exd, hf)


def _qbit_avr():
    sig = Tridiagonal.unary(k=1, plus=[1.0])
    return [[_diag(2, [1.0, 0.0])], [_diag(2, [0.0, 1.0])], [sig], [sig]]


def qbit(deltaA=-10.0, gamma=0.0, eta=0.0, picture="Sch"):
    """qbit::make(ParsPumpedLossy, picture): a two-level mode pumped with -eta (Qbit.cc:217-246)"""
    eta = complex(eta)
    z = complex(gamma, -deltaA)
    exd = None
    if picture == "Sch":
        h = _mode_h_over_i(z, -eta, 2)
        if gamma and _nonzero(eta):
            hf = max(abs(z.conjugate()), abs(eta))
        elif gamma:
            hf = max(abs(gamma), abs(deltaA))
        else:
            hf = max(abs(deltaA), abs(eta))
    else:
        if not gamma or not _nonzero(eta):
            raise ValueError("interaction-picture qbit needs gamma and eta")
        zI = z if picture == "IP" else complex(0.0, -deltaA)
        exd = [0j, zI]
        ex = _exact_free(2, exd)
        if picture == "IP":
            h, hf = _mode_h_over_i(0j, -eta, 2, ex), abs(z.conjugate())
        else:
            h, hf = _mode_h_over_i(complex(gamma, 0.0), -eta, 2, ex), max(abs(gamma), abs(deltaA))
        hf = max(hf, abs(eta))
    jumps = []
    if gamma:
        jumps.append([Tridiagonal.unary(k=1, plus=[complex(math.sqrt(2.0 * gamma) * math.sqrt(1.0), 0.0)])])
    return Free(2, [h] if h is not None else [], jumps, _qbit_avr(), [0, 0, 0, 1], [-1, -1, -1, -1], [0, 0, 0, 0],
                exd, hf)


def pumped_two_level_atom_sch(gamma=10.0, deltaA=-10.0, eta=0.0):
    """PumpedTwoLevelAtomSch: H/i = [[0, eta^*], [-eta, za]], za = (-gamma, deltaA) (PumpedTwoLevelAtom.cc:13-49)"""
    eta = complex(eta)
    za = complex(-gamma, deltaA)
    h = Tridiagonal((2,), (1,), [np.array([0j, za]), np.array([-eta]), np.array([eta.conjugate()])])
    jump = Tridiagonal.unary(k=1, plus=[complex(math.sqrt(-2.0 * za.real), 0.0)])
    hf = max(abs(complex(gamma, deltaA)), abs(eta))
    return Free(2, [h], [[jump]], _qbit_avr(), [0, 0, 0, 1], [-1, -1, -1, -1], [0, 0, 0, 0], None, hf)


def exp_inkx(dim, K, free=None):
    """e^{iKx} on the momentum grid: sum_q |q+K><q| (expINKX, Particle.cc:60-72)"""
    a = abs(K)
    ones = [1.0 + 0j] * max(0, dim - a)
    t = Tridiagonal.unary(k=a, minus=ones) if K >= 0 else Tridiagonal.unary(k=a, plus=ones)
    return _with_freqs(t, free)


def mode_function(dim, mf, nK, free=None):
    """mfNKX (Particle.cc:274-289)"""
    if mf == "sin":
        return (exp_inkx(dim, nK, free) + exp_inkx(dim, -nK, free).scaled(-1.0)).scaled(1.0 / complex(0.0, 2.0))
    if mf == "cos":
        return (exp_inkx(dim, nK, free) + exp_inkx(dim, -nK, free)).scaled(0.5)
    return exp_inkx(dim, nK if mf == "plus" else -nK, free)


def _identity(dim):
    return _diag(dim, [1.0] * dim)


def mf_composition(dim, a, b, nK, free=None):
    """mfComposition(particle, (a, K), (b, K)) for positive wavenumbers (Particle.cc:290-345)"""
    def e(K):
        return exp_inkx(dim, K, free)

    def idk(K):
        return _with_k(_identity(dim), K)
    K2 = 2 * nK
    cos2 = (e(K2) + e(-K2)).scaled(0.5)
    sin2 = (e(K2) + e(-K2).scaled(-1.0)).scaled(1.0 / complex(0.0, 2.0))
    half_i = complex(0.0, 0.5)
    if a == "sin" and b == "sin":
        return (idk(K2) + cos2.scaled(-1.0)).scaled(0.5)
    if a == "cos" and b == "cos":
        return (idk(K2) + cos2).scaled(0.5)
    if {a, b} == {"sin", "cos"}:
        return sin2.scaled(0.5)
    if a == b:
        return _identity(dim)
    if (a, b) == ("plus", "minus"):
        return e(-K2)
    if (a, b) == ("minus", "plus"):
        return e(K2)
    if (a, b) in (("plus", "sin"), ("sin", "minus")):
        return (e(-K2) + idk(K2).scaled(-1.0)).scaled(half_i)
    if (a, b) in (("plus", "cos"), ("cos", "minus")):
        return (idk(K2) + e(-K2)).scaled(0.5)
    if (a, b) in (("minus", "cos"), ("cos", "plus")):
        return (idk(K2) + e(K2)).scaled(0.5)
    return (e(K2) + idk(K2).scaled(-1.0)).scaled(-half_i)  # minus-sin, sin-plus


def _cpp_mod(x, m):
    """C's truncating remainder"""
    r = abs(x) % m
    return r if x >= 0 else -r


def particle(fin=6, omrec=1.0, vclass=0.0, kpart=1, mf="sin", picture="Sch"):
    """(Pumped)Particle(Sch) in the momentum basis (Particle.cc:79-110, 138-167, 210-252)"""
    if picture == "UIP":
        raise ValueError("the particle has no unitary interaction picture of its own")
    dim = 1 << int(fin)
    pumped = vclass != 0.0
    exd = None if picture == "Sch" else [complex(0.0, omrec * (n - dim // 2) * (n - dim // 2)) for n in range(dim)]
    ex = None if exd is None else _exact_free(dim, exd)
    pump = None
    if pumped and mf in ("sin", "cos") and 2 * kpart < dim:
        c = (exp_inkx(dim, 2 * kpart, ex) + exp_inkx(dim, -2 * kpart, ex)).scaled(0.5)
        pump = c.scaled(complex(vclass * (-1.0 if mf == "sin" else 1.0), 0.0) / complex(0.0, 2.0))
    if picture == "Sch":
        kin = _diag(dim, [complex(0.0, -omrec * (n - dim // 2) * (n - dim // 2)) for n in range(dim)])
        h = [kin if pump is None else _with_k(kin, 2 * kpart) + pump]
        hf = abs(omrec) * float(dim) * dim
    else:
        h = [] if pump is None else [pump]
        hf = abs(omrec) * float(dim)
    if pumped:
        hf = max(hf, abs(vclass))
    # averages: <P>, <P^2>, <X>, <X^2> with X = F^dagger x F (LazyDensityOperatorFFT.cc:16-40)
    pi = 3.14159265358979323846
    dx, xmax = 2.0 * pi / dim, pi
    byd = {}
    for dlt in range(-(dim - 1), dim):
        a = b = 0j
        for j in range(dim):
            x = j * dx - xmax
            ph = _cpp_mod((j + dim // 2) * dlt, dim)
            e = cmath.rect(1.0 / dim, 2.0 * pi * float(ph) / dim)
            a += x * e
            b += x * x * e
        byd[dlt] = (a, b)
    rows = [n for n in range(dim) for m in range(dim)]
    cols = [m for n in range(dim) for m in range(dim)]
    X = Sparse((dim,), rows, cols, [byd[m - n][0] for n, m in zip(rows, cols)])
    X2 = Sparse((dim,), rows, cols, [byd[m - n][1] for n, m in zip(rows, cols)])
    k = [float(n - dim // 2) for n in range(dim)]
    avr = [[_diag(dim, k)], [_diag(dim, [v * v for v in k])], [X], [X2]]
    f = Free(dim, h, [], avr, [0, 0, 0, 0], [-1, 0, -1, 2], [0, 0, 0, 1], exd, hf)
    f.pumped, f.vclass, f.mf, f.kpart = pumped, vclass, mf, kpart
    return f


# ---------------------------------------------------------------- interactions

def jaynes_cummings(q, m, iq, im, g):
    """conj(g) sigma (x) a^dagger - h.c. (JaynesCummings_.h:54-66); q, m: the frees, iq, im their axes"""
    g = complex(g)
    hf = abs(g * math.sqrt(float(q.dim) * m.dim))
    if not _nonzero(g):
        return Interaction((iq, im), [], hf)
    t = aop(q.dim, q).scaled(g.conjugate()) * aop(m.dim, m).dagger()
    return Interaction((iq, im), [t + t.dagger().scaled(-1.0)], hf)


def coupled_modes(m1, m2, i1, i2, u):
    """CoupledModes<CM_XX_RWA>: conj(u) a_1 a_2^dagger - h.c. (CoupledModes.cc:36-41)"""
    u = complex(u)
    t = aop(m1.dim, m1).scaled(u.conjugate()) * aop(m2.dim, m2).dagger()
    return Interaction((i1, i2), [t + t.dagger().scaled(-1.0)], abs(u * (m1.dim * math.sqrt(float(m2.dim)))))


def zz(a, b, ia, ib, J):
    """J sigma_z (x) sigma_z, sigma_z = diag(-1, +1); H/i = -i J sz sz"""
    sz = _diag(2, [-1.0, 1.0])
    return Interaction((ia, ib), [(sz * sz).scaled(complex(0.0, -J))], abs(J))


def _etaeff(u0, v):
    return (1.0 if u0 > 0 else -1.0 if u0 < 0 else 0.0) * math.sqrt(u0 * v)


def particle_orthogonal_to_cavity(m, p, im, ip, u0=-1.0):
    """etaeff (a (x) mf(kPart x) + h.c.) / i (ParticleCavity.cc:47-57, 79-85)"""
    if u0 * p.vclass < 0.0:
        raise ValueError("Unot VClass sign discrepancy")
    ee = _etaeff(u0, p.vclass)
    a = aop(m.dim, m) * mode_function(p.dim, p.mf, p.kpart, p)
    hf = max(abs(u0) * m.dim, abs(ee) * math.sqrt(float(m.dim)))
    return Interaction((im, ip), [(a + a.dagger()).scaled(complex(ee, 0.0) / complex(0.0, 1.0))], hf)


def particle_along_cavity(m, p, im, ip, u0=0.0, kcav=1, mf="cos", vclass=None):
    """ParticleAlongCavity (ParticleCavity.cc:26-53, 87-110): the dispersive term and the
    interference term, each a Tridiagonal of its own (the TridiagonalHamiltonian's list)"""
    pumped_ctor = p.pumped and vclass is None
    v = p.vclass if pumped_ctor else (vclass or 0.0)
    if u0 * v < 0.0:
        raise ValueError("Unot VClass sign discrepancy")
    ee = _etaeff(u0, v)
    sgn = -1.0 if mf == "sin" else 1.0
    K = 2 * kcav
    h = []
    if u0 != 0.0 and K < p.dim and mf in ("cos", "sin"):
        n = _diag(m.dim, [float(a) for a in range(m.dim)])
        c = Tridiagonal.unary(k=K, minus=[0.5] * (p.dim - K), plus=[0.5] * (p.dim - K))
        c = c if p.exact is None else c.furnish_with_freqs(p.exact)
        h.append((n * c).scaled(complex(u0 * sgn, 0.0) / complex(0.0, 2.0)))
    special = not pumped_ctor or kcav == abs(p.kpart)
    if ee != 0.0 and special:
        if pumped_ctor:
            x = aop(m.dim, m).dagger() * mf_composition(p.dim, mf, p.mf, kcav, p)
        else:
            x = aop(m.dim, m) * mode_function(p.dim, mf, kcav, p)
        h.append((x + x.dagger()).scaled(complex(ee, 0.0) / complex(0.0, 1.0)))
    if ee != 0.0 and not special:
        h.append(_chained_pcav(m, p, ee, mf, kcav))
    return Interaction((im, ip), h, max(abs(u0) * m.dim, abs(ee) * math.sqrt(float(m.dim))))


def _mf_matrix(dim, mf, nK):
    """mfNKX on the truncated momentum grid as a dense matrix (e^{iKx}: |q> -> |q+K>)"""
    e = np.zeros((dim, dim), complex)
    for q in range(dim - abs(nK)):
        if nK >= 0:
            e[q + nK, q] = 1.0
        else:
            e[q, q - nK] = 1.0
    if mf == "sin":
        return (e + e.conj().T * -1.0) * (1.0 / complex(0.0, 2.0))
    if mf == "cos":
        return (e + e.conj().T) * 0.5
    return e if mf == "plus" else e.conj().T


def _chained_pcav(m, p, ee, mf, kcav):
    """ParticleAlongCavity with a pumped particle of another wavenumber (ParticleCavity.cc:102-140):
    secondH (firstH psi) + secondHT (firstHT psi), firstH = ee aop(mode) mfNKX(particle, (mf, kcav)) / i,
    secondH = mfNKX(particle, particle's mf)^dagger, i.e. ee (a (x) mfPart^dagger mf + a^dagger (x)
    mfPart mf^dagger) / i: a product of Tridiagonals, handed over as a sparse term.  Its entries
    shift the particle by +-kPart +- kcav (never 0), so with both frees Exact each rotates with
    exact_r - exact_c over both axes; with only the particle Exact it cannot be one entry."""
    if p.exact is not None and m.exact is None:
        raise ValueError("the chained particle-cavity term needs the mode in the interaction picture when the particle is")
    mc, mp = _mf_matrix(p.dim, mf, kcav), _mf_matrix(p.dim, p.mf, p.kpart)
    f = complex(ee, 0.0) / complex(0.0, 1.0)
    rows, cols, vals, freqs = [], [], [], []
    both = p.exact is not None and m.exact is not None
    for P, up in ((mp.conj().T @ mc, False), (mp @ mc.conj().T, True)):
        for n in range(m.dim - 1):
            rm, cm = (n + 1, n) if up else (n, n + 1)    # a: |n><n+1|, a^dagger: |n+1><n|
            av = complex(math.sqrt(n + 1.0), 0.0)
            for r, c in zip(*np.nonzero(P)):
                rows.append(rm * p.dim + r)
                cols.append(cm * p.dim + c)
                vals.append((av * P[r, c]) * f)
                freqs.append(((0.0 + (m.exact[rm] - m.exact[cm])) + (p.exact[r] - p.exact[c])) if both else 0.0)
    return Sparse((m.dim, p.dim), rows, cols, vals, freqs if both else None)


def particle_two_modes(m0, m1, p, i0, i1, ip, u0=-1.0, u1=-1.0, phi=0.0, k=1, mf0="plus", mf1="plus"):
    """ParticleTwoModes with equal wavenumbers (ParticleTwoModes.cc:35-46)"""
    if u0 * u1 < 0:
        raise ValueError("Unots sign discrepancy")
    fac = (1.0 if u0 > 0 else -1.0 if u0 < 0 else 0.0) * math.sqrt(u0 * u1) * cmath.rect(1.0, phi)
    x = ((aop(m0.dim, m0).dagger() * aop(m1.dim, m1)) * mf_composition(p.dim, mf0, mf1, k, p)).scaled(fac)
    hf = max(abs(u0) * math.sqrt(float(m0.dim)), abs(u1) * math.sqrt(float(m1.dim)))
    return Interaction((i0, i1, ip), [(x + x.dagger()).scaled(1.0 / complex(0.0, 1.0))], hf)


# ---------------------------------------------------------------- systems

class _Tri(C.Structure):
    _fields_ = [("rank", C.c_int), ("axes", C.c_void_p), ("k", C.c_void_p), ("diags", C.c_void_p),
                ("freqs", C.c_void_p)]


class _Coo(C.Structure):
    _fields_ = [("rank", C.c_int), ("axes", C.c_void_p), ("nnz", C.c_longlong), ("rows", C.c_void_p),
                ("cols", C.c_void_p), ("vals", C.c_void_p), ("freqs", C.c_void_p)]


class _Op(C.Structure):
    _fields_ = [("n_tri", C.c_int), ("tri", C.c_void_p), ("n_coo", C.c_int), ("coo", C.c_void_p)]


class _FreeDesc(C.Structure):
    _fields_ = [("dim", C.c_int), ("exact", C.c_void_p)]


class _SysDesc(C.Structure):
    _fields_ = [("n_frees", C.c_int), ("frees", C.c_void_p), ("h", _Op), ("n_jumps", C.c_int), ("jumps", C.c_void_p),
                ("n_avr", C.c_int), ("avr", C.c_void_p), ("avr_imag", C.c_void_p), ("avr_var_of", C.c_void_p),
                ("avr_sqrt", C.c_void_p), ("highest_frequency", C.c_double)]


def build_system(frees, interactions=(), binary=None):
    """A device system from frees and interactions (qt_sys_build_terms): H/i = the frees' terms,
    then the interactions' in order; jumps and averages of the frees in order; the highest
    frequency is the frees' (and the interaction's for a binary system, DynamicsBase.cc:49-66).
    binary: None = two frees with at most one interaction."""
    keep = []

    def arr(a, dt):
        a = np.ascontiguousarray(a, dt)
        keep.append(a)
        return a.ctypes.data

    def op(terms, axes):
        tris, coos = [], []
        for t in terms:
            if isinstance(t, Sparse):
                coos.append(_Coo(len(axes), arr(axes, np.int32), len(t.vals), arr(t.rows, np.int64),
                                 arr(t.cols, np.int64), arr(t.vals.view(np.float64), np.float64),
                                 None if t.freqs is None else arr(t.freqs.view(np.float64), np.float64)))
                continue
            if len(t.dims) != len(axes):
                raise ValueError("term rank does not match its axes")
            n = 3 ** t.rank
            dp = (C.c_void_p * n)(*[None if d is None else arr(d.ravel().view(np.float64), np.float64) for d in t.diags])
            fp = (C.c_void_p * n)(*[None if f is None else arr(f.ravel().view(np.float64), np.float64) for f in t.freqs])
            keep.extend([dp, fp])
            tris.append(_Tri(t.rank, arr(axes, np.int32), arr(t.k, np.int32), C.cast(dp, C.c_void_p).value,
                             C.cast(fp, C.c_void_p).value))
        ta = (_Tri * max(1, len(tris)))(*tris)
        ca = (_Coo * max(1, len(coos)))(*coos)
        keep.extend([ta, ca])
        return _Op(len(tris), C.cast(ta, C.c_void_p).value, len(coos), C.cast(ca, C.c_void_p).value)

    fd = (_FreeDesc * len(frees))(*[_FreeDesc(f.dim, None if f.exact is None else arr(f.exact.view(np.float64), np.float64))
                                    for f in frees])
    h_terms, h_axes = [], []
    for a, f in enumerate(frees):
        for t in f.h:
            h_terms.append(t)
            h_axes.append((a,))
    for ia in interactions:
        for t in ia.h:
            h_terms.append(t)
            h_axes.append(ia.axes)
    # one operator holding every H/i term: the Tridiagonals keep their own axes
    tris, coos = [], []
    for t, ax in zip(h_terms, h_axes):
        o = op([t], ax)
        if o.n_tri:
            tris.append(C.cast(o.tri, C.POINTER(_Tri))[0])
        else:
            coos.append(C.cast(o.coo, C.POINTER(_Coo))[0])
    ht = (_Tri * max(1, len(tris)))(*tris)
    hc = (_Coo * max(1, len(coos)))(*coos)
    keep.extend([ht, hc])
    h = _Op(len(tris), C.cast(ht, C.c_void_p).value, len(coos), C.cast(hc, C.c_void_p).value)
    jumps, avr, imag, var_of, sq = [], [], [], [], []
    for a, f in enumerate(frees):
        for j in f.jumps:
            jumps.append(op(j, (a,)))
        base = len(avr)
        for q, o in enumerate(f.avr):
            avr.append(op(o, (a,)))
            imag.append(f.imag[q])
            var_of.append(-1 if f.var_of[q] < 0 else base + f.var_of[q])
            sq.append(f.sqrt[q])
    ja = (_Op * max(1, len(jumps)))(*jumps)
    aa = (_Op * max(1, len(avr)))(*avr)
    keep.extend([ja, aa, fd])
    if binary is None:
        binary = len(frees) == 2 and len(interactions) <= 1
    hf = max(f.hf for f in frees)
    if binary:
        hf = max([hf] + [ia.hf for ia in interactions])
    desc = _SysDesc(len(frees), C.cast(fd, C.c_void_p).value, h, len(jumps), C.cast(ja, C.c_void_p).value,
                    len(avr), C.cast(aa, C.c_void_p).value, arr(imag, np.int32), arr(var_of, np.int32),
                    arr(sq, np.int32), hf)
    l = lib()
    l.qt_sys_build_terms.argtypes = [C.POINTER(_SysDesc), C.POINTER(C.c_void_p)]
    handle = C.c_void_p()
    check(l.qt_sys_build_terms(C.byref(desc), C.byref(handle)))
    return System.adopt(handle)
