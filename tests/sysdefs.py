"""Element specs and initial states of the systems the tests and the benchmark use.

The spec grammar is shared by the product (qt_sys_build) and the oracle.
"""
import numpy as np

# PumpedTwoLevelAtomSch of the reference's PTLA_C++QED golden runs (gamma=10, deltaA=-10, etat=8)
PTLA = "free ptla gamma=10 deltaA=-10 eta=8"

# PumpedLossyMode, cutoff 20, Delta=-1, eta=2, kappa=1 (Schroedinger picture)
PLM20 = "free mode cutoff=20 delta=-1 kappa=1 eta=2"

# driven Jaynes-Cummings, qubit (x) mode cutoff 30: g=3, deltaC=-1, deltaA=0.5, eta=1, kappa=1, gamma=0.2
JC60 = ("free qbit deltaA=0.5 gamma=0.2 eta=0\n"
        "free mode cutoff=30 delta=-1 kappa=1 eta=1\n"
        "ia jc 0 1 g=3")


# Mixed batch of composite systems (all FP32 derivatives).  Parameters the workload leaves open
# (losses, pumps, detunings, initial states) are fixed here.
# (a) mode cutoff 16 (x) particle with 64 momentum states, cos^2 coupling U0 = -1: the dispersive
#     term U0 n cos(2kx)/2 of ParticleAlongCavity plus U0 n / 2 folded into the mode detuning
#     (deltaC = -1 -> -1 - U0/2, as CPPQEDscripts/1particle1mode.cc does), recoil frequency 0.1
PCAV1024 = ("free mode cutoff=16 delta=-0.5 kappa=1 eta=1\n"
            "free particle fin=6 omrec=0.1\n"
            "ia pcav 0 1 u0=-1 kcav=1 mf=cos")
# (b) four modes, cutoff 8 each (dim 4096), RWA beam splitters g = 0.5 between neighbours,
#     kappa = 1 and deltaC = -1 on every mode, the first one pumped with eta = 1
MODES4096 = "\n".join([f"free mode cutoff=8 delta=-1 kappa=1 eta={1 if i == 0 else 0}" for i in range(4)]
                      + [f"ia bs {i} {i + 1} g=0.5" for i in range(3)] + ["system composite"])


def qubits64_spec(seed=20240611):
    """(c) six qubits (dim 64): ZZ couplings J_ij ~ U(-1, 1) on all 15 pairs (numpy PCG64 seed),
    transverse drive eta = 0.7 and decay gamma = 0.2 on every qubit, deltaA = 0."""
    rng = np.random.default_rng(seed)
    lines = ["free qbit deltaA=0 gamma=0.2 eta=0.7"] * 6
    for i in range(6):
        for j in range(i + 1, 6):
            lines.append(f"ia zz {i} {j} J={rng.uniform(-1, 1)!r}")
    return "\n".join(lines + ["system composite"])


QUBITS64 = qubits64_spec()

# the four builds with their initial basis states (flat index): vacuum (x) particle at k = 0, vacuum,
# all qubits in the ground state, ground (x) vacuum
MIXED = [("pcav1024", PCAV1024, 32), ("modes4096", MODES4096, 0), ("qubits64", QUBITS64, 0), ("jc60", JC60, 0)]


def mixed_counts(n_total=8192, seed=7):
    """Trajectories per build: each of n_total drawn uniformly from the four builds."""
    rng = np.random.default_rng(seed)
    return np.bincount(rng.integers(0, len(MIXED), n_total), minlength=len(MIXED))


def basis(dim, index):
    psi = np.zeros(dim, np.complex128)
    psi[index] = 1
    return psi


def jc_master_spec(cutoff=40, eta=1.0, g=10.0):
    return (f"free qbit deltaA=0 gamma=0 eta=0\n"
            f"free mode cutoff={cutoff} delta=0 kappa=1 eta={eta!r}\n"
            f"ia jc 0 1 g={g!r}")


# CPPQEDscripts/1particle1mode, configuration 1 (pumped particle moving orthogonal to the pumped
# cavity), with the golden runs' options --cutoff 3 --fin 3 --eta 10 and the parameter defaults
# (deltaC = -10, kappa = 10, omrec = 1, vClass = -10, kPart = 1, Sin mode function, Unot = -1),
# interaction picture; the Master runs switch the lossy mode to its unitary interaction picture.
P1M1 = ("free mode cutoff=3 delta=-10 kappa=10 eta=10 picture=IP\n"
        "free particle fin=3 omrec=1 vclass=-10 kpart=1 mf=sin picture=IP\n"
        "ia pcav_orth 0 1 u0=-1")
P1M1_MASTER = P1M1.replace("eta=10 picture=IP", "eta=10 picture=UIP")

# 1particle1mode --1p1mconf 2, 3, 4 (CPPQEDscripts/1particle1mode.cc:36-68): ParticleAlongCavity;
# the cavity detuning is shifted by -Unot/2 = +0.5
_P1M1_MODE = "free mode cutoff=3 delta=-9.5 kappa=10 eta=10 picture=IP\n"
P1M1_CONF = {
    2: _P1M1_MODE + "free particle fin=3 omrec=1 picture=IP\nia pcav 0 1 u0=-1 kcav=1 mf=sin vclass=-10",
    3: _P1M1_MODE + "free particle fin=3 omrec=1 vclass=-10 kpart=1 mf=sin picture=IP\nia pcav 0 1 u0=-1 kcav=1 mf=sin",
    4: _P1M1_MODE + "free particle fin=3 omrec=1 vclass=-10 kpart=1 mf=sin picture=IP\n"
                    "ia pcav 0 1 u0=-1 kcav=1 mf=sin vclass=0",
}


def particle_wavepacket(fin, x0=-0.5, k0=0.0, sig=0.1):
    """particle::wavePacket in x (Particle.cc:347-369), exp(-(x - pi x0)^2 / (4 sig^2) + i k0 x) on
    x_j = j 2pi/N - pi, taken to the momentum basis by quantumdata::ffTransform(DIR_XK)
    (LazyDensityOperatorFFT.cc:16-40: GSL forward transform, halves swapped with N^-1/2, odd
    entries negated), normalised.  The defaults are the reference's pinit (ParsParticle.cc:14)."""
    n = 1 << fin
    x = np.arange(n) * (2 * np.pi / n) - np.pi
    psi = np.exp(-(x - np.pi * x0) ** 2 / (4 * sig ** 2) + 1j * x * k0)
    y = np.fft.fft(psi)
    out = np.concatenate([y[n // 2:], y[:n // 2]]) / np.sqrt(n)
    out[1::2] *= -1
    return out / np.linalg.norm(out)


def p1m1_initial():
    """mode vacuum (mode::init with minit = 0, Mode.cc:336-339) (x) the particle's wave packet"""
    return np.kron(basis(3, 0), particle_wavepacket(3))


# CPPQEDscripts/Ring with the golden runs' options (--etaM 10 --etaP 10 --cutoffM 3 --cutoffP 3
# --fin 3): particle (x) mode "plus" (x) mode "minus", composite, interaction picture; each mode's
# detuning -10 is shifted by -Unot (complex mode functions, Ring.cc:22-23); the two
# ParticleAlongCavity acts have no Hamiltonian part (complex mode functions, unpumped particle);
# the three-body ParticleTwoModes act on (plus, minus, particle).
RING = ("free particle fin=3 omrec=1 picture=IP\n"
        "free mode cutoff=3 delta=-9 kappa=10 eta=10 picture=IP\n"
        "free mode cutoff=3 delta=-9 kappa=10 eta=10 picture=IP\n"
        "ia pcav 1 0 u0=-1 kcav=1 mf=plus\n"
        "ia pcav 2 0 u0=-1 kcav=1 mf=minus\n"
        "ia ptm 1 2 0 u0=-1 u1=-1 phi=0 k0=1 k1=1 mf0=plus mf1=minus\n"
        "system composite")
RING_MASTER = RING.replace("eta=10 picture=IP", "eta=10 picture=UIP")


def ring_initial():
    """init(pp) (x) init(pmP) (x) init(pmM): the particle's wave packet and two vacua"""
    return np.kron(np.kron(particle_wavepacket(3), basis(3, 0)), basis(3, 0))


def ground(dim):
    psi = np.zeros(dim, np.complex128)
    psi[0] = 1
    return psi


def coherent(alpha, dim):
    """mode::coherent (Mode.cc:321-324) with coherentElement (MathExtensions.cc:35-41), small n."""
    from math import factorial, sqrt
    n = np.arange(dim)
    amp = np.array([alpha ** k / sqrt(factorial(k)) for k in n], np.complex128)
    return np.exp(-abs(alpha) ** 2 / 2) * amp


# ParticleAlongCavity with a pumped particle of another wavenumber than the cavity's (the chained
# product term of ParticleCavity.cc:102-140): mode cutoff 3 (x) particle with 8 momentum states (in
# the interaction picture its entries take 31 distinct frequencies, within the device's 32)
PCAV_OTHER_K = ("free mode cutoff=3 delta=-1 kappa=1 eta=0.5\n"
                "free particle fin=3 omrec=0.7 vclass=-2 kpart=1 mf=sin\nia pcav 0 1 u0=-1.5 kcav=2 mf=cos")
PCAV_OTHER_K_IP = PCAV_OTHER_K.replace("eta=0.5", "eta=0.5 picture=IP").replace("mf=sin", "mf=sin picture=IP")
