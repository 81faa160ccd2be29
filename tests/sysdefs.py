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


def jc_master_spec(cutoff=40, eta=1.0, g=10.0):
    return (f"free qbit deltaA=0 gamma=0 eta=0\n"
            f"free mode cutoff={cutoff} delta=0 kappa=1 eta={eta!r}\n"
            f"ia jc 0 1 g={g!r}")


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
