"""Shared input generators and error metrics for the tests."""
import ctypes

import numpy as np

KP = {  # kernel -> singularity -> parameters
    1: {1: [0.5], 2: [0.5]},
    2: {1: [0.5], 2: [0.5]},
    3: {1: [0.05], 2: [0.3, 0.05]},
    8: {1: [1.5]},  # sin-over-r (Hermite cases only; skipping)
}


def supported_combos():
    """(compute_type, kernel, singularity) the reference accepts for Lagrange
    (src/interaction_compute/interaction_compute_{pc,cp,cc}.c)."""
    out = []
    for ct in (1, 2, 3):
        for k in (1, 2, 3):
            for s in (1, 2):
                if k == 3 and s == 2 and ct != 1:
                    continue
                out.append((ct, k, s))
    return out


def uniform_cube(n, m, seed, wlo=0.5, whi=1.5):
    """Sources (x,y,z,q,w) and targets (x,y,z,q), uniform in [-1,1]^3,
    charges uniform in [-1,1], weights uniform in [wlo, whi]."""
    rng = np.random.default_rng(seed)
    s = [rng.uniform(-1, 1, n) for _ in range(4)] + [rng.uniform(wlo, whi, n)]
    t = [rng.uniform(-1, 1, m) for _ in range(4)]
    return t, s


def kat_inputs():
    """Inputs of interfaces/python/testBaryTreeInterface.py:41-46: numpy legacy
    MT19937 seed(1), drawn in the order RHO, X, Y, Z; weights 1."""
    rs = np.random.RandomState(1)
    n = 5000
    rho = rs.rand(n); x = rs.rand(n); y = rs.rand(n); z = rs.rand(n)
    return x, y, z, rho, np.ones(n)


def rel_l2(ref, val):
    ref = np.asarray(ref); val = np.asarray(val)
    return float(np.sqrt(np.sum((ref - val) ** 2) / np.sum(ref ** 2)))


def hermite_combos():
    """(compute_type, kernel, singularity) the reference accepts for Hermite:
    PC with either singularity, CP with skipping, no CC, and sin-over-r in
    PC and CP with skipping
    (src/interaction_compute/interaction_compute_{pc,cp,cc}.c,
    interaction_compute_downpass.c:126-128)."""
    out = [(1, k, s) for k in (1, 2, 3) for s in (1, 2)]
    out += [(2, k, 1) for k in (1, 2, 3)]
    return out + [(1, 8, 1), (2, 8, 1)]


# the kernels beside the Coulomb family: kernel id -> parameters
KP_OTHER = {4: [0.5, 0.1], 5: [2.0, 0.3], 8: [1.5], 9: [2.0, 0.3], 10: [0.1], 11: [0.1], 12: [0.5]}


def other_combos():
    """(compute_type, kernel, singularity) for the kernels beside the Coulomb
    family, where the reference dispatchers compute them as written
    (src/interaction_compute/interaction_compute_{pc,cp,cc}.c): Lagrange;
    ATAN and MQ in PC only; subtraction for regularized Yukawa in PC only."""
    out = [(ct, k, 1) for k in (4, 8, 10, 11, 12) for ct in (1, 2, 3)]
    return out + [(1, 5, 1), (1, 9, 1), (1, 4, 2)]


def glibc_rand_particles(n, with_targets=False):
    """srand(1); rand()/RAND_MAX draws as in tests/serial_tests.c:580-588 and
    :1452-1465 (glibc's rand, so the same stream as the reference test)."""
    libc = ctypes.CDLL("libc.so.6")
    libc.rand.restype = ctypes.c_int
    libc.srand(1)
    rmax = 2147483647.0

    def r():
        return libc.rand() / rmax

    if with_targets:
        t = [np.zeros(n) for _ in range(4)]
        s = [np.zeros(n) for _ in range(5)]
        for i in range(n):
            for a in t:
                a[i] = r() * 2.0 - 1.0
            for a in s[:4]:
                a[i] = r() * 2.0 - 1.0
            s[4][i] = r()
        return t, s
    s = [np.zeros(n) for _ in range(5)]
    for i in range(n):
        for a in s[:4]:
            a[i] = r() * 2.0 - 1.0
        s[4][i] = r()
    return s


# The reference's treecode tests (tests/serial_tests.c): approximation
# (1 Lagrange, 2 Hermite) x kernel (1 Coulomb, 2 Yukawa) x singularity
# (1 skipping, 2 subtraction), kernel parameter 0.5.
SCENARIO_COMBOS = [(ap, k, s) for ap in (1, 2) for k in (1, 2) for s in (1, 2)]
# serial_tests.c:536-840: PC, theta 0.8, degree 4, leaves 100/100, size_check 0
ONE_TARGET_ARGS = (1, 0.8, 4, 100, 100, 0.0)
ONE_TARGET_TOL = {1: 2e-4, 2: 2e-5}
# serial_tests.c:965-1060: (theta, degree) of the three parameter sets
MONOTONE_SETS = [(0.9, 3), (0.9, 6), (0.4, 3)]


def one_target():
    """The single target of serial_tests.c:571-577: the origin, charge 1."""
    return [np.zeros(1), np.zeros(1), np.zeros(1), np.ones(1)]
