"""Pins of the oracle's kernels beside the Coulomb family (regularized
Yukawa, atan, sin-over-r, MQ, RBS-u, RBS-v and the reference's user-kernel
example), CPU only.

- Bit-equal to the reference's own compiled kernel sources (oracle/_ref) in
  every combination the reference computes as written, including CC with
  cluster-cluster lists.
- The treecode converges to the direct sum as the degree grows, and at a
  tiny theta it is the direct sum (all pairs near).
- Closed forms of the two 2-D RBS kernels and the periodic 1-D kernels on
  hand-placed particles (the direct sum against formulas from the kernels'
  definitions), and the rejected combinations.
"""
import math

import numpy as np
import pytest

from tests.helpers import KP_OTHER, other_combos, rel_l2, uniform_cube


@pytest.mark.parametrize("ct,k,s", other_combos())
def test_other_kernels_bit_equal_to_reference_sources(oracle_mod, ct, k, s):
    if not oracle_mod.ref_available():
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    t, src = uniform_cube(2500, 2000, seed=40 + ct * 7 + k)
    args = (k, KP_OTHER[k], s, 1, ct, 0.7, 4, 60, 50)
    assert np.array_equal(oracle_mod.treecode(*t, *src, *args), oracle_mod.ref_treecode(*t, *src, *args))


@pytest.mark.parametrize("k", [4, 8, 10, 11, 12])
def test_other_kernels_cc_lists_bit_equal(oracle_mod, k):
    if not oracle_mod.ref_available():
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    t, src = uniform_cube(8000, 8000, seed=9)
    args = (k, KP_OTHER[k], 1, 1, 3, 0.8, 2, 30, 30)
    a = oracle_mod.treecode(*t, *src, *args)
    st = oracle_mod.last_stats()
    assert st["num_cc"] > 0 and st["num_pc"] > 0 and st["num_cp"] > 0
    assert np.array_equal(a, oracle_mod.ref_treecode(*t, *src, *args))


@pytest.mark.parametrize("ct,k,s", other_combos())
def test_other_kernels_converge(oracle_mod, ct, k, s):
    t, src = uniform_cube(2000, 2000, seed=90 + ct + k)
    direct = oracle_mod.direct(*t, *src, k, KP_OTHER[k], s)
    errs = [rel_l2(direct, oracle_mod.treecode(*t, *src, k, KP_OTHER[k], s, 1, ct, 0.7, d, 50, 50))
            for d in (2, 5, 8)]
    assert errs[2] < 0.05 * errs[0], errs
    tiny = oracle_mod.treecode(*t, *src, k, KP_OTHER[k], s, 1, ct, 1e-6, 3, 50, 50)
    assert rel_l2(direct, tiny) < 1e-13


def test_other_kernels_closed_forms(oracle_mod):
    # one source of charge q at the origin, targets placed by hand
    tx = np.array([0.3, -0.2, 0.05]); ty = np.array([0.1, 0.4, -0.3]); tz = np.array([0.2, -0.45, 0.7])
    sx = np.zeros(1); sy = np.zeros(1); sz = np.zeros(1); q = np.array([0.7]); w = np.ones(1)
    tq = np.zeros(3)
    d = 0.1
    # RBS u, v: 2-D regularized Biot-Savart, u = q y / (2 pi (x^2 + y^2 + d^2)), v = -q x / (...)
    u = oracle_mod.direct(tx, ty, tz, tq, sx, sy, sz, q, w, 10, [d], 1)
    v = oracle_mod.direct(tx, ty, tz, tq, sx, sy, sz, q, w, 11, [d], 1)
    den = 2 * math.pi * (tx ** 2 + ty ** 2 + d ** 2)
    assert np.allclose(u, 0.7 * ty / den, rtol=1e-14) and np.allclose(v, -0.7 * tx / den, rtol=1e-14)
    # sin-over-r and the user example exp(-k r)/r
    r = np.sqrt(tx ** 2 + ty ** 2 + tz ** 2)
    so = oracle_mod.direct(tx, ty, tz, tq, sx, sy, sz, q, w, 8, [1.5], 1)
    us = oracle_mod.direct(tx, ty, tz, tq, sx, sy, sz, q, w, 12, [0.5], 1)
    assert np.allclose(so, 0.7 * np.sin(1.5 * r) / r, rtol=1e-14)
    assert np.allclose(us, 0.7 * np.exp(-0.5 * r) / r, rtol=1e-14)
    # MQ, period L: odd in the wrapped offset, zero at offset 0 and 1/2
    L, dl = 2.0, 0.3
    zt = np.array([0.0, 1.0, 0.37, -0.37])
    mq = oracle_mod.direct(np.zeros(4), np.zeros(4), zt, np.zeros(4), sx, sy, sz, q, w, 9, [L, dl], 1)
    assert abs(mq[0]) == 0.0 and abs(mq[1]) < 1e-15 and abs(mq[2] + mq[3]) < 1e-15
    # ATAN: the same symmetry, scaled by wadj
    at = oracle_mod.direct(np.zeros(4), np.zeros(4), zt, np.zeros(4), sx, sy, sz, q, w, 5, [L, dl], 1)
    assert abs(at[0]) == 0.0 and abs(at[2] + at[3]) < 1e-15
    dz = 0.37 / L
    f = math.sqrt(1 + 1 / dl ** 2)
    wadj = 1 / (1 - dl / math.sqrt(1 + dl * dl))
    assert abs(at[2] - 0.7 * wadj * (math.atan(f * math.tan(math.pi * dz)) / math.pi - dz)) < 1e-15


def test_other_kernels_unsupported(oracle_mod):
    t, src = uniform_cube(300, 300, seed=1)
    for ct, k, s, apx in [(2, 5, 1, 1), (3, 9, 1, 1), (2, 4, 2, 1), (1, 8, 2, 1), (3, 8, 1, 2), (1, 4, 1, 2)]:
        with pytest.raises(RuntimeError):
            oracle_mod.treecode(*t, *src, k, KP_OTHER[k], s, apx, ct, 0.7, 3, 50, 50)
