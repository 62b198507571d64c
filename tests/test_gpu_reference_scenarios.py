"""The reference's own treecode tests (tests/serial_tests.c) on the HIP path.

Each scenario runs through the drop-in entry point BaryTreeInterface with the
reference test's inputs (glibc srand(1)/rand() streams) and parameters, and
is checked twice: against the CPU oracle on the same inputs at 1e-12 (the
plain FP64 bar), and against the reference test's own tolerance or property
(the FP64 direct sum as the judge).
"""
import ctypes

import numpy as np
import pytest

from tests.helpers import (MONOTONE_SETS, ONE_TARGET_ARGS, ONE_TARGET_TOL, SCENARIO_COMBOS,
                           glibc_rand_particles, one_target, rel_l2)

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-12


@pytest.fixture(scope="module")
def bt():
    import barytree_amd
    barytree_amd.lib()
    return barytree_amd


def interface(bt, t, s, kernel, kp, sing, approx, ct, theta, degree, msl, mtl, size_check):
    """One BaryTreeInterface call (src/interface/BaryTreeInterface.h:52-59),
    size_check passed as given (0 included, as the reference tests set it)."""
    def p(a):
        return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    t = [np.ascontiguousarray(a, dtype=np.float64) for a in t]
    s = [np.ascontiguousarray(a, dtype=np.float64) for a in s]
    kpa = np.asarray(kp, dtype=np.float64)
    out = np.zeros(len(t[0]))
    bt.lib().BaryTreeInterface(
        ctypes.c_int(len(t[0])), ctypes.c_int(len(s[0])), *[p(a) for a in t], *[p(a) for a in s], p(out),
        ctypes.c_int(kernel), ctypes.c_int(len(kpa)), p(kpa), ctypes.c_int(sing), ctypes.c_int(approx),
        ctypes.c_int(ct), ctypes.c_double(theta), ctypes.c_int(degree), ctypes.c_int(msl),
        ctypes.c_int(mtl), ctypes.c_double(size_check), ctypes.c_double(-1.0), ctypes.c_int(0))
    return out


@pytest.mark.parametrize("ap,k,s", SCENARIO_COMBOS)
def test_treecode_on_1_target_10000_sources(bt, oracle_mod, ap, k, s):
    """serial_tests.c:536-840: relative error < 2e-4 (Lagrange), < 2e-5
    (Hermite) against the direct sum."""
    src = glibc_rand_particles(10000)
    t = one_target()
    got = interface(bt, t, src, k, [0.5], s, ap, *ONE_TARGET_ARGS)
    ref = oracle_mod.treecode(*t, *src, k, [0.5], s, ap, *ONE_TARGET_ARGS)
    d = oracle_mod.direct(*t, *src, k, [0.5], s)
    print(f"approx={ap} k={k} s={s}: gpu {got[0]:.15e} oracle {ref[0]:.15e} direct {d[0]:.15e}")
    assert abs(got[0] - ref[0]) <= TOL_FP64 * abs(ref[0])
    assert abs(got[0] - d[0]) / abs(d[0]) < ONE_TARGET_TOL[ap]


@pytest.mark.parametrize("ap,k,s", SCENARIO_COMBOS)
def test_treecode_parameters_on_1_target_5000_sources(bt, oracle_mod, ap, k, s):
    """serial_tests.c:965-1411: a higher degree and a smaller theta each
    reduce the error (leaves 50/5, size_check 0)."""
    src = glibc_rand_particles(5000)
    t = one_target()
    d = oracle_mod.direct(*t, *src, k, [0.5], s)[0]
    errs = []
    for th, deg in MONOTONE_SETS:
        args = (1, th, deg, 50, 5, 0.0)
        got = interface(bt, t, src, k, [0.5], s, ap, *args)[0]
        ref = oracle_mod.treecode(*t, *src, k, [0.5], s, ap, *args)[0]
        assert abs(got - ref) <= TOL_FP64 * abs(ref)
        errs.append(abs(got - d))
    assert errs[1] < errs[0] and errs[2] < errs[0]


def test_BLDTT(bt, oracle_mod):
    """serial_tests.c:1417-1537: 5000 targets and sources from srand(1),
    cluster-cluster, theta 0.9, degree 3, leaves 20/20, size_check 1, Coulomb
    skipping: the summed potential within 6e-3 of the direct sum's."""
    t, s = glibc_rand_particles(5000, with_targets=True)
    args = (1, [0.5], 1, 1, 3, 0.9, 3, 20, 20, 1.0)
    got = interface(bt, t, s, *args)
    ref = oracle_mod.treecode(*t, *s, *args)
    assert rel_l2(ref, got) <= TOL_FP64
    d = oracle_mod.direct(*t, *s, 1, [0.5], 1)
    assert abs(got.sum() - d.sum()) / abs(d.sum()) < 6e-3
