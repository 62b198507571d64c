"""GPU parity of the kernels beside the Coulomb family (regularized
Yukawa, atan, sin-over-r, MQ, RBS-u, RBS-v, the reference's user-kernel
example) through the C ABI against the CPU oracle, which is bit-equal to
the reference's compiled kernels (test_oracle_kernels.py).  They run in
FP64 whatever precision is asked.  Bar: relative L2 <= 1e-12.
"""
import numpy as np
import pytest

from tests.helpers import KP_OTHER, other_combos, rel_l2, uniform_cube

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-12


@pytest.fixture(scope="module")
def bt():
    import barytree_amd
    barytree_amd.lib()
    return barytree_amd


@pytest.mark.parametrize("ct,k,s", other_combos())
def test_other_kernels_small(bt, oracle_mod, ct, k, s):
    print(f"case ct={ct} k={k} s={s}", flush=True)
    t, src = uniform_cube(3000, 2500, seed=700 + 10 * ct + k)
    args = (k, KP_OTHER[k], s, 1, ct, 0.7, 4, 60, 50)
    ref = oracle_mod.treecode(*t, *src, *args)
    got = bt.solve(t, src, *args)
    err = rel_l2(ref, got)
    print(f"relL2={err:.3e}", flush=True)
    assert err <= TOL_FP64
    # the mixed precision request runs these kernels in FP64
    assert np.array_equal(got, bt.solve(t, src, *args, precision=bt.PRECISION_MIXED))


@pytest.mark.parametrize("k", [4, 8, 10, 11, 12])
def test_other_kernels_cc_lists(bt, oracle_mod, k):
    print(f"case k={k}", flush=True)
    t, src = uniform_cube(12000, 12000, seed=17)
    args = (k, KP_OTHER[k], 1, 1, 3, 0.8, 2, 30, 30)
    ref = oracle_mod.treecode(*t, *src, *args)
    st = oracle_mod.last_stats()
    assert st["num_cc"] > 0
    assert rel_l2(ref, bt.solve(t, src, *args)) <= TOL_FP64


def test_other_kernels_direct_sample(bt, oracle_mod):
    t, src = uniform_cube(3000, 3000, seed=4)
    idx = np.arange(0, 3000, 7)
    for k, s in [(4, 2), (5, 1), (9, 1), (10, 1)]:
        with bt.Plan(t, src, k, KP_OTHER[k], s, 1, 1, 0.7, 4, 60, 60) as p:
            got = p.direct_sample(idx)
        ref = oracle_mod.direct(*[a[idx] for a in t], *src, k, KP_OTHER[k], s)
        assert rel_l2(ref, got) <= TOL_FP64


def test_other_kernels_reject_learned_degrees(bt):
    t, src = uniform_cube(2000, 2000, seed=2)
    with bt.Plan(t, src, 8, [1.5], 1, 1, 1, 0.7, 3, 60, 60) as p:
        with pytest.raises(RuntimeError, match="per-pair degrees"):
            p.set_pair_degrees(np.full(p.num_pairs(), 3))
