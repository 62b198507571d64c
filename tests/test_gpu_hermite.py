"""GPU parity of the Hermite approximation (PC with either singularity, CP
with skipping) through the C ABI against the CPU oracle, whose Hermite path
is bit-equal to the reference's compiled Hermite kernels and downpass
(test_oracle_hermite.py).  Bar: the plain-mode bar, relative L2 <= 1e-12.
"""
import numpy as np
import pytest

from tests.helpers import KP, hermite_combos, rel_l2, uniform_cube

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-12


@pytest.fixture(scope="module")
def bt():
    import barytree_amd
    barytree_amd.lib()
    return barytree_amd


@pytest.mark.parametrize("ct,k,s", hermite_combos())
def test_hermite_small(bt, oracle_mod, ct, k, s):
    print(f"case ct={ct} k={k} s={s}", flush=True)
    t, src = uniform_cube(3000, 2500, seed=300 + 10 * ct + k)
    args = (k, KP[k][s], s, 2, ct, 0.7, 3, 60, 50)
    ref = oracle_mod.treecode(*t, *src, *args)
    st = oracle_mod.last_stats()
    assert st["num_pc"] + st["num_cp"] > 0
    got = bt.solve(t, src, *args)
    err = rel_l2(ref, got)
    print(f"relL2={err:.3e}", flush=True)
    assert err <= TOL_FP64
    # a target's sum has one fixed order: runs repeat bit for bit
    assert np.array_equal(got, bt.solve(t, src, *args))


@pytest.mark.parametrize("ct,k,s,deg", [(1, 1, 2, 1), (1, 2, 1, 6), (1, 3, 2, 9), (1, 1, 1, 12),
                                        (2, 1, 1, 1), (2, 2, 1, 8), (2, 3, 1, 12)])
def test_hermite_degrees(bt, oracle_mod, ct, k, s, deg):
    print(f"case ct={ct} k={k} s={s} degree={deg}", flush=True)
    t, src = uniform_cube(2500, 2500, seed=400 + deg)
    args = (k, KP[k][s], s, 2, ct, 0.6, deg, 200, 200)
    ref = oracle_mod.treecode(*t, *src, *args)
    got = bt.solve(t, src, *args)
    assert rel_l2(ref, got) <= TOL_FP64


def test_hermite_large_nodes_in_parts(bt, oracle_mod):
    """A source root of 40000 particles: its upward pass runs in parts of
    16384 particles summed in order."""
    t, src = uniform_cube(2000, 40000, seed=5)
    for s in (1, 2):
        args = (1, KP[1][s], s, 2, 1, 0.7, 2, 800, 800)
        ref = oracle_mod.treecode(*t, *src, *args)
        got = bt.solve(t, src, *args)
        assert rel_l2(ref, got) <= TOL_FP64


def test_hermite_cp_keys_split_by_sources(bt, oracle_mod):
    """Cluster-particle keys with more than 16384 sources run as several keys
    of the same node (rows summed in order): more grid keys than target
    nodes, and the plain-mode bar against the oracle."""
    print("case Hermite CP N=40000 leaf 200 degree 3", flush=True)
    rng = np.random.default_rng(5)
    x, y, z, q = (rng.uniform(-1, 1, 40000) for _ in range(4))
    t, src = (x.copy(), y.copy(), z.copy(), q.copy()), (x, y, z, q, np.ones(40000))
    args = (1, [0.5], 1, 2, 2, 0.7, 3, 200, 200)
    a = bt.analyze(t, src, *args)
    # one tile per key (64 grid points), at most 7 padding slots
    assert a["far_tiles"] - 7 > a["target_nodes"], a
    ref = oracle_mod.treecode(*t, *src, *args)
    got = bt.solve(t, src, *args)
    err = rel_l2(ref, got)
    print(f"relL2={err:.3e}", flush=True)
    assert err <= TOL_FP64


@pytest.mark.parametrize("ct,k,s", [(1, 1, 2), (1, 2, 1), (1, 3, 1), (2, 1, 1), (2, 2, 1)])
def test_hermite_targets_are_sources(bt, oracle_mod, ct, k, s):
    """Targets equal to the sources: the near field runs in the symmetric
    launch (the approximation concerns only the far field).  The plain-mode
    bar against the oracle, and the one-directional near field (symmetric
    form off) gives the same sums up to their order."""
    print(f"case Hermite ct={ct} k={k} s={s} targets=sources N=6000", flush=True)
    _, src = uniform_cube(6000, 1, seed=700 + 10 * ct + k)
    t = [a.copy() for a in src[:4]]
    args = (k, KP[k][s], s, 2, ct, 0.7, 4, 300, 300)
    ref = oracle_mod.treecode(*t, *src, *args)
    with bt.Plan(t, src, *args) as p:
        got = p.execute()
        st = p.stats()
        p.set_symmetric(False)
        one = p.execute()
    assert st["sym_ptwise"] == st["pp_ptwise"] > 0, st
    err = rel_l2(ref, got)
    print(f"relL2={err:.3e}", flush=True)
    assert err <= TOL_FP64
    assert rel_l2(one, got) <= 1e-13


def test_hermite_batched_equals_single_solves(bt):
    systems = []
    for i, n in enumerate([300, 2500, 40, 1800]):
        t, s = uniform_cube(n, int(n * 0.9) + 1, seed=600 + i)
        systems.append((t, s))
    leaf = [50, 70, 20, 60]
    for ct, k, sing in [(1, 1, 2), (1, 2, 1), (2, 3, 1)]:
        args = (k, KP[k][sing], sing, 2, ct, 0.7, 3)
        ref = [bt.solve(t, s, *args, lf, lf) for (t, s), lf in zip(systems, leaf)]
        got = bt.solve_batched(systems, *args, max_source_leaf=leaf, max_target_leaf=leaf,
                               schedule=bt.Schedule.make({}, 2, 1))
        for r, g in zip(ref, got):
            assert np.array_equal(r, g)


def test_hermite_mixed_precision(bt, oracle_mod):
    """Mixed precision: the near field in FP32 pairs, the Hermite far field
    stays FP64."""
    t, src = uniform_cube(4000, 4000, seed=8)
    for ct, k, s in [(1, 1, 2), (2, 2, 1)]:
        args = (k, KP[k][s], s, 2, ct, 0.7, 4, 80, 80)
        ref = oracle_mod.treecode(*t, *src, *args)
        got = bt.solve(t, src, *args, precision=bt.PRECISION_MIXED)
        assert rel_l2(ref, got) <= 2e-6


def test_hermite_rejects_learned_degrees(bt):
    t, src = uniform_cube(2000, 2000, seed=2)
    with bt.Plan(t, src, 1, [0.5], 1, 2, 1, 0.7, 3, 60, 60) as p:
        n = p.num_pairs()
        with pytest.raises(RuntimeError, match="Lagrange"):
            p.set_pair_degrees(np.full(n, 3))
