"""GPU tests of batching: many independent systems in one set of launches.

Each system of a batch must come out exactly as a single-system solve of the
same system with the same leaf sizes: its targets see only its own sources,
and a target's sum runs over the same sources in the same order whatever
the tile shape (256 or 512 targets) or launch form (one block per tile, or
resident blocks popping per-XCD queues).  So the bar against single solves
is bit equality; against the oracle it is the plain-mode bar, 1e-12.
"""
import numpy as np
import pytest

from barytree_amd.workloads import batched_sizes, gaussian_mixture
from tests.helpers import KP, rel_l2, uniform_cube

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-12


@pytest.fixture(scope="module")
def bt():
    import barytree_amd
    barytree_amd.lib()
    return barytree_amd


def systems_uneven(seed, sizes, m_over_n=1.0):
    out = []
    for i, n in enumerate(sizes):
        m = max(1, int(round(n * m_over_n)))
        t, s = uniform_cube(int(n), m, seed=seed + i)
        out.append((t, s))
    return out


def single(bt, sysm, args, leaf):
    t, s = sysm
    return bt.solve(t, s, *args, leaf, leaf)


@pytest.mark.parametrize("ct,k,sing", [(1, 1, 1), (1, 2, 2), (1, 3, 2), (2, 1, 2), (2, 2, 1),
                                       (3, 1, 2), (3, 2, 1), (3, 3, 1)])
def test_batched_equals_single_solves(bt, ct, k, sing):
    print(f"case ct={ct} k={k} sing={sing}", flush=True)
    sizes = [1, 7, 300, 1200, 2500, 64, 4000, 900]
    systems = systems_uneven(50 + ct * 10 + k, sizes, m_over_n=0.8)
    leaf = [60, 20, 80, 40, 100, 30, 120, 50]
    args = (k, KP[k][sing], sing, 1, ct, 0.7, 3)
    ref = [bt.solve(t, s, *args, lf, lf) for (t, s), lf in zip(systems, leaf)]
    for tpt in (1, 2):
        for pers in (0, 1):
            sc = bt.Schedule.make({}, targets_per_thread=tpt, persistent=pers)
            got = bt.solve_batched(systems, *args, max_source_leaf=leaf, max_target_leaf=leaf,
                                   schedule=sc)
            st = bt.last_stats()
            assert st["num_systems"] == len(systems)
            assert st["targets_per_thread"] == tpt and st["persistent"] == pers
            for i, (r, g) in enumerate(zip(ref, got)):
                assert np.array_equal(r, g), (tpt, pers, i, rel_l2(r, g))


@pytest.mark.parametrize("ct,k,sing", [(1, 1, 1), (2, 3, 1), (3, 2, 2)])
def test_batched_against_oracle(bt, oracle_mod, ct, k, sing):
    print(f"case ct={ct} k={k} sing={sing}", flush=True)
    sizes = [500, 3000, 1500, 40, 2200]
    systems = systems_uneven(900 + ct, sizes)
    leaf = 70
    args = (k, KP[k][sing], sing, 1, ct, 0.75, 4)
    got = bt.solve_batched(systems, *args, max_source_leaf=leaf, max_target_leaf=leaf,
                           schedule=bt.Schedule.make({}, 2, 1))
    for (t, s), g in zip(systems, got):
        ref = oracle_mod.treecode(*t, *s, *args, leaf, leaf)
        assert rel_l2(ref, g) <= TOL_FP64


def test_schedule_leaf_by_size_class(bt, oracle_mod):
    # classes: 2^c <= max(M, N) < 2^(c+1); unset classes take the nearest set one
    systems = systems_uneven(77, [700, 3000, 20000])
    args = (1, [0.5], 1, 1, 1, 0.7, 4)
    sc = bt.Schedule.make({9: 64, 11: 200, 14: 500}, 2, 1)
    got = bt.solve_batched(systems, *args, schedule=sc)
    for (t, s), g, leaf in zip(systems, got, (64, 200, 500)):
        ref = oracle_mod.treecode(*t, *s, *args, leaf, leaf)
        assert rel_l2(ref, g) <= TOL_FP64
    sc2 = bt.Schedule.make({12: 300}, 1, 0)
    got2 = bt.solve_batched(systems, *args, schedule=sc2)
    for (t, s), g in zip(systems, got2):
        ref = oracle_mod.treecode(*t, *s, *args, 300, 300)
        assert rel_l2(ref, g) <= TOL_FP64


def test_autotuner_schedule_and_solve(bt, oracle_mod):
    sizes = batched_sizes(24, 3, nmin=1000, nmax=16000)
    systems = []
    for i, n in enumerate(sizes):
        x, y, z, q, w = gaussian_mixture(int(n), 3000 + i)
        systems.append(((x, y, z, q), (x.copy(), y.copy(), z.copy(), q.copy(), w)))
    args = (1, [0.0], 1, 1, 1, 0.7, 6)
    sc = bt.autotune(systems, *args)
    d = sc.as_dict()
    print("schedule", d, flush=True)
    P = 7 ** 3
    assert d["targets_per_thread"] in (1, 2) and d["persistent"] in (0, 1)
    classes = {int(np.floor(np.log2(n))) for n in sizes}
    assert set(d["leaf"]) == classes
    assert all(v in (P // 2, P, 2 * P, 4 * P, 8 * P) for v in d["leaf"].values())
    assert d["trials"] == 5 * len(classes) + 4
    # the same schedule through the solve, checked against the oracle
    got = bt.solve_batched(systems, *args, schedule=sc)
    for ((t, s), g, n) in zip(systems, got, sizes):
        leaf = d["leaf"][int(np.floor(np.log2(n)))]
        ref = oracle_mod.treecode(*t, *s, *args, leaf, leaf)
        assert rel_l2(ref, g) <= TOL_FP64
    # no leaf sizes and no schedule: the autotuner runs (cached afterwards)
    got2 = bt.solve_batched(systems[:5], *args)
    assert bt.last_stats()["num_systems"] == 5
    for (t, s), g in zip(systems[:5], got2):
        assert np.all(np.isfinite(g))


def test_batched_plan_mixed_and_direct(bt, oracle_mod):
    systems = systems_uneven(5, [4000, 1000, 9000, 2500])
    args = (1, [0.5], 2, 1, 3, 0.7, 8)
    leaf = [300, 300, 500, 300]
    plan = bt.Plan.batched(systems, *args, max_source_leaf=leaf, max_target_leaf=leaf,
                           schedule=bt.Schedule.make({}, 2, 1), precision=bt.PRECISION_MIXED)
    try:
        mixed = plan.execute()
        off = plan.offsets
        idx = np.concatenate([np.arange(off[i], off[i + 1], 37) for i in range(len(systems))])
        direct = plan.direct_sample(idx)
        # FP64 direct sums per system, each against its own sources
        for i, (t, s) in enumerate(systems):
            sel = (idx >= off[i]) & (idx < off[i + 1])
            loc = idx[sel] - off[i]
            ref = oracle_mod.direct(t[0][loc], t[1][loc], t[2][loc], t[3][loc], *s, 1, [0.5], 2)
            assert rel_l2(ref, direct[sel]) <= 1e-13
        assert rel_l2(direct, mixed[idx]) <= 1e-6
        plan.set_precision(bt.PRECISION_FP64)
        plain = plan.execute()
        for i, ((t, s), lf) in enumerate(zip(systems, leaf)):
            ref = oracle_mod.treecode(*t, *s, *args, lf, lf)
            assert rel_l2(ref, plain[off[i]:off[i + 1]]) <= TOL_FP64
        again = plan.execute()
        assert np.array_equal(plain, again)
    finally:
        plan.close()


def test_batched_beta_sets_every_leaf(bt, oracle_mod):
    systems = systems_uneven(11, [3000, 5000])
    args = (2, [0.5], 1, 1, 1, -1.0, -1)
    got = bt.solve_batched(systems, *args, beta=0.4)
    for (t, s), g in zip(systems, got):
        ref = oracle_mod.treecode(*t, *s, 2, [0.5], 1, 1, 1, -1.0, -1, 1, 1, beta=0.4)
        assert rel_l2(ref, g) <= TOL_FP64
