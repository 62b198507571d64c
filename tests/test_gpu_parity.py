"""GPU parity: the HIP path through the C ABI against the CPU oracle.

Bar (plain FP64 mode): relative L2 difference <= 1e-12 against the oracle on
the same inputs (the oracle is bit-identical to the reference's compiled
sources, see test_oracle.py).  Expected differences are summation-order
rounding, about 1e-15.
"""
import numpy as np
import pytest

from tests.helpers import KP, kat_inputs, rel_l2, supported_combos, uniform_cube

pytestmark = pytest.mark.gpu

TOL_FP64 = 1e-12


@pytest.fixture(scope="module")
def bt():
    import barytree_amd
    barytree_amd.lib()
    return barytree_amd


@pytest.mark.parametrize("ct,k,s", supported_combos())
def test_modes_small(bt, oracle_mod, ct, k, s):
    # degree 3, leaves 60/50: every list type is populated
    t, src = uniform_cube(3000, 2500, seed=100 + 10 * ct + k)
    args = (k, KP[k][s], s, 1, ct, 0.7, 3, 60, 50)
    ref = oracle_mod.treecode(*t, *src, *args)
    got = bt.solve(t, src, *args)
    err = rel_l2(ref, got)
    print(f"ct={ct} k={k} s={s} relL2={err:.3e}")
    assert err <= TOL_FP64


@pytest.mark.parametrize("ct,k,s", [(3, 1, 1), (3, 2, 2), (3, 1, 2), (3, 3, 1)])
def test_cc_with_cluster_cluster_lists(bt, oracle_mod, ct, k, s):
    # degree 2, leaves 30: the dual traversal yields CC, PC and CP entries
    t, src = uniform_cube(12000, 12000, seed=7)
    args = (k, KP[k][s], s, 1, ct, 0.8, 2, 30, 30)
    ref = oracle_mod.treecode(*t, *src, *args)
    st = oracle_mod.last_stats()
    assert st["num_cc"] > 0 and st["num_pc"] > 0 and st["num_cp"] > 0
    got = bt.solve(t, src, *args)
    assert rel_l2(ref, got) <= TOL_FP64


def test_known_answer_yukawa_ss(bt, oracle_mod):
    # interfaces/python/testBaryTreeInterface.py:41-61 and :68-76
    x, y, z, rho, w = kat_inputs()
    n = len(x)
    out = bt.callTreedriver(n, n, x, y, z, rho, x.copy(), y.copy(), z.copy(), rho.copy(),
                            w.copy(), bt.Kernel.YUKAWA, 1, np.array([0.5]),
                            bt.Singularity.SUBTRACTION, bt.Approximation.LAGRANGE,
                            bt.ComputeType.PARTICLE_CLUSTER, True, 0, theta=0.8, degree=4,
                            sourceLeafSize=50, targetLeafSize=10, sizeCheck=1.0)
    assert abs(out[0] - 588.7432483318685) < 1e-12 * 588.74
    ref = oracle_mod.treecode(x, y, z, rho, x.copy(), y.copy(), z.copy(), rho.copy(), w,
                              2, [0.5], 2, 1, 1, 0.8, 4, 50, 10)
    assert rel_l2(ref, out) <= TOL_FP64
    out_b = bt.callTreedriver(n, n, x, y, z, rho, x.copy(), y.copy(), z.copy(), rho.copy(),
                              w.copy(), bt.Kernel.YUKAWA, 1, np.array([0.5]),
                              bt.Singularity.SUBTRACTION, bt.Approximation.LAGRANGE,
                              bt.ComputeType.PARTICLE_CLUSTER, True, 0, beta=0.1, sizeCheck=1.0)
    assert abs(out_b[0] - 588.7445889051367) < 1e-12 * 588.74


def test_caller_arrays_untouched_and_aliasing(bt, oracle_mod):
    t, s = uniform_cube(4000, 4000, seed=3)
    # targets alias the sources (src/particles/particles.c:195-217 duplicates them)
    tx, ty, tz, tq = s[0], s[1], s[2], s[3]
    keep = [a.copy() for a in s]
    out = bt.solve((tx, ty, tz, tq), s, 1, [0.5], 2, 1, 3, 0.7, 4, 200, 200)
    for a, b in zip(keep, s):
        assert np.array_equal(a, b)
    ref = oracle_mod.treecode(tx, ty, tz, tq, *s, 1, [0.5], 2, 1, 3, 0.7, 4, 200, 200)
    assert rel_l2(ref, out) <= TOL_FP64


def test_edge_cases(bt, oracle_mod):
    # one particle each; all particles in one leaf; coincident points
    one = [np.array([0.3])] * 4
    out = bt.solve(one, one + [np.array([1.0])], 1, [0.5], 1, 1, 1, 0.7, 3, 10, 10)
    assert out[0] == 0.0  # self-interaction skipped (coulomb_pp.c:48)
    t, s = uniform_cube(300, 200, seed=11)
    for ct in (1, 2, 3):
        ref = oracle_mod.treecode(*t, *s, 1, [0.5], 1, 1, ct, 0.7, 3, 1000, 1000)
        got = bt.solve(t, s, 1, [0.5], 1, 1, ct, 0.7, 3, 1000, 1000)
        assert rel_l2(ref, got) <= TOL_FP64
    # duplicated source points and targets on sources (exact node hits)
    rng = np.random.default_rng(5)
    xs = np.repeat(rng.uniform(-1, 1, 500), 2)
    pts = [xs, np.repeat(rng.uniform(-1, 1, 500), 2), np.repeat(rng.uniform(-1, 1, 500), 2)]
    q = rng.uniform(-1, 1, 1000)
    src = pts + [q, np.ones(1000)]
    tg = [p.copy() for p in pts] + [q.copy()]
    for ct in (1, 2, 3):
        ref = oracle_mod.treecode(*tg, *src, 1, [0.5], 1, 1, ct, 0.7, 3, 20, 20)
        got = bt.solve(tg, src, 1, [0.5], 1, 1, ct, 0.7, 3, 20, 20)
        assert rel_l2(ref, got) <= TOL_FP64


def test_errors_are_reported(bt):
    t, s = uniform_cube(100, 100, seed=1)
    with pytest.raises(RuntimeError, match="weights"):
        bt.solve(t, s[:4] + [None], 1, [0.5], 2, 1, 1, 0.7, 3, 10, 10)
    with pytest.raises(RuntimeError, match="CC HERMITE"):
        bt.solve(t, s, 1, [0.5], 1, 2, 3, 0.7, 3, 10, 10)
    with pytest.raises(RuntimeError, match="REGULARIZED"):
        bt.solve(t, s, 3, [0.3, 0.05], 2, 1, 3, 0.7, 3, 10, 10)
    with pytest.raises(RuntimeError):
        bt.solve(t, s, 7, [0.5], 1, 1, 1, 0.7, 3, 10, 10)
    # the reference entry point prints and leaves zeros instead of exiting
    out = bt.callTreedriver(100, 100, *t, *s[:4], None, bt.Kernel.COULOMB, 1, [0.5],
                            bt.Singularity.SUBTRACTION, bt.Approximation.LAGRANGE,
                            bt.ComputeType.PARTICLE_CLUSTER, True, 0, theta=0.7, degree=3,
                            sourceLeafSize=10, targetLeafSize=10)
    assert np.all(out == 0.0)
    assert "weights" in bt.last_error()


def test_plan_reuse_and_direct(bt, oracle_mod):
    t, s = uniform_cube(20000, 20000, seed=21)
    args = (1, [0.5], 1, 1, 3, 0.7, 6, 500, 500)
    with bt.Plan(t, s, *args) as plan:
        a = plan.execute()
        b = plan.execute()
        assert np.array_equal(a, b)
        idx = np.arange(0, 20000, 20)
        d = plan.direct_sample(idx)
    ref_d = oracle_mod.direct(*[x[idx] for x in t], *s, 1, [0.5], 1)
    assert rel_l2(ref_d, d) <= 1e-13
    err = rel_l2(d, a[idx])
    assert err < 1e-4


@pytest.mark.parametrize("ct,k,s", [(1, 1, 1), (3, 2, 1), (3, 1, 2), (2, 3, 1)])
def test_mixed_precision_accuracy(bt, ct, k, s):
    t, src = uniform_cube(200000, 200000, seed=31, wlo=1.0, whi=1.0)
    args = (k, KP[k][s], s, 1, ct, 0.7, 8, 2000, 2000)
    fp64 = bt.solve(t, src, *args, precision=bt.PRECISION_FP64)
    mixed = bt.solve(t, src, *args, precision=bt.PRECISION_MIXED)
    d = rel_l2(fp64, mixed)
    print(f"mixed vs fp64 ct={ct} k={k} s={s}: {d:.3e}")
    assert d < 1e-6


def test_against_reference_outputs(bt):
    """The HIP path against the reference's own compiled sources (committed
    outputs, tests/golden/make_golden.py)."""
    import os
    from tests.golden.make_golden import M, N, cases
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_outputs.npz"))
    for name, seed, args in cases():
        t, s = uniform_cube(N, M, seed)
        got = bt.solve(t, s, *args)
        assert rel_l2(g[name], got) <= TOL_FP64, name


@pytest.mark.parametrize("ct,k,s", [(1, 1, 2), (3, 1, 2), (2, 2, 1), (1, 3, 1), (3, 2, 2)])
def test_mixed_near_field_close_and_coincident_pairs(bt, oracle_mod, ct, k, s):
    # targets that coincide with sources (skipped) and targets 1e-9 and 1e-6
    # away from a source: the mixed near field must recompute such pairs from
    # coordinate differences instead of the expanded r^2 of the matrix cores
    print(f"case ct={ct} k={k} s={s}", flush=True)
    t, src = uniform_cube(6000, 4000, seed=31)
    rng = np.random.default_rng(5)
    pick = rng.choice(6000, 1500, replace=False)
    for j, i in enumerate(pick[:2000]):
        d = 0.0 if j % 3 == 0 else (1e-9 if j % 3 == 1 else 1e-6)
        t[0][j] = src[0][i] + d
        t[1][j] = src[1][i]
        t[2][j] = src[2][i] - d
    args = (k, KP[k][s], s, 1, ct, 0.7, 4, 80, 80)
    ref = oracle_mod.treecode(*t, *src, *args)
    got = bt.solve(t, src, *args, precision=bt.PRECISION_MIXED)
    err = rel_l2(ref, got)
    # at a planted close target the close pair dominates; its FP32 term has an
    # error of a few 1e-7 of the magnitude of its parts: |q_s w_s| (+ |q_t w_s|
    # for subtraction) over r
    worst = 0.0
    for j, i in enumerate(pick[:1500]):
        if j % 3 == 0:
            continue
        r = np.sqrt(2.0) * (1e-9 if j % 3 == 1 else 1e-6)
        mag = abs(src[3][i] * (src[4][i] if s == 2 else 1.0))
        if s == 2:
            mag += abs(t[3][j] * src[4][i])
        worst = max(worst, abs(ref[j] - got[j]) / (mag / r))
    print(f"relL2 {err:.2e}, worst close-pair error / term magnitude {worst:.2e}")
    assert err <= 1e-6 and worst <= 2e-6


def test_verbosity_diagnostics(bt, capfd):
    """verbosity > 2 prints the reference's diagnostics (treedriver.c:53-86,
    989-1136): start and end lines, sizes, total time, the clusterwise and
    pointwise interaction tables with the counts of the stats, and the
    point-wise interactions each kind accounts for; verbosity 0 prints
    nothing."""
    import re
    t, s = uniform_cube(20000, 20000, seed=17)
    args = (1, KP[1][2], 2, 1, 3, 0.8, 5, 500, 500)
    bt.solve(t, s, *args, verbosity=3)
    out = capfd.readouterr().out
    st = bt.last_stats()
    assert "BaryTree has started" in out and "BaryTree has finished" in out
    assert "20000 target and 20000 source particles" in out and "Total BaryTree time" in out
    cw = out[out.index("Clusterwise interaction information"):out.index("Pointwise interaction information")]
    rows = {m.group(1): int(m.group(2)) for m in re.finditer(r"\| (PP|PC|CP|CC) +\| +(\d+)", cw)}
    assert rows["PP"] == st["num_pp"] and rows["CC"] == st["num_cc"]
    pt = re.search(r"Pointwise interaction information.*?\| PP +\| +(\d+)", out, re.S)
    assert int(pt.group(1)) == st["pp_ptwise"]
    acc = re.search(r"accounted for.*?\| Total \| +(\d+)", out, re.S)
    assert int(acc.group(1)) == st["pp_ptwise"] + st["pc_replaced"] + st["cp_replaced"] + st["cc_replaced"]
    assert "GPU: near field" in out
    bt.solve(t, s, *args, verbosity=0)
    assert capfd.readouterr().out == ""
