"""Symmetric near field: mixed executions of a system whose targets are its
sources in the same tree order (Coulomb, Yukawa, or regularized Coulomb with
skipping; PC, CP or CC with equal leaf sizes) evaluate a pair of leaves once
for both particles.  Checked against the oracle (the
reference's FP64 arithmetic) at the mixed-precision bar, against the same
system solved without the symmetric path (a batched plan of two systems
never takes it), for run-to-run reproducibility, on pairs planted across
leaf boundaries far closer than the expansion of r^2 resolves, and with
per-pair degrees.
"""
import numpy as np
import pytest

from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu

KPS = {1: [0.5], 2: [0.5], 3: [0.05]}
CASES = [(1, 2), (2, 2), (1, 1), (2, 1), (3, 1)]  # (kernel, singularity)


@pytest.fixture(scope="module")
def bt():
    import barytree_amd
    barytree_amd.lib()
    return barytree_amd


def same_set(n, seed, wlo=0.5, whi=1.5):
    """Targets equal to the sources, particle for particle (separate arrays)."""
    rng = np.random.default_rng(seed)
    x, y, z = (rng.uniform(-1, 1, n) for _ in range(3))
    q = rng.uniform(-1, 1, n)
    w = rng.uniform(wlo, whi, n)
    return (x.copy(), y.copy(), z.copy(), q.copy()), (x, y, z, q, w)


def tiny_system():
    rng = np.random.default_rng(99)
    x, y, z, q = (rng.uniform(-1, 1, 50) for _ in range(4))
    return (x.copy(), y.copy(), z.copy(), q.copy()), (x, y, z, q, np.ones(50))


@pytest.mark.parametrize("k,sg", CASES)
def test_symmetric_matches_oracle_and_is_reproducible(bt, oracle_mod, k, sg):
    print(f"case kernel={k} singularity={sg}", flush=True)
    t, s = same_set(20000, 10 + k + sg)
    args = (k, KPS[k], sg, 1, 3, 0.8, 5, 400, 400)
    with bt.Plan(t, s, *args, precision=bt.PRECISION_MIXED) as plan:
        st = plan.stats()
        assert st["sym_ptwise"] > 0.8 * st["pp_ptwise"], st
        got = plan.execute()
        assert np.array_equal(got, plan.execute())  # every sum has one fixed order
        plan.set_precision(bt.PRECISION_FP64)
        fp64 = plan.execute()
    ref = oracle_mod.treecode(*t, *s, *args)
    assert rel_l2(ref, fp64) <= 1e-12
    err = rel_l2(ref, got)
    print(f"symmetric mixed vs oracle: {err:.3e}", flush=True)
    assert err < 1e-6


@pytest.mark.parametrize("ct,k,sg", [(1, 1, 2), (1, 2, 1), (1, 3, 1), (2, 1, 1), (2, 2, 2), (2, 3, 1)])
def test_symmetric_particle_cluster_and_cluster_particle(bt, oracle_mod, ct, k, sg):
    """PC (target batches) and CP (source batches) plans with equal leaf
    sizes order targets and sources alike and take the symmetric path."""
    print(f"case ct={ct} kernel={k} singularity={sg}", flush=True)
    t, s = same_set(20000, 50 + 3 * ct + k + sg)
    args = (k, KPS[k], sg, 1, ct, 0.7, 5, 300, 300)
    with bt.Plan(t, s, *args, precision=bt.PRECISION_MIXED) as plan:
        st = plan.stats()
        assert st["sym_ptwise"] > 0.8 * st["pp_ptwise"], st
        got = plan.execute()
        assert np.array_equal(got, plan.execute())
    ref = oracle_mod.treecode(*t, *s, *args)
    err = rel_l2(ref, got)
    print(f"symmetric mixed vs oracle: {err:.3e}", flush=True)
    assert err < 1e-6


@pytest.mark.parametrize("k,sg", CASES)
def test_symmetric_against_one_directional_mixed(bt, k, sg):
    t, s = same_set(200000, 20 + k + sg, 1.0, 1.0)
    args = (k, KPS[k], sg, 1, 3, 0.8, 8, 2000, 2000)
    with bt.Plan(t, s, *args, precision=bt.PRECISION_MIXED) as plan:
        assert plan.stats()["sym_ptwise"] > 0
        sym = plan.execute()
        plan.set_precision(bt.PRECISION_FP64)
        fp64 = plan.execute()
    with bt.Plan.batched([(t, s), tiny_system()], *args[:7], 2000, 2000,
                         precision=bt.PRECISION_MIXED) as bp:
        assert bp.stats()["sym_ptwise"] == 0
        one = bp.execute()[:len(t[0])]
    d_sym, d_one, d_pair = rel_l2(fp64, sym), rel_l2(fp64, one), rel_l2(one, sym)
    print(f"vs fp64: symmetric {d_sym:.3e}, one-directional {d_one:.3e}; between {d_pair:.3e}", flush=True)
    assert d_sym < 1e-6 and d_one < 1e-6
    assert d_sym < 4 * d_one + 1e-9


@pytest.mark.parametrize("k,sg", CASES)
def test_symmetric_close_pairs_across_leaves(bt, oracle_mod, k, sg):
    """Pairs 2e-10 and 2e-7 apart straddling the root's x split (x = 0: the
    box is [-1, 1] in x), so each pair's particles sit in different leaves:
    the expanded r^2 cannot resolve them, the close recomputation must."""
    t, s = same_set(12000, 30 + k + sg)
    x, y, z, q, w = s
    x[0], x[1] = -1.0, 1.0
    rng = np.random.default_rng(7)
    for j in range(100):
        i = 2 + 2 * j
        d = 1e-10 if j % 2 == 0 else 1e-7
        x[i], x[i + 1] = -d, d
        y[i + 1], z[i + 1] = y[i], z[i]
    t = (x.copy(), y.copy(), z.copy(), q.copy())
    args = (k, KPS[k], sg, 1, 3, 0.8, 4, 300, 300)
    with bt.Plan(t, s, *args, precision=bt.PRECISION_MIXED) as plan:
        assert plan.stats()["sym_ptwise"] > 0
        got = plan.execute()
    ref = oracle_mod.treecode(*t, *s, *args)
    assert np.all(np.isfinite(got))
    # a planted particle's sum is dominated by its partner's term, whose FP32
    # evaluation keeps a few 1e-7 of the magnitude of its parts (regularized:
    # the pair's term is bounded by 1/eps)
    for j in range(100):
        i = 2 + 2 * j
        r = 2e-10 if j % 2 == 0 else 2e-7
        if k == 3:
            r = np.sqrt(r * r + KPS[3][0] ** 2)
        for a, b in ((i, i + 1), (i + 1, i)):
            mag = (abs(q[b] * w[b]) + (abs(q[a] * w[b]) if sg == 2 else 0.0)) / r
            assert abs(got[a] - ref[a]) <= 2e-6 * mag + 1e-6 * abs(ref[a]), (j, a)
    rest = np.ones(len(got), bool)
    rest[2:202] = False
    assert rel_l2(ref[rest], got[rest]) < 1e-6


@pytest.mark.parametrize("leaf", [1, 7, 33])
def test_symmetric_tiny_leaves(bt, oracle_mod, leaf):
    """Leaves far smaller than a pass (128 targets) or a staged chunk (256
    sources): every item is mostly padding lanes and dummy sources."""
    t, s = same_set(3000, 60 + leaf)
    args = (1, KPS[1], 2, 1, 3, 0.8, 3, leaf, leaf)
    with bt.Plan(t, s, *args, precision=bt.PRECISION_MIXED) as plan:
        assert plan.stats()["sym_ptwise"] > 0
        got = plan.execute()
    ref = oracle_mod.treecode(*t, *s, *args)
    assert rel_l2(ref, got) < 1e-6


def test_symmetric_with_pair_degrees(bt):
    t, s = same_set(100000, 40)
    args = (1, KPS[1], 2, 1, 3, 0.8, 8, 1000, 1000)
    with bt.Plan(t, s, *args, precision=bt.PRECISION_MIXED) as plan:
        rng = np.random.default_rng(3)
        deg = rng.integers(6, 11, plan.num_pairs())
        plan.set_pair_degrees(deg)
        mixed = plan.execute()
        plan.set_precision(bt.PRECISION_FP64)
        fp64 = plan.execute()
    d = rel_l2(fp64, mixed)
    print(f"per-pair degrees, symmetric mixed vs fp64: {d:.3e}", flush=True)
    assert d < 1e-6


# ---- FP64 (plain) executions: the same items in FP64 arithmetic ----------

@pytest.mark.parametrize("ct,k,sg", [(3, 1, 2), (3, 1, 1), (3, 2, 2), (3, 2, 1), (3, 3, 1),
                                     (1, 1, 2), (2, 2, 1), (1, 3, 1)])
def test_symmetric_fp64_matches_oracle(bt, oracle_mod, ct, k, sg):
    """Plain executions take the symmetric path too: the plain-mode bar
    (1e-12 against the reference's arithmetic), repeat runs bit-equal, and
    the one-directional FP64 run of the same plan agrees to rounding."""
    print(f"case ct={ct} kernel={k} singularity={sg}", flush=True)
    t, s = same_set(20000, 70 + 3 * ct + k + sg)
    args = (k, KPS[k], sg, 1, ct, 0.8, 5, 300, 300)
    with bt.Plan(t, s, *args, precision=bt.PRECISION_FP64) as plan:
        st = plan.stats()
        assert st["sym_ptwise"] > 0.8 * st["pp_ptwise"], st
        got = plan.execute()
        st = plan.stats()
        assert st["gpu_near_sym_s"] > 0.0
        assert np.array_equal(got, plan.execute())
        plan.set_symmetric(False)
        assert plan.stats()["sym_ptwise"] == 0
        one = plan.execute()
        assert plan.stats()["gpu_near_sym_s"] == 0.0
    ref = oracle_mod.treecode(*t, *s, *args)
    e_sym, e_one, e_pair = rel_l2(ref, got), rel_l2(ref, one), rel_l2(one, got)
    print(f"fp64 vs oracle: symmetric {e_sym:.3e}, one-directional {e_one:.3e}; between {e_pair:.3e}",
          flush=True)
    assert e_sym <= 1e-12 and e_one <= 1e-12
    assert e_pair <= 1e-13


@pytest.mark.parametrize("k,sg", CASES)
def test_symmetric_fp64_close_pairs_across_leaves(bt, oracle_mod, k, sg):
    """Planted pairs 2e-10 and 2e-7 apart across the root's x split: FP64
    differences resolve them; every particle within 1e-12 of the oracle
    relative to the magnitude of its terms."""
    t, s = same_set(12000, 80 + k + sg)
    x, y, z, q, w = s
    x[0], x[1] = -1.0, 1.0
    for j in range(100):
        i = 2 + 2 * j
        d = 1e-10 if j % 2 == 0 else 1e-7
        x[i], x[i + 1] = -d, d
        y[i + 1], z[i + 1] = y[i], z[i]
    t = (x.copy(), y.copy(), z.copy(), q.copy())
    args = (k, KPS[k], sg, 1, 3, 0.8, 4, 300, 300)
    with bt.Plan(t, s, *args, precision=bt.PRECISION_FP64) as plan:
        assert plan.stats()["sym_ptwise"] > 0
        got = plan.execute()
    ref = oracle_mod.treecode(*t, *s, *args)
    assert np.all(np.isfinite(got))
    for j in range(100):
        i = 2 + 2 * j
        r = 2e-10 if j % 2 == 0 else 2e-7
        if k == 3:
            r = np.sqrt(r * r + KPS[3][0] ** 2)
        for a, b in ((i, i + 1), (i + 1, i)):
            mag = (abs(q[b] * w[b]) + (abs(q[a] * w[b]) if sg == 2 else 0.0)) / r
            assert abs(got[a] - ref[a]) <= 1e-12 * (mag + abs(ref[a])), (j, a)
    assert rel_l2(ref, got) <= 1e-12


@pytest.mark.parametrize("leaf", [1, 7, 33])
def test_symmetric_fp64_tiny_leaves(bt, oracle_mod, leaf):
    t, s = same_set(3000, 90 + leaf)
    args = (1, KPS[1], 2, 1, 3, 0.8, 3, leaf, leaf)
    with bt.Plan(t, s, *args, precision=bt.PRECISION_FP64) as plan:
        assert plan.stats()["sym_ptwise"] > 0
        got = plan.execute()
    ref = oracle_mod.treecode(*t, *s, *args)
    assert rel_l2(ref, got) <= 1e-12


def test_symmetric_fp64_through_the_reference_entry_point(bt, oracle_mod):
    """The drop-in BaryTreeInterface (plain mode) on a system whose targets
    are its sources takes the symmetric path and keeps the plain-mode bar."""
    t, s = same_set(30000, 95)
    args = (1, KPS[1], 2, 1, 3, 0.8, 6, 500, 500)
    got = bt.callTreedriver(len(t[0]), len(s[0]), *t, *s[:4], s[4], bt.Kernel.COULOMB, 1, KPS[1],
                            bt.Singularity.SUBTRACTION, bt.Approximation.LAGRANGE,
                            bt.ComputeType.CLUSTER_CLUSTER, True, 0, theta=0.8, degree=6,
                            sourceLeafSize=500, targetLeafSize=500)
    st = bt.last_stats()
    assert st["sym_ptwise"] > 0, st
    ref = oracle_mod.treecode(*t, *s, *args)
    assert rel_l2(ref, got) <= 1e-12


@pytest.mark.parametrize("ct,k", [(1, 1), (2, 2), (3, 3), (3, 1)])
def test_one_sided_items_for_other_targets(bt, oracle_mod, ct, k):
    """One system whose targets differ from its sources, skipping kernel:
    in mixed precision every near particle pair runs as a one-sided item
    (ordered pairs delivered = pairs evaluated = the plan's PP count).
    Against the oracle and against the one-directional near launch (items
    switched off) at the mixed bar; FP64 executions keep the one-directional
    launch (the same bits with items on or off)."""
    print(f"case ct={ct} kernel={k} other targets, mixed", flush=True)
    rng = np.random.default_rng(40 + 10 * ct + k)
    s = tuple(rng.uniform(-1, 1, 20000) for _ in range(4)) + (np.ones(20000),)
    t = tuple(rng.uniform(-1, 1, 15000) for _ in range(4))
    args = (k, KPS[k], 1, 1, ct, 0.7, 6, 500, 500)
    ref = oracle_mod.treecode(*t, *s, *args)
    with bt.Plan(t, s, *args, precision=bt.PRECISION_MIXED) as plan:
        got = plan.execute()
        st = plan.stats()
        plan.set_symmetric(False)
        one = plan.execute()
        plan.set_symmetric(True)
        plan.set_precision(bt.PRECISION_FP64)
        fp64 = plan.execute()
        st64 = plan.stats()
        plan.set_symmetric(False)
        fp64_off = plan.execute()
    assert st["sym_ptwise"] == st["sym_evals"] == st["pp_ptwise"] > 0, st
    assert st["gpu_near_sym_s"] > 0.0
    assert st64["sym_ptwise"] == 0 and st64["sym_slots"] == 0
    assert np.array_equal(fp64, fp64_off)
    assert rel_l2(ref, fp64) <= 1e-12
    err, diff = rel_l2(ref, got), rel_l2(one, got)
    print(f"items vs oracle {err:.3e}, vs one-directional launch {diff:.3e}", flush=True)
    assert err < 1e-6 and diff < 1e-6


def test_two_stream_executions_repeat_across_setting_changes(bt):
    """A symmetric plan's executions (the item launch on the plan's stream,
    the other phases on the second stream) give the same bits as fresh
    plans with the same settings, whatever ran before on the plan: precision
    switches, the symmetric near field switched off and on, per-pair degrees
    set and cleared, and the same through a learned one-call solve against
    its plan form."""
    from barytree_amd import learned
    rng = np.random.default_rng(77)
    n = 60000
    x, y, z, q = (rng.uniform(-1, 1, n) for _ in range(4))
    w = np.ones(n)
    src, tgt = (x, y, z, q, w), (x.copy(), y.copy(), z.copy(), q.copy())
    args = (1, [0.5], 2, 1, 3, 0.8, 8, 1000, 1000)

    def fresh(precision, sym=True, degrees=None):
        with bt.Plan(tgt, src, *args, precision=precision) as p:
            p.set_symmetric(sym)
            if degrees is not None:
                p.set_pair_degrees(degrees)
            return p.execute()

    with bt.Plan(tgt, src, *args) as p:
        degs = np.random.default_rng(3).integers(2, 12, p.num_pairs())
        seq = []
        for prec, sym, d in ((bt.PRECISION_FP64, True, None), (bt.PRECISION_MIXED, True, None),
                             (bt.PRECISION_MIXED, False, None), (bt.PRECISION_MIXED, True, degs),
                             (bt.PRECISION_FP64, True, degs), (bt.PRECISION_MIXED, True, None),
                             (bt.PRECISION_FP64, True, None)):
            p.set_precision(prec)
            p.set_symmetric(sym)
            p.set_pair_degrees(d)
            seq.append((prec, sym, d, p.execute(), p.stats()))
    for prec, sym, d, got, st in seq:
        print(f"case precision={prec} symmetric={sym} degrees={'set' if d is not None else 'fixed'}", flush=True)
        assert (st["sym_ptwise"] > 0) == sym
        assert np.array_equal(got, fresh(prec, sym, d))
    model = learned.load_model("coulomb_cc")
    one_call = bt.solve_learned(tgt, src, *args, model=model, delta=0.5, precision=bt.PRECISION_MIXED)
    with bt.Plan(tgt, src, *args, precision=bt.PRECISION_MIXED) as p:
        p.use_model(model, 0.5)
        planned = p.execute()
    assert np.array_equal(one_call, planned)
