"""GPU tests of the learned degree selection, through the C ABI.

- variable-degree execution against the oracle's variable-degree treecode
  (plain FP64 bar, 1e-12), and with every pair at the plan's degree against
  the fixed-degree path;
- pair labels of the HIP kernel against the oracle's labels;
- the MLP's forward/backward kernels: gradients and logits against the
  oracle (FP32 kernels vs FP64 oracle: relative 1e-4);
- training lowers the loss; the trained network's gain over fixed degrees.
"""
import numpy as np
import pytest

from tests.helpers import KP, rel_l2, supported_combos, uniform_cube
from tests.test_learn_oracle import make_pairs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def bt():
    import barytree_amd
    barytree_amd.lib()
    return barytree_amd


@pytest.mark.parametrize("ct,k,s", supported_combos())
def test_uniform_pair_degree_equals_fixed(bt, ct, k, s):
    print(f"case ct={ct} k={k} s={s}", flush=True)
    t, src = uniform_cube(3000, 2500, seed=60 + 10 * ct + k)
    args = (k, KP[k][s], s, 1, ct, 0.7, 4, 60, 50)
    with bt.Plan(t, src, *args) as p:
        fixed = p.execute()
        n = p.num_pairs()
        assert n > 0
        p.set_pair_degrees(np.full(n, 4))
        st = p.stats()
        assert st["variable_degree"] == 1 and st["mean_pair_degree"] == 4.0
        var = p.execute()
        assert rel_l2(fixed, var) <= 1e-12
        p.set_pair_degrees(None)
        assert p.stats()["variable_degree"] == 0
        assert np.array_equal(p.execute(), fixed)


@pytest.mark.parametrize("ct,k,s", [(1, 1, 2), (2, 2, 1), (3, 1, 2), (3, 3, 1)])
def test_pair_degrees_below_and_above_the_plan_degree(bt, oracle_mod, ct, k, s):
    """Degrees up to the plan's own (7 here) go through the plan's degree-7
    grids (change of degree on the source and target side), higher ones
    from and to the particles; both against the oracle."""
    print(f"case ct={ct} k={k} s={s}", flush=True)
    t, src = uniform_cube(2500, 2000, seed=90 + 10 * ct + k)
    args = (k, KP[k][s], s, 1, ct, 0.7, 7, 60, 50)
    with bt.Plan(t, src, *args) as p:
        n = p.num_pairs()
        deg = np.random.default_rng(ct * 7 + k).integers(1, 13, n)
        p.set_pair_degrees(deg)
        got = p.execute()
        ref = oracle_mod.treecode_vardeg(*t, *src, *args, deg)
        assert rel_l2(ref, got) <= 1e-12


@pytest.mark.parametrize("ct,k,s", supported_combos())
def test_random_pair_degrees_match_oracle(bt, oracle_mod, ct, k, s):
    print(f"case ct={ct} k={k} s={s}", flush=True)
    t, src = uniform_cube(2500, 2000, seed=80 + 10 * ct + k)
    args = (k, KP[k][s], s, 1, ct, 0.75, 3, 50, 40)
    with bt.Plan(t, src, *args) as p:
        n = p.num_pairs()
        deg = np.random.default_rng(ct * 10 + k).integers(1, 13, n)
        p.set_pair_degrees(deg)
        got = p.execute()
        ref = oracle_mod.treecode_vardeg(*t, *src, *args, deg)
        assert rel_l2(ref, got) <= 1e-12
        p.set_precision(bt.PRECISION_MIXED)
        mixed = p.execute()
        assert rel_l2(ref, mixed) <= 2e-6


def test_random_pair_degrees_batched_persistent(bt, oracle_mod):
    systems = []
    for i, n in enumerate((3000, 800, 2200)):
        t, s = uniform_cube(n, n, seed=500 + i)
        systems.append((t, s))
    args = (1, [0.5], 2, 1, 3, 0.7, 3)
    plan = bt.Plan.batched(systems, *args, max_source_leaf=40, max_target_leaf=40,
                           schedule=bt.Schedule.make({}, 1, 1))
    try:
        kind, system = plan.pair_info()
        deg = np.random.default_rng(9).integers(2, 10, len(kind))
        plan.set_pair_degrees(deg)
        out = plan.execute()
        off = plan.offsets
        for i, (t, s) in enumerate(systems):
            ref = oracle_mod.treecode_vardeg(*t, *s, *args, 40, 40, deg[system == i])
            assert rel_l2(ref, out[off[i]:off[i + 1]]) <= 1e-12
    finally:
        plan.close()


def test_labels_match_oracle(bt, oracle_mod):
    rng = np.random.default_rng(21)
    kinds, tb, tc, sb, sc, tbox, sbox, T, S, Q = make_pairs(rng, 60)
    for kernel, p0, tol in ((1, 0.0, 1e-7), (2, 0.5, 1e-6), (3, 0.01, 1e-8)):
        lg, eg = bt.label_pairs(kinds, tb, tc, sb, sc, tbox, sbox, (T[:, 0], T[:, 1], T[:, 2]),
                                (S[:, 0], S[:, 1], S[:, 2], Q), kernel, p0, tol)
        lo, eo = oracle_mod.label_pairs(kinds, tb, tc, sb, sc, tbox, sbox, T[:, 0], T[:, 1], T[:, 2],
                                        S[:, 0], S[:, 1], S[:, 2], Q, kernel, p0, tol)
        both = ~np.isnan(eo) & ~np.isnan(eg)
        # an error is a difference of O(1) potentials summed in another
        # order: it agrees to about 1e-15 absolute, not relative
        assert np.allclose(eg[both], eo[both], rtol=1e-8, atol=1e-13)
        # labels agree except where an error sits within rounding of tol
        border = np.any(np.abs(eo[both] - tol) <= 1e-13)
        assert np.array_equal(lg, lo) or border


def test_subtraction_labels_match_oracle(bt, oracle_mod):
    """Labels of the subtraction kernels (both terms) against the oracle's
    restatement, which scipy's interpolation pins."""
    rng = np.random.default_rng(22)
    kinds, tb, tc, sb, sc, tbox, sbox, T, S, Q = make_pairs(rng, 60)
    tq = rng.uniform(-1, 1, len(T))
    w = rng.uniform(0.5, 1.5, len(S))
    for kernel, kp, tol in ((1, [0.5], 1e-7), (2, [0.5], 1e-6), (3, [0.3, 0.01], 1e-8)):
        print(f"case kernel={kernel} kp={kp}", flush=True)
        lg, eg = bt.label_pairs_subtraction(kinds, tb, tc, sb, sc, tbox, sbox, (T[:, 0], T[:, 1], T[:, 2], tq),
                                            (S[:, 0], S[:, 1], S[:, 2], Q, w), kernel, kp, tol)
        p0 = {1: 0.0, 2: kp[0], 3: kp[-1]}[kernel]
        p1 = {1: 1.0 / kp[0] ** 2, 2: 0.0, 3: 1.0 / kp[0] ** 2}[kernel]
        lo, eo = oracle_mod.label_pairs_ss(kinds, tb, tc, sb, sc, tbox, sbox, T[:, 0], T[:, 1], T[:, 2], tq,
                                           S[:, 0], S[:, 1], S[:, 2], Q, w, kernel, p0, p1, tol)
        both = ~np.isnan(eo) & ~np.isnan(eg)
        assert np.allclose(eg[both], eo[both], rtol=1e-8, atol=1e-13)
        border = np.any(np.abs(eo[both] - tol) <= 1e-13)
        assert np.array_equal(lg, lo) or border


def test_subtraction_plan_labels_follow_the_rule(bt):
    """A subtraction plan labels its pairs with the subtraction kernel: the
    rule holds, and zero target charges give the skipping kernel's labels."""
    t, src = uniform_cube(20000, 20000, seed=6)
    args = (1, [0.5], 2, 1, 3, 0.8, 6, 300, 300)
    with bt.Plan(t, src, *args) as p:
        lab, err = p.label_pairs(1e-7, with_errors=True)
        for i in range(len(lab)):
            L = lab[i]
            if L <= 12:
                assert err[i, L - 2] <= 1e-7 and np.all(err[i, :L - 2] > 1e-7)
    t0 = (t[0], t[1], t[2], np.zeros(len(t[0])))
    ones = (src[0], src[1], src[2], src[3], np.ones(len(src[0])))
    with bt.Plan(t0, ones, *args) as p0, bt.Plan(t0, ones, 1, [0.5], 1, 1, 3, 0.8, 6, 300, 300) as pk:
        l0, e0 = p0.label_pairs(1e-7, with_errors=True)
        lk, ek = pk.label_pairs(1e-7, with_errors=True)
    assert np.array_equal(l0, lk)


def test_learned_subtraction_small_kappa_against_direct(bt):
    """The shipped Coulomb subtraction model (trained at kappa 0.5) on a
    system with kappa 0.1: the learned mixed run keeps the error bar against
    the FP64 direct sum."""
    from barytree_amd import learned
    n = 400_000
    rng = np.random.default_rng(31)
    x, y, z = (rng.uniform(-1, 1, n) for _ in range(3))
    q = rng.uniform(-1, 1, n)
    src = (x, y, z, q, np.ones(n))
    tgt = (x.copy(), y.copy(), z.copy(), q.copy())
    m = learned.load_model("coulomb_cc")
    idx = np.arange(0, n, n // 1000)[:1000]
    with bt.Plan(tgt, src, 1, [0.1], 2, 1, 3, 0.8, 10, 2000, 2000, precision=bt.PRECISION_MIXED) as p:
        direct = p.direct_sample(idx)
        p.use_model(m, 0.5)
        st = p.stats()
        out = p.execute()
    e = rel_l2(direct, out[idx])
    print(f"kappa 0.1 learned mixed: rel-L2 {e:.2e}, mean degree {st['mean_pair_degree']:.2f}, "
          f"fallback pairs {st['model_fallback_pairs']}", flush=True)
    assert e <= 1e-6


def test_plan_labels_follow_the_rule(bt):
    t, src = uniform_cube(20000, 20000, seed=3)
    with bt.Plan(t, src, 1, [0.0], 1, 1, 3, 0.8, 6, 300, 300) as p:
        lab, err = p.label_pairs(1e-6, with_errors=True)
        kind, _ = p.pair_info()
        assert len(lab) == p.num_pairs() and set(np.unique(kind)) <= {1, 2, 3}
        for i in range(len(lab)):
            L = lab[i]
            if L <= 12:
                assert err[i, L - 2] <= 1e-6 and np.all(err[i, :L - 2] > 1e-6)
        f = p.pair_features()
        assert f.shape == (len(lab), 8) and np.all(np.isfinite(f))
        assert np.all(f[:, 2 - 2] > 0) and np.all(f[:, 6] == 0.0)


def random_model(bt, seed):
    rng = np.random.default_rng(seed)
    m = bt.DegreeModel(seed=seed)
    p, _, _ = m.get_params()
    m.set_params(p, rng.normal(0, 1, 8), rng.uniform(0.5, 2.0, 8))
    return m, rng


def test_mlp_gradient_and_logits_match_oracle(bt, oracle_mod):
    m, rng = random_model(bt, 4)
    p, mu, sc = m.get_params()
    x = rng.normal(0, 1, (3000, 8)).astype(np.float32)
    y = rng.integers(2, 14, 3000)
    loss, g = m.gradient(x, y)
    lo, go = oracle_mod.mlp_loss_grad(p.astype(np.float64), mu, sc, x, y)
    assert abs(loss - lo) <= 1e-5 * abs(lo)
    assert rel_l2(go, g) <= 1e-4
    d, z = m.predict(x, delta=0.5, with_logits=True)
    zo = oracle_mod.mlp_logits(p.astype(np.float64), mu, sc, x)
    assert np.max(np.abs(z - zo)) <= 1e-4 * (1 + np.max(np.abs(zo)))
    sig = 1 / (1 + np.exp(-zo))
    expect = np.array([max([2] + [3 + k for k in range(10) if s_[k] >= 0.5]) for s_ in sig])
    clear = np.all(np.abs(sig - 0.5) > 1e-4, axis=1)
    assert np.array_equal(d[clear], expect[clear])


def test_training_lowers_loss_and_learns_a_rule(bt):
    rng = np.random.default_rng(8)
    x = rng.uniform(-1, 1, (40000, 8)).astype(np.float32)
    # the least degree grows with the first feature (like r/d)
    y = np.clip(np.round(2 + 10 * (x[:, 0] + 1) / 2), 2, 12).astype(np.int32)
    m = bt.DegreeModel(seed=1)
    losses = m.train(x, y, epochs=15, batch=1024, lr=3e-3, seed=2)
    assert losses[-1] < 0.2 * losses[0]
    d = m.predict(x[:5000], delta=0.5)
    assert np.mean(np.abs(d - y[:5000]) <= 1) > 0.95


def test_trained_network_gain_over_fixed_degree(bt):
    """Regularized Coulomb, Plummer sphere, cluster-particle, theta 0.7,
    degree 10 (the fixed-degree method): a network trained on two other
    Plummer realizations picks per-pair degrees for a third.  It must cut the
    far-field work, keep the error <= 1e-6 against the FP64 direct sum, and
    beat the uniform degree of equal or higher cost."""
    from barytree_amd import learned
    cfg = dict(learned.MODELS["regularized_coulomb_cp"], n=300_000, seeds=(101, 102), leaf=1000,
               max_pairs=10 ** 7)
    x, y = learned.training_set(cfg, log=lambda s: print(s, flush=True))
    m, losses = learned.train(x, y, epochs=10)
    print("losses", losses, flush=True)
    assert losses[-1] < 0.8 * losses[0]
    pts = learned.particles("plummer", 300_000, 103)
    idx = np.arange(0, 300_000, 300)
    with learned.plan_for(cfg, *pts) as p:
        fixed = p.execute()
        st = p.stats()
        t_fixed = st["gpu_total_s"]
        ev_fixed = st["cp_ptwise"]
        direct = p.direct_sample(idx)
        e_fixed = rel_l2(direct, fixed[idx])
        p.use_model(m, 0.5)
        out = p.execute()
        st = p.stats()
        t_learned, ev_learned = st["gpu_total_s"], st["far_ptwise_variable"]
        e_learned = rel_l2(direct, out[idx])
        n = p.num_pairs()
        # the highest uniform degree whose far field costs no more than the learned one
        d_eq, e_eq = None, None
        for d in range(10, 1, -1):
            p.set_pair_degrees(np.full(n, d))
            if p.stats()["far_ptwise_variable"] <= ev_learned:
                d_eq, e_eq = d, rel_l2(direct, p.execute()[idx])
                break
    print(f"fixed degree 10: err {e_fixed:.2e}, far evals {ev_fixed:.3e}, gpu {t_fixed:.4f}s; "
          f"learned: err {e_learned:.2e}, far evals {ev_learned:.3e}, gpu {t_learned:.4f}s, "
          f"mean degree {st['mean_pair_degree']:.2f}; uniform degree {d_eq} at equal cost: "
          f"err {e_eq}", flush=True)
    assert ev_learned < 0.7 * ev_fixed
    assert e_learned <= 1e-6
    assert d_eq is not None and e_learned < e_eq


def test_model_guard_falls_back_outside_the_training_range(bt):
    """Pairs whose features lie outside the model's training range (by more
    than 5% of its width) keep the fixed degree: the count equals the pairs
    outside the shipped range computed here from the plan's features; with a
    range no pair fits, the learned run equals the fixed-degree run to
    rounding; without a range there is no guard."""
    from barytree_amd import learned
    print("case: regularized Coulomb CP Plummer 200k, shipped model with its range, then an empty range",
          flush=True)
    cfg = dict(learned.MODELS["regularized_coulomb_cp"], leaf=1000)
    m = learned.load_model("regularized_coulomb_cp")
    rng_lo, rng_hi = m.feature_range()
    pts = learned.particles("plummer", 200_000, 77)
    with learned.plan_for(cfg, *pts) as p:
        fixed = p.execute()
        n = p.num_pairs()
        f = p.pair_features()
        mg = 0.05 * (rng_hi - rng_lo)
        outside = int(np.sum(np.any(~((f >= rng_lo - mg) & (f <= rng_hi + mg)), axis=1)))
        p.use_model(m, 0.5)
        st = p.stats()
        assert st["variable_degree"] == 1
        print(f"pairs {n}, outside the shipped range {outside}", flush=True)
        assert st["model_fallback_pairs"] == outside
        # a range far from every feature: all pairs keep the fixed degree
        m.set_feature_range(rng_lo + 100.0, rng_hi + 100.0)
        p.use_model(m, 0.5)
        st = p.stats()
        assert st["model_fallback_pairs"] == n
        assert st["mean_pair_degree"] == cfg["degree"]
        assert rel_l2(fixed, p.execute()) <= 1e-12
        # no range: no guard
        m.set_feature_range(None, None)
        assert m.feature_range() is None
        p.use_model(m, 0.5)
        assert p.stats()["model_fallback_pairs"] == 0
        for bad in (0.0, 1.0, -0.5, float("nan")):
            with pytest.raises(RuntimeError, match="delta"):
                p.use_model(m, bad)
