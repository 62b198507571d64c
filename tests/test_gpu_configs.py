"""The benchmark configurations at parity-test size (against the oracle,
<= 1e-12) and at full size (against the FP64 direct sum at 1000 sampled
targets, the error metric of examples/support_fns.c:694-730).
Distributions: barytree_amd/workloads.py."""
import numpy as np
import pytest

from barytree_amd import workloads as W
from tests.helpers import rel_l2

pytestmark = pytest.mark.gpu

# name: (generator, n, kernel, params, singularity, compute, theta, degree, leaf)
# kernel, singularity, compute type, theta and distribution of each
# configuration; degree and leaf size reduced so that clusters are accepted at
# sizes the oracle finishes in seconds
SMALL = {
    "C1": (W.uniform_cube, 20000, 1, [0.5], 1, 1, 0.7, 5, 300),
    "C2": (W.uniform_cube, 20000, 2, [0.5], 1, 3, 0.8, 4, 300),
    "C3": (W.gaussian_mixture, 8000, 1, [0.5], 1, 1, 0.7, 6, 300),
    "C4": (W.plummer, 20000, 3, [0.01], 1, 2, 0.7, 5, 300),
    "C5": (W.uniform_cube, 20000, 1, [0.5], 2, 3, 0.8, 4, 300),
}


@pytest.fixture(scope="module")
def bt():
    import barytree_amd
    barytree_amd.lib()
    return barytree_amd


@pytest.mark.parametrize("name", list(SMALL))
def test_config_shapes_against_oracle(bt, oracle_mod, name):
    gen, n, k, kp, s, ct, theta, deg, leaf = SMALL[name]
    x, y, z, q, w = gen(n, 7)
    t, src = (x.copy(), y.copy(), z.copy(), q.copy()), (x, y, z, q, w)
    args = (k, kp, s, 1, ct, theta, deg, leaf, leaf)
    ref = oracle_mod.treecode(*t, *src, *args)
    got = bt.solve(t, src, *args)
    assert rel_l2(ref, got) <= 1e-12


FULL = {
    "C1": (W.uniform_cube, 100_000, 1, [0.5], 1, 1, 0.7, 8, 2000, 1e-6),
    "C2": (W.uniform_cube, 1_000_000, 2, [0.5], 1, 3, 0.8, 8, 4000, 1e-5),
    "C4": (W.plummer, 1_000_000, 3, [0.01], 1, 2, 0.7, 10, 4000, 1e-6),
}


@pytest.mark.parametrize("name", list(FULL))
def test_full_size_against_direct(bt, name):
    gen, n, k, kp, s, ct, theta, deg, leaf, tol = FULL[name]
    x, y, z, q, w = gen(n, 11)
    args = (k, kp, s, 1, ct, theta, deg, leaf, leaf)
    idx = np.arange(0, n, n // 1000)[:1000]
    with bt.Plan((x.copy(), y.copy(), z.copy(), q.copy()), (x, y, z, q, w), *args) as p:
        out = p.execute()
        d = p.direct_sample(idx)
        p.set_precision(bt.PRECISION_MIXED)
        mixed = p.execute()
    e64, e32 = rel_l2(d, out[idx]), rel_l2(d, mixed[idx])
    print(f"{name}: fp64 relL2 {e64:.2e}, mixed relL2 {e32:.2e}")
    assert e64 < tol and e32 < tol


# ---- full-size configurations with their own forms and properties ---------

def test_c5_full_size_learned_mixed(bt):
    """16M Coulomb with subtraction, CC, theta 0.8, degree 10, leaf 8000:
    the learned-degree mixed form against the FP64 direct sum (<= 1e-6) and
    exactly linear in the charges (charges and weights enter every term
    linearly; doubling them is exact in binary floating point and the
    per-pair degrees depend only on geometry)."""
    from barytree_amd import learned
    n = 16_000_000
    print(f"case C5 N={n} mixed + learned", flush=True)
    x, y, z, q, w = W.uniform_cube(n, 5)
    args = (1, [0.5], 2, 1, 3, 0.8, 10, 8000, 8000)
    model = learned.load_model("coulomb_cc")
    idx = np.arange(0, n, n // 1000)[:1000]
    with bt.Plan((x, y, z, q), (x, y, z, q, w), *args, precision=bt.PRECISION_MIXED) as p:
        p.use_model(model, 0.5)
        out = p.execute()
        d = p.direct_sample(idx)
        st = p.stats()
    err = rel_l2(d, out[idx])
    print(f"C5 learned mixed relL2 {err:.2e}, mean degree {st['mean_pair_degree']:.2f}", flush=True)
    assert err <= 1e-6
    assert 2 <= st["mean_pair_degree"] < 10
    q2 = 2.0 * q
    with bt.Plan((x, y, z, q2), (x, y, z, q2, w), *args, precision=bt.PRECISION_MIXED) as p:
        p.use_model(model, 0.5)
        out2 = p.execute()
    assert np.array_equal(out2, 2.0 * out)


def test_c4_full_size_learned(bt):
    """4M Plummer, regularized Coulomb eps 0.01, CP, theta 0.7, degree 10,
    leaf 2000: fixed-degree FP64 and learned-degree mixed forms against the
    FP64 direct sum (<= 1e-6); the learned degrees evaluate fewer far-field
    terms than degree 10."""
    from barytree_amd import learned
    cfg = learned.MODELS["regularized_coulomb_cp"]
    n = 4_000_000
    print(f"case C4 N={n} fixed FP64 and learned mixed", flush=True)
    pts = learned.particles("plummer", n, 4)
    model = learned.load_model("regularized_coulomb_cp")
    idx = np.arange(0, n, n // 1000)[:1000]
    with learned.plan_for(cfg, *pts) as p:
        fixed = p.execute()
        d = p.direct_sample(idx)
        cp_fixed = p.stats()["cp_ptwise"]
        p.use_model(model, 0.5)
        p.set_precision(bt.PRECISION_MIXED)
        lm = p.execute()
        st = p.stats()
    e_fixed, e_lm = rel_l2(d, fixed[idx]), rel_l2(d, lm[idx])
    print(f"C4 fixed FP64 {e_fixed:.2e}, learned mixed {e_lm:.2e}, far terms "
          f"{st['far_ptwise_variable']:.3e} vs {cp_fixed:.3e}", flush=True)
    assert e_fixed <= 1e-6 and e_lm <= 1e-6
    assert 0 < st["far_ptwise_variable"] < cp_fixed


def test_c3_full_size_batched(bt):
    """256 Gaussian-mixture Coulomb systems, N log-uniform in [1k, 64k], PC,
    theta 0.7, degree 6, one batched call: each system agrees with its own
    single-system plan (one-directional near field, the batched form's;
    checked on every 16th system) to 1e-14 -- not bit for bit, since a plan
    with fewer than 1024 source leaves anterpolates its leaves in parts and
    adds the parts' rows, another order of the same sums -- and the batch
    matches the FP64 direct sum at 1000 targets spread over all systems."""
    sizes = W.batched_sizes(256, 3, nmin=1000, nmax=64000)
    print(f"case C3: 256 systems, {int(sizes.sum())} particles", flush=True)
    systems = []
    for i, n in enumerate(sizes):
        x, y, z, q, w = W.gaussian_mixture(int(n), 3000 + i)
        systems.append(((x.copy(), y.copy(), z.copy(), q.copy()), (x, y, z, q, w)))
    kargs = (1, [0.0], 1, 1, 1, 0.7, 6)
    leaf = [1000 if n < 8000 else 2000 for n in sizes]
    got = bt.solve_batched(systems, *kargs, max_source_leaf=leaf, max_target_leaf=leaf)
    assert bt.last_stats()["num_systems"] == 256
    for i in range(0, 256, 16):
        t, s = systems[i]
        with bt.Plan(t, s, *kargs, leaf[i], leaf[i]) as p:
            p.set_symmetric(False)
            one = p.execute()
        assert rel_l2(one, got[i]) <= 1e-14, (i, rel_l2(one, got[i]))
    total = int(sizes.sum())
    idx = np.linspace(0, total - 1, 1000).astype(np.int64)
    with bt.Plan.batched(systems, *kargs, max_source_leaf=leaf, max_target_leaf=leaf) as p:
        d = p.direct_sample(idx)
    err = rel_l2(d, np.concatenate(got)[idx])
    print(f"C3 batched relL2 vs direct {err:.2e}", flush=True)
    assert err <= 1e-5
