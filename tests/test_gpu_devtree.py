"""Trees built on the GPU (every GPU plan) against the host builder
(BaryTree_Analyze, which the oracle comparisons pin): the interaction counts
of the lists depend on every node's members and box, and must agree exactly;
the results then agree with the oracle at the plain-mode bar."""
import numpy as np
import pytest

from tests.helpers import KP, rel_l2, uniform_cube

pytestmark = pytest.mark.gpu

COUNTS = ("num_pp", "num_pc", "num_cp", "num_cc", "pp_ptwise", "pc_ptwise", "cp_ptwise", "cc_ptwise",
          "pc_replaced", "cp_replaced", "cc_replaced", "source_nodes", "target_nodes")


@pytest.fixture(scope="module")
def bt():
    import barytree_amd
    barytree_amd.lib()
    return barytree_amd


def plummer(n, seed):
    rng = np.random.default_rng(seed)
    r = 0.1 / np.sqrt(rng.uniform(1e-3, 1, n) ** (-2 / 3) - 1)
    v = rng.normal(size=(3, n))
    v /= np.linalg.norm(v, axis=0)
    x, y, z = np.minimum(r, 1.0) * v
    return [x, y, z, rng.uniform(-1, 1, n), np.ones(n)]


@pytest.mark.parametrize("ct", [1, 2, 3])
@pytest.mark.parametrize("same", [True, False])
@pytest.mark.parametrize("leaves", [(60, 60), (40, 90)])
def test_device_trees_give_the_host_lists(bt, ct, same, leaves):
    t, s = uniform_cube(20000, 15000, seed=300 + ct)
    if same:
        t = [a.copy() for a in s[:4]]
    args = (1, [0.5], 1, 1, ct, 0.7, 3, *leaves)
    host = bt.analyze(t, s, *args)
    p = bt.Plan(t, s, *args)
    dev = p.stats()
    p.close()
    for k in COUNTS:
        assert host[k] == dev[k], k


@pytest.mark.parametrize("ct", [1, 2, 3])
def test_clustered_and_lattice_inputs(bt, oracle_mod, ct):
    # Plummer sphere (uneven depths) and a lattice (midpoints on particles)
    s = plummer(30000, 9)
    t = [a.copy() for a in s[:4]]
    args = (3, [0.05], 1, 1, ct, 0.7, 4, 200, 200)
    host = bt.analyze(t, s, *args)
    p = bt.Plan(t, s, *args)
    dev = p.stats()
    got = p.execute()
    p.close()
    for k in COUNTS:
        assert host[k] == dev[k], k
    ref = oracle_mod.treecode(*t, *s, *args)
    assert rel_l2(ref, got) <= 1e-12
    g = np.linspace(-1, 1, 21)
    X, Y, Z = np.meshgrid(g, g, g[:15], indexing="ij")
    n = X.size
    s = [X.ravel(), Y.ravel(), Z.ravel(), np.random.default_rng(2).uniform(-1, 1, n), np.ones(n)]
    t = [a.copy() for a in s[:4]]
    args = (1, [0.5], 1, 1, ct, 0.7, 3, 7, 7)
    host = bt.analyze(t, s, *args)
    p = bt.Plan(t, s, *args)
    dev = p.stats()
    got = p.execute()
    p.close()
    for k in COUNTS:
        assert host[k] == dev[k], k
    ref = oracle_mod.treecode(*t, *s, *args)
    assert rel_l2(ref, got) <= 1e-12


def test_coincident_particles_are_refused(bt):
    # more coincident particles than the leaf size: no split separates them
    n = 400
    s = [np.zeros(n), np.zeros(n), np.zeros(n), np.ones(n), np.ones(n)]
    t = [a.copy() for a in s[:4]]
    with pytest.raises(RuntimeError, match="coincident"):
        bt.solve(t, s, 1, [0.5], 1, 1, 1, 0.7, 3, 100, 100)
    with pytest.raises(RuntimeError, match="coincident"):
        bt.analyze(t, s, 1, [0.5], 1, 1, 1, 0.7, 3, 100, 100)


@pytest.mark.parametrize("k,sg", [(1, 2), (2, 1)])
def test_shared_and_separate_targets_agree(bt, k, sg):
    # targets equal to the sources share the sources' tree and arrays; a
    # separate copy that differs in one bit builds its own tree: the results
    # agree to rounding
    t, s = uniform_cube(30000, 30000, seed=41)
    t = [a.copy() for a in s[:4]]
    args = (k, KP[k][sg], sg, 1, 3, 0.8, 5, 400, 400)
    a = bt.solve(t, s, *args)
    t2 = [a_.copy() for a_ in t]
    t2[3][0] = np.nextafter(t2[3][0], 2.0)  # one target charge differs by one ulp
    b = bt.solve(t2, s, *args)
    assert rel_l2(a, b) <= 1e-12
