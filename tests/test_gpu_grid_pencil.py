"""The far field on target grids as tensor grids (k_grid_pencil for mixed
executions, k_grid_pencil64 for FP64 ones): every grid size the degrees
1..12 give (2..13 points per dimension: one thread per pencil, 1 to 64
groups of threads sharing the staged sources), in one launch when pairs
carry their own degrees, for every kernel and singularity the launch
serves, in cluster-particle and cluster-cluster plans.  FP64 results are
checked against the oracle (the plain-mode bar, 1e-12), mixed ones against
the FP64 execution of the same plan (the FP32 pairs' bar, 1e-6)."""
import numpy as np
import pytest

from tests.helpers import KP, rel_l2, uniform_cube

COMBOS = [(2, 1, 1), (2, 1, 2), (2, 2, 2), (2, 3, 1), (3, 1, 2), (3, 2, 1), (3, 3, 1)]

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def bt():
    import barytree_amd
    barytree_amd.lib()
    return barytree_amd


@pytest.mark.parametrize("deg", [1, 3, 4, 7, 10, 11, 12])
@pytest.mark.parametrize("ct,k,s", COMBOS)
def test_pencil_grids_every_degree(bt, ct, k, s, deg):
    print(f"case ct={ct} kernel={k} singularity={s} degree={deg}", flush=True)
    t, src = uniform_cube(30000, 24000, seed=70 + ct + 3 * k + s)
    args = (k, KP[k][s], s, 1, ct, 0.7, deg, 400, 400)
    with bt.Plan(t, src, *args) as p:
        fp64 = p.execute()
        p.set_precision(bt.PRECISION_MIXED)
        mixed = p.execute()
        assert p.stats()["num_cc"] + p.stats()["num_cp"] > 0
    err = rel_l2(fp64, mixed)
    print(f"mixed vs FP64 {err:.2e}", flush=True)
    assert err < 1e-6


@pytest.mark.parametrize("ct,k,s", [(2, 3, 1), (3, 1, 2), (3, 2, 1), (2, 1, 2)])
def test_pencil_grids_mixed_degrees_in_one_launch(bt, ct, k, s):
    print(f"case ct={ct} kernel={k} singularity={s} per-pair degrees 1..12", flush=True)
    t, src = uniform_cube(30000, 30000, seed=90 + ct + k + s)
    args = (k, KP[k][s], s, 1, ct, 0.7, 8, 400, 400)
    with bt.Plan(t, src, *args) as p:
        rng = np.random.default_rng(ct + 10 * k + s)
        p.set_pair_degrees(rng.integers(1, 13, p.num_pairs()))
        fp64 = p.execute()
        p.set_precision(bt.PRECISION_MIXED)
        mixed = p.execute()
    err = rel_l2(fp64, mixed)
    print(f"mixed vs FP64 {err:.2e}", flush=True)
    assert err < 1e-6


@pytest.mark.parametrize("deg", [1, 2, 5, 8, 11, 12])
@pytest.mark.parametrize("ct,k,s", COMBOS)
def test_pencil_grids_fp64_every_degree_match_oracle(bt, oracle_mod, ct, k, s, deg):
    print(f"case FP64 ct={ct} kernel={k} singularity={s} degree={deg}", flush=True)
    t, src = uniform_cube(6000, 5000, seed=110 + ct + 3 * k + s + deg)
    # size check off: far-field pairs at every degree (CC: fewer, larger leaves)
    leaf = 250 if ct == 3 else 110
    args = (k, KP[k][s], s, 1, ct, 0.7, deg, leaf, leaf)
    ref = oracle_mod.treecode(*t, *src, *args, size_check=0.0)
    st = oracle_mod.last_stats()
    assert st["num_cc"] + st["num_cp"] > 0
    with bt.Plan(t, src, *args, size_check=0.0) as p:
        got = p.execute()
        again = p.execute()
    err = rel_l2(ref, got)
    print(f"FP64 vs oracle {err:.2e}", flush=True)
    assert err <= 1e-12
    assert np.array_equal(got, again)


@pytest.mark.parametrize("ct,k,s", [(2, 3, 1), (3, 1, 2), (2, 2, 1)])
def test_pencil_grids_fp64_mixed_degrees_in_one_launch(bt, oracle_mod, ct, k, s):
    """Per-pair degrees 1..12 (one FP64 launch over every grid size) against
    the oracle, at a plan degree (10) that sends the low degrees through the
    change of degree and the high ones through their own grids."""
    print(f"case FP64 ct={ct} kernel={k} singularity={s} per-pair degrees 1..12", flush=True)
    t, src = uniform_cube(6000, 6000, seed=130 + ct + k + s)
    leaf = 250 if ct == 3 else 110
    args = (k, KP[k][s], s, 1, ct, 0.7, 10, leaf, leaf)
    with bt.Plan(t, src, *args, size_check=0.0) as p:
        deg = np.random.default_rng(ct + 5 * k + s).integers(1, 13, p.num_pairs())
        p.set_pair_degrees(deg)
        got = p.execute()
    ref = oracle_mod.treecode_vardeg(*t, *src, *args, deg, size_check=0.0)
    err = rel_l2(ref, got)
    print(f"FP64 vs oracle {err:.2e}", flush=True)
    assert err <= 1e-12
