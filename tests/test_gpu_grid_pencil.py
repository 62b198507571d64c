"""The far field of mixed executions on target grids as tensor grids
(k_grid_pencil): every grid size the degrees 1..12 give (2..13 points per
dimension: one thread per pencil, 1 to 64 groups of threads sharing the
staged sources), in one launch when pairs carry
their own degrees, for every kernel and singularity the mixed path serves,
in cluster-particle and cluster-cluster plans.  The mixed result is checked
against the FP64 execution of the same plan (the FP32 pairs' bar, 1e-6)."""
import numpy as np
import pytest

from tests.helpers import KP, rel_l2, uniform_cube

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def bt():
    import barytree_amd
    barytree_amd.lib()
    return barytree_amd


@pytest.mark.parametrize("deg", [1, 3, 4, 7, 10, 11, 12])
@pytest.mark.parametrize("ct,k,s", [(2, 1, 1), (2, 1, 2), (2, 2, 2), (2, 3, 1), (3, 1, 2), (3, 2, 1), (3, 3, 1)])
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
