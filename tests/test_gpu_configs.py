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
