"""BaryTreePlan_DirectSampleNoAlloc: the FP64 direct sum of sampled targets
with its device buffers in the library's fixed scratch array (no device
allocation).  Called through the C ABI and compared with
BaryTreePlan_DirectSample (the same kernels and chunking, so bit for bit) on
two calls in a row, the second with a smaller input, on a batched plan with
repeated and out-of-order indices, and at the size limit.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_ip = ctypes.POINTER(ctypes.c_int)
_dp = ctypes.POINTER(ctypes.c_double)
MAX_SAMPLE = 8192


@pytest.fixture(scope="module")
def bt():
    import barytree_amd
    barytree_amd.lib()
    return barytree_amd


def noalloc(bt, plan, idx):
    idx = np.ascontiguousarray(idx, dtype=np.int32)
    out = np.full(len(idx), np.nan)
    rc = bt.lib().BaryTreePlan_DirectSampleNoAlloc(plan._h, ctypes.c_int(len(idx)), idx.ctypes.data_as(_ip),
                                                   out.ctypes.data_as(_dp))
    return rc, out


def system(n, seed):
    rng = np.random.default_rng(seed)
    x, y, z, q = (rng.uniform(-1, 1, n) for _ in range(4))
    w = rng.uniform(0.5, 1.5, n)
    return (x.copy(), y.copy(), z.copy(), q.copy()), (x, y, z, q, w)


def test_noalloc_equals_direct_sample_two_calls(bt):
    """A large call, then a smaller one (fewer targets, a smaller plan and
    another kernel): each equals BaryTreePlan_DirectSample bit for bit, so the
    second sees nothing the first left in the scratch."""
    print("case: Coulomb-SS N=200k m=6000, then Yukawa N=30k m=500", flush=True)
    t, s = system(200_000, 1)
    with bt.Plan(t, s, 1, [0.5], 2, 1, 3, 0.8, 6, 2000, 2000) as plan:
        idx = np.random.default_rng(2).integers(0, 200_000, 6000)
        rc, got = noalloc(bt, plan, idx)
        assert rc == 0, bt.last_error()
        ref = plan.direct_sample(idx)
        assert np.array_equal(got, ref)
    t, s = system(30_000, 3)
    with bt.Plan(t, s, 2, [0.5], 1, 1, 1, 0.7, 5, 500, 500) as plan:
        idx = np.arange(0, 30_000, 60)[:500]
        rc, got = noalloc(bt, plan, idx)
        assert rc == 0, bt.last_error()
        assert np.array_equal(got, plan.direct_sample(idx))


def test_noalloc_batched_repeated_indices_and_limit(bt):
    print("case: batched 3 systems, repeated indices; m = 8192 and 8193", flush=True)
    systems = [system(n, 10 + i) for i, n in enumerate((5000, 40_000, 12_000))]
    with bt.Plan.batched(systems, 1, [0.5], 1, 1, 1, 0.7, 5, 800, 800) as plan:
        total = 57_000
        rng = np.random.default_rng(4)
        idx = rng.integers(0, total, 3000)
        idx[:10] = idx[10:20]  # repeats, in no order
        rc, got = noalloc(bt, plan, idx)
        assert rc == 0, bt.last_error()
        assert np.array_equal(got, plan.direct_sample(idx))
        idx = rng.integers(0, total, MAX_SAMPLE)
        rc, got = noalloc(bt, plan, idx)
        assert rc == 0, bt.last_error()
        assert np.array_equal(got, plan.direct_sample(idx))
        rc, _ = noalloc(bt, plan, rng.integers(0, total, MAX_SAMPLE + 1))
        assert rc != 0 and "8192" in bt.last_error()
        rc, _ = noalloc(bt, plan, np.array([total]))
        assert rc != 0 and "range" in bt.last_error()
