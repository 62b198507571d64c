"""Solves from particles already in GPU memory (BaryTree_SolveDevice,
BaryTree_SolveLearnedDevice): every array is a device pointer, and the
results equal the host-array entry points' bit for bit -- the same trees,
lists and kernels; only the copies differ."""
import numpy as np
import pytest

from tests.helpers import KP, uniform_cube

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def bt():
    import barytree_amd
    barytree_amd.lib()
    return barytree_amd


def dev(a):
    import barytree_amd
    return None if a is None else barytree_amd.DeviceBuffer.from_numpy(a)


@pytest.mark.parametrize("ct,k,s,same", [(1, 1, 1, False), (2, 2, 1, False), (3, 1, 2, False),
                                         (3, 1, 2, True), (3, 2, 1, True), (1, 3, 2, False)])
def test_device_solve_equals_host_solve(bt, ct, k, s, same):
    print(f"case ct={ct} kernel={k} singularity={s} targets=sources={same}", flush=True)
    t, src = uniform_cube(30000, 25000 if not same else 30000, seed=100 + ct + k + s)
    if same:
        t = [a.copy() for a in src[:4]]
    args = (k, KP[k][s], s, 1, ct, 0.8, 6, 400, 400)
    for precision in (bt.PRECISION_FP64, bt.PRECISION_MIXED):
        host = bt.solve(t, src, *args, precision=precision)
        dt, ds = [dev(a) for a in t], [dev(a) for a in src]
        out = dev(np.full(len(t[0]), np.nan))
        bt.solve_device(dt, ds, out, *args, precision=precision)
        got = out.to_numpy()
        assert np.array_equal(got, host), (precision, np.max(np.abs(got - host)))
        if same:
            # the same tensors as targets and sources (aliased pointers)
            out2 = dev(np.zeros(len(t[0])))
            bt.solve_device(ds[:4], ds, out2, *args, precision=precision)
            assert np.array_equal(out2.to_numpy(), host)


def test_device_learned_solve_equals_host(bt):
    from barytree_amd import learned
    cfg = learned.MODELS["regularized_coulomb_cp"]
    pts = learned.particles("plummer", 200_000, 9)
    t = (pts[0].copy(), pts[1].copy(), pts[2].copy(), pts[3].copy())
    args = (cfg["kernel"], cfg["kp"], cfg["sing"], 1, cfg["ct"], cfg["theta"], cfg["degree"], 1000, 1000)
    m = learned.load_model("regularized_coulomb_cp")
    host = bt.solve_learned(t, pts, *args, model=m, delta=0.5, precision=bt.PRECISION_MIXED)
    out = dev(np.zeros(len(t[0])))
    bt.solve_device([dev(a) for a in t], [dev(a) for a in pts], out, *args, precision=bt.PRECISION_MIXED,
                    model=m, delta=0.5)
    assert np.array_equal(out.to_numpy(), host)
    assert bt.last_stats()["variable_degree"] == 1


def test_device_solve_errors(bt):
    t, src = uniform_cube(1000, 1000, seed=3)
    out = dev(np.zeros(1000))
    with pytest.raises(ValueError):
        bt.solve_device([dev(a) for a in t], [dev(a) for a in src], np.zeros(1000), 1, [0.5], 1, 1, 1, 0.7, 4,
                        100, 100)
    with pytest.raises(RuntimeError, match="weights"):
        bt.solve_device([dev(a) for a in t], [dev(a) for a in src[:4]] + [None], out, 1, [0.5], 2, 1, 1, 0.7, 4,
                        100, 100)


def test_device_memory_cache(bt):
    """Destroyed plans leave their buffers in the library's per-GPU cache;
    the next solve of the same size takes them back (results unchanged), a
    limit of 0 keeps nothing, and emptying the cache frees everything."""
    print("case device-memory cache, N=200k PC Coulomb", flush=True)
    t, s = uniform_cube(200_000, 200_000, seed=91)
    args = (1, [0.5], 1, 1, 1, 0.7, 6, 2000, 2000)
    bt.empty_device_cache()
    first = bt.solve(t, s, *args, precision=bt.PRECISION_MIXED)
    cached = bt.cached_device_bytes()
    assert cached > 8 * 200_000 * 4
    second = bt.solve(t, s, *args, precision=bt.PRECISION_MIXED)
    assert np.array_equal(first, second)
    assert bt.cached_device_bytes() == cached
    bt.set_device_cache_limit(0)
    assert bt.cached_device_bytes() == 0
    third = bt.solve(t, s, *args, precision=bt.PRECISION_MIXED)
    assert np.array_equal(first, third)
    assert bt.cached_device_bytes() == 0
    bt.set_device_cache_limit(32 << 30)
    bt.solve(t, s, *args)
    assert bt.cached_device_bytes() > 0
    assert bt.empty_device_cache() > 0
    assert bt.cached_device_bytes() == 0
