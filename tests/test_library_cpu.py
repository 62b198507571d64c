"""The C-ABI library on a machine without a GPU: it loads, exports every
function include/BaryTreeInterface.h declares, keeps the reference's enum
numbering, and its host-side trees and interaction lists (BaryTree_Analyze,
no GPU work) agree with the oracle's."""
import os
import re

import numpy as np
import pytest

from tests.helpers import KP, KP_OTHER, hermite_combos, other_combos, uniform_cube

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "BaryTreeInterface.h")


@pytest.fixture(scope="module")
def bt():
    import barytree_amd
    if not os.path.exists(barytree_amd.LIB_PATH):
        from barytree_amd import build
        build.build()
    barytree_amd.lib()
    return barytree_amd


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\(", text, re.M)
    return sorted(set(n for n in names if n not in ("if", "defined")))


def test_exports_every_declared_symbol(bt):
    names = declared_functions()
    assert "BaryTreeInterface" in names and len(names) >= 10
    lib = bt.lib()
    for n in names:
        assert hasattr(lib, n), n


def test_enum_numbering_matches_reference_abi(bt):
    # src/interface/BaryTreeInterface.h:7-47
    assert [bt.Kernel.COULOMB, bt.Kernel.YUKAWA, bt.Kernel.REGULARIZED_COULOMB, bt.Kernel.USER] == [1, 2, 3, 12]
    assert [bt.Singularity.SKIPPING, bt.Singularity.SUBTRACTION] == [1, 2]
    assert [bt.Approximation.LAGRANGE, bt.Approximation.HERMITE] == [1, 2]
    assert [bt.ComputeType.PARTICLE_CLUSTER, bt.ComputeType.CLUSTER_PARTICLE,
            bt.ComputeType.CLUSTER_CLUSTER] == [1, 2, 3]
    text = open(HEADER).read()
    body = re.search(r"typedef enum KERNEL \{(.*?)\} KERNEL;", text, re.S).group(1)
    assert [s.strip() for s in body.split(",") if s.strip()][:4] == [
        "NO_KERNEL", "COULOMB", "YUKAWA", "REGULARIZED_COULOMB"]


KEYS = ["num_pp", "num_pc", "num_cp", "num_cc", "pp_ptwise", "pc_ptwise", "cp_ptwise",
        "cc_ptwise", "pc_replaced", "cp_replaced", "cc_replaced"]


@pytest.mark.parametrize("ct", [1, 2, 3])
@pytest.mark.parametrize("n,m,deg,msl,mtl,theta", [
    (3000, 2500, 3, 60, 50, 0.7), (12000, 12000, 2, 30, 30, 0.8), (5000, 5000, 4, 50, 10, 0.8)])
def test_host_lists_match_oracle(bt, oracle_mod, ct, n, m, deg, msl, mtl, theta):
    t, s = uniform_cube(n, m, seed=n + ct)
    a = bt.analyze(t, s, 1, [0.5], 1, 1, ct, theta, deg, msl, mtl)
    oracle_mod.treecode(*t, *s, 1, [0.5], 1, 1, ct, theta, deg, msl, mtl)
    o = oracle_mod.last_stats()
    assert {k: a[k] for k in KEYS} == {k: o[k] for k in KEYS}
    assert a["pp_ptwise"] + a["pc_replaced"] + a["cp_replaced"] + a["cc_replaced"] == n * m


@pytest.mark.parametrize("ct,k,s", hermite_combos())
def test_hermite_host_work_matches_oracle(bt, oracle_mod, ct, k, s):
    """Hermite plans build (and check) their work lists on the host; the
    interaction counts are the Lagrange ones (the lists do not depend on the
    approximation)."""
    t, src = uniform_cube(4000, 3000, seed=11 + ct + k)
    a = bt.analyze(t, src, k, KP[k][s], s, 2, ct, 0.7, 3, 60, 50)
    oracle_mod.treecode(*t, *src, k, KP[k][s], s, 2, ct, 0.7, 3, 60, 50)
    o = oracle_mod.last_stats()
    assert {x: a[x] for x in KEYS} == {x: o[x] for x in KEYS}


@pytest.mark.parametrize("ct,k,s", other_combos())
def test_other_kernels_host_work_matches_oracle(bt, oracle_mod, ct, k, s):
    t, src = uniform_cube(3000, 2500, seed=21 + ct + k)
    a = bt.analyze(t, src, k, KP_OTHER[k], s, 1, ct, 0.7, 4, 60, 50)
    oracle_mod.treecode(*t, *src, k, KP_OTHER[k], s, 1, ct, 0.7, 4, 60, 50)
    o = oracle_mod.last_stats()
    assert {x: a[x] for x in KEYS} == {x: o[x] for x in KEYS}


def test_other_kernels_rejected_combinations(bt):
    t, s = uniform_cube(200, 200, seed=3)
    with pytest.raises(RuntimeError, match="PARTICLE-CLUSTER ONLY"):
        bt.analyze(t, s, 5, [2.0, 0.3], 1, 1, 2, 0.7, 3, 10, 10)
    with pytest.raises(RuntimeError, match="SUBTRACTION WITH THIS KERNEL"):
        bt.analyze(t, s, 4, [0.5, 0.1], 2, 1, 2, 0.7, 3, 10, 10)
    with pytest.raises(RuntimeError, match="HERMITE WITH THIS KERNEL"):
        bt.analyze(t, s, 4, [0.5, 0.1], 1, 2, 1, 0.7, 3, 10, 10)
    with pytest.raises(RuntimeError, match="HERMITE WITH THIS KERNEL"):
        bt.analyze(t, s, 8, [1.5], 1, 2, 3, 0.7, 3, 10, 10)
    with pytest.raises(RuntimeError, match="more kernel parameters"):
        bt.analyze(t, s, 9, [2.0], 1, 1, 1, 0.7, 3, 10, 10)


def test_beta_resolution_matches_oracle(bt, oracle_mod):
    t, s = uniform_cube(3000, 3000, seed=4)
    for ct in (1, 2, 3):
        a = bt.analyze(t, s, 2, [0.5], 1, 1, ct, -1.0, -1, 1, 1, 1.0, 0.3)
        oracle_mod.treecode(*t, *s, 2, [0.5], 1, 1, ct, -1.0, -1, 1, 1, 1.0, 0.3)
        o = oracle_mod.last_stats()
        assert a["interp_degree"] == o["interp_degree"] and a["theta"] == o["theta"]
        assert (a["max_per_source_leaf"], a["max_per_target_leaf"]) == (
            o["max_per_source_leaf"], o["max_per_target_leaf"])
        assert {k: a[k] for k in KEYS} == {k: o[k] for k in KEYS}


def test_beta_leaf_sizes_of_the_gpu_build(bt):
    """The GPU build's beta leaf sizes (run_params.c:100-103: 3000/3000)
    on request, the CPU build's by default; theta and degree unchanged."""
    t, s = uniform_cube(20000, 20000, seed=5)
    try:
        cpu = [bt.analyze(t, s, 1, [0.5], 1, 1, ct, -1.0, -1, 1, 1, 1.0, 0.5) for ct in (1, 2, 3)]
        bt.set_beta_leaf_sizes(bt.BETA_LEAVES_GPU)
        gpu = [bt.analyze(t, s, 1, [0.5], 1, 1, ct, -1.0, -1, 1, 1, 1.0, 0.5) for ct in (1, 2, 3)]
        with pytest.raises(RuntimeError):
            bt.set_beta_leaf_sizes(7)
    finally:
        bt.set_beta_leaf_sizes(bt.BETA_LEAVES_CPU)
    for c, g in zip(cpu, gpu):
        assert (g["max_per_source_leaf"], g["max_per_target_leaf"]) == (3000, 3000)
        assert (c["max_per_source_leaf"], c["max_per_target_leaf"]) != (3000, 3000)
        assert (g["theta"], g["interp_degree"]) == (c["theta"], c["interp_degree"])
    again = bt.analyze(t, s, 1, [0.5], 1, 1, 3, -1.0, -1, 1, 1, 1.0, 0.5)
    assert (again["max_per_source_leaf"], again["max_per_target_leaf"]) == (50, 50)


def test_unsupported_combinations_are_errors(bt):
    t, s = uniform_cube(100, 100, seed=1)
    # Hermite: no CC, no CP with subtraction (interaction_compute_cc.c:146,
    # interaction_compute_downpass.c:126-128)
    with pytest.raises(RuntimeError, match="CC HERMITE"):
        bt.analyze(t, s, 1, [0.5], 1, 2, 3, 0.7, 3, 10, 10)
    with pytest.raises(RuntimeError, match="CP HERMITE SS"):
        bt.analyze(t, s, 1, [0.5], 2, 2, 2, 0.7, 3, 10, 10)
    with pytest.raises(RuntimeError, match="REGULARIZED"):
        bt.analyze(t, s, 3, [0.3, 0.05], 2, 1, 2, 0.7, 3, 10, 10)
    with pytest.raises(RuntimeError, match="KERNEL"):
        bt.analyze(t, s, 6, [0.5], 1, 1, 1, 0.7, 3, 10, 10)
    with pytest.raises(RuntimeError, match="degree"):
        bt.analyze(t, s, 1, [0.5], 1, 1, 1, 0.7, 13, 10, 10)
    with pytest.raises(RuntimeError, match="positive"):
        bt.analyze([a[:0] for a in t], s, 1, [0.5], 1, 1, 1, 0.7, 3, 10, 10)


def _systems(sizes, seed):
    out = []
    for i, n in enumerate(sizes):
        t, s = uniform_cube(n, max(1, (3 * n) // 4), seed=seed + i)
        out.append((t, s))
    return out


@pytest.mark.parametrize("ct", [1, 2, 3])
def test_batched_lists_are_the_sum_of_single_systems(bt, oracle_mod, ct):
    # one plan over several systems: every system gets its own trees and lists
    sizes = [1, 40, 900, 2500, 3000]
    leaf = [10, 8, 30, 60, 45]
    systems = _systems(sizes, 300 + ct)
    a = bt.analyze_batched(systems, 1, [0.5], 1, 1, ct, 0.7, 3, max_source_leaf=leaf,
                           max_target_leaf=leaf)
    tot = {k: 0 for k in KEYS}
    for (t, s), lf in zip(systems, leaf):
        oracle_mod.treecode(*t, *s, 1, [0.5], 1, 1, ct, 0.7, 3, lf, lf)
        o = oracle_mod.last_stats()
        for k in KEYS:
            tot[k] += o[k]
    assert {k: a[k] for k in KEYS} == tot
    assert a["num_systems"] == len(sizes)
    pairs = sum(len(t[0]) * len(s[0]) for t, s in systems)
    assert a["pp_ptwise"] + a["pc_replaced"] + a["cp_replaced"] + a["cc_replaced"] == pairs


def test_batched_schedule_classes_pick_leaf_sizes(bt, oracle_mod):
    # size class c holds systems with 2^c <= max(M, N) < 2^(c+1); an unset
    # class takes the nearest set one
    systems = _systems([700, 3000, 20000], 41)
    sc = bt.Schedule.make({9: 64, 14: 500})
    a = bt.analyze_batched(systems, 1, [0.5], 1, 1, 1, 0.7, 4, schedule=sc)
    tot = {k: 0 for k in KEYS}
    for (t, s), lf in zip(systems, (64, 64, 500)):   # 3000 is class 11: nearest set is 9
        oracle_mod.treecode(*t, *s, 1, [0.5], 1, 1, 1, 0.7, 4, lf, lf)
        o = oracle_mod.last_stats()
        for k in KEYS:
            tot[k] += o[k]
    assert {k: a[k] for k in KEYS} == tot


def test_batched_argument_errors(bt):
    systems = _systems([100, 200], 5)
    args = (1, [0.5], 1, 1, 1, 0.7, 3)
    with pytest.raises(RuntimeError, match="leaf sizes or a schedule"):
        bt.analyze_batched(systems, *args)
    with pytest.raises(RuntimeError, match="targets_per_thread"):
        bt.analyze_batched(systems, *args, schedule=bt.Schedule.make({}, targets_per_thread=3))
    with pytest.raises(RuntimeError, match="CC HERMITE"):
        bt.analyze_batched(systems, 1, [0.5], 1, 2, 3, 0.7, 3, max_source_leaf=10,
                           max_target_leaf=10)
    with pytest.raises(RuntimeError, match="leaf sizes must be positive"):
        bt.analyze_batched(systems, *args, max_source_leaf=[10, 0], max_target_leaf=10)
    with pytest.raises(RuntimeError, match="at least one target"):
        bt.analyze_batched([systems[0], ([a[:0] for a in systems[1][0]], systems[1][1])], *args,
                           max_source_leaf=10, max_target_leaf=10)
    # solve entry points reject the same inputs before any GPU work
    with pytest.raises(RuntimeError, match="at least one target"):
        bt.solve_batched([systems[0], ([a[:0] for a in systems[1][0]], systems[1][1])], *args,
                         max_source_leaf=10, max_target_leaf=10)
    with pytest.raises(RuntimeError, match="numSystems"):
        bt.solve_batched([], *args, max_source_leaf=10, max_target_leaf=10)


@pytest.mark.parametrize("ct", [1, 2, 3])
def test_stats_layout_far_pair_count(bt, ct):
    # the last fields of BaryTreeStats: far-field pairs = entries of the
    # approximation lists (PC, CP and CC)
    t, s = uniform_cube(4000, 3000, seed=12)
    a = bt.analyze(t, s, 1, [0.5], 1, 1, ct, 0.7, 3, 40, 40)
    assert a["num_far_pairs"] == a["num_pc"] + a["num_cp"] + a["num_cc"] > 0
    assert a["variable_degree"] == 0 and a["num_systems"] == 1


@pytest.mark.parametrize("ct,k,s", [(1, 1, 1), (1, 3, 2), (2, 2, 1), (2, 3, 1), (3, 1, 2), (3, 2, 1)])
def test_variable_degree_work_on_host(bt, ct, k, s):
    # the per-degree work (slots, parts, one near and one far work list,
    # split rows, downpass tiles) is built and every index checked on the host
    from tests.helpers import KP
    t, src = uniform_cube(6000, 5000, seed=70 + ct)
    args = (k, KP[k][s], s, 1, ct, 0.7, 5, 60, 60)
    st, f = bt.analyze_degrees(t, src, *args)
    n = st["num_far_pairs"]
    assert n == st["num_pc"] + st["num_cp"] + st["num_cc"] > 0
    assert f.shape == (n, 8) and np.all(np.isfinite(f)) and np.all(f[:, 0] >= 0)
    kinds = f[:, 4] + 2 * f[:, 5]            # 1 PC, 2 CP, 0 CC
    assert set(np.unique(kinds)) <= ({1} if ct == 1 else {2} if ct == 2 else {0, 1, 2})
    # every pair at the plan's degree: the fixed method's far-field work
    st5, _ = bt.analyze_degrees(t, src, *args, degrees=np.full(n, 5))
    assert st5["variable_degree"] == 1 and st5["mean_pair_degree"] == 5.0
    assert st5["far_ptwise_variable"] == st["pc_ptwise"] + st["cp_ptwise"] + st["cc_ptwise"]
    deg = np.random.default_rng(ct).integers(1, 13, n)
    stv, _ = bt.analyze_degrees(t, src, *args, degrees=deg)
    assert np.isclose(stv["mean_pair_degree"], deg.mean())
    with pytest.raises(RuntimeError, match="degree outside"):
        bt.analyze_degrees(t, src, *args, degrees=np.full(n, 13))


def test_analyze_degrees_checks_the_pair_count_of_every_buffer(bt):
    # a C caller that passes a features buffer on the counting call (numPairs
    # 0) or with a stale count gets an error and an untouched buffer
    import ctypes
    from barytree_amd.interface import Stats, _kp, _ptr
    t, src = uniform_cube(3000, 3000, seed=31)
    kp = _kp([0.5])
    args = (ctypes.c_int(3000), ctypes.c_int(3000), *[_ptr(a) for a in t[:3]], *[_ptr(a) for a in src[:3]],
            ctypes.c_int(1), ctypes.c_int(1), _ptr(kp), ctypes.c_int(1), ctypes.c_int(1), ctypes.c_int(3),
            ctypes.c_double(0.7), ctypes.c_int(4), ctypes.c_int(50), ctypes.c_int(50), ctypes.c_double(1.0),
            ctypes.c_double(-1.0))
    s = Stats()
    assert bt.lib().BaryTree_AnalyzeDegrees(*args, 0, None, None, ctypes.byref(s)) == 0
    n = s.num_far_pairs
    assert n > 1
    fptr = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))  # noqa: E731
    for count in (0, n - 1, n + 1):
        f = np.full((n + 2) * 8, -7.0, dtype=np.float32)
        rc = bt.lib().BaryTree_AnalyzeDegrees(*args, ctypes.c_int(count), None, fptr(f), ctypes.byref(s))
        assert rc != 0 and "numPairs" in bt.last_error()
        assert np.all(f == -7.0)
    f = np.full((n + 2) * 8, -7.0, dtype=np.float32)
    assert bt.lib().BaryTree_AnalyzeDegrees(*args, ctypes.c_int(n), None, fptr(f), ctypes.byref(s)) == 0
    assert np.all(f[n * 8:] == -7.0) and np.all(np.isfinite(f[:n * 8]))


def test_symmetric_near_field_eligibility(bt):
    """The symmetric near field covers every near particle pair of a system
    (Coulomb, Yukawa, or regularized Coulomb with skipping) whose targets are
    its sources in the same tree order, and nothing otherwise: pairs of
    leaves in each other's near list once for both particles, the rest
    (each leaf with itself among them) once for the target."""
    rng = np.random.default_rng(8)
    n = 60000
    x, y, z, q = (rng.uniform(-1, 1, n) for _ in range(4))
    w = np.ones(n)
    same = (x.copy(), y.copy(), z.copy(), q.copy())
    src = (x, y, z, q, w)
    for k, s in ((1, 2), (2, 2), (1, 1), (2, 1), (3, 1)):
        a = bt.analyze(same, src, k, [0.5], s, 1, 3, 0.8, 6, 800, 800)
        assert a["sym_ptwise"] == a["pp_ptwise"], a
        # most pairs run once for both particles; the one-sided ones
        # (ordered pairs evaluated once: 2 evals - ptwise) include every
        # particle with itself, and every particle has a slot of its own leaf
        assert a["pp_ptwise"] // 2 < a["sym_evals"] < 0.6 * a["pp_ptwise"], a
        assert 2 * a["sym_evals"] - a["sym_ptwise"] >= n, a
        assert n <= a["sym_slots"] < a["sym_evals"], a
    for ct in (1, 2):  # particle-cluster and cluster-particle with equal leaf sizes
        a = bt.analyze(same, src, 1, [0.5], 1, 1, ct, 0.8, 6, 800, 800)
        assert a["sym_ptwise"] == a["pp_ptwise"] and a["sym_evals"] < 0.6 * a["pp_ptwise"], (ct, a)
    # the approximation concerns only the far field: Hermite plans too
    herm = bt.analyze(same, src, 1, [0.5], 1, 2, 2, 0.8, 6, 800, 800)  # CP Hermite
    lag = bt.analyze(same, src, 1, [0.5], 1, 1, 2, 0.8, 6, 800, 800)
    assert herm["sym_ptwise"] == herm["pp_ptwise"] == lag["sym_ptwise"]
    assert herm["sym_slots"] == lag["sym_slots"] > 0
    assert bt.analyze(same, src, 1, [0.5], 1, 1, 1, 0.8, 6, 800, 100)["sym_ptwise"] == 0  # PC, other order
    assert bt.analyze(same, src, 8, [1.5], 1, 1, 3, 0.8, 6, 800, 800)["sym_ptwise"] == 0  # sin-over-r
    # different trees
    assert bt.analyze(same, src, 1, [0.5], 2, 1, 3, 0.8, 6, 800, 100)["sym_ptwise"] == 0
    other = (x.copy(), y.copy(), z.copy(), q.copy())
    other[0][17] += 1e-12
    assert bt.analyze(other, src, 1, [0.5], 2, 1, 3, 0.8, 6, 800, 800)["sym_ptwise"] == 0  # positions


def test_model_feature_range_round_trip(bt, tmp_path):
    """The training range is saved and loaded with the model; a file of the
    earlier layout (no range) loads without one; bad ranges are refused."""
    import struct
    m = bt.DegreeModel(seed=3)
    assert m.feature_range() is None
    lo = np.linspace(-1, 0, 8).astype(np.float32)
    hi = lo + np.arange(1, 9, dtype=np.float32)
    m.set_feature_range(lo, hi)
    path = tmp_path / "m.bin"
    m.save(path)
    r = bt.DegreeModel(path=path).feature_range()
    assert r is not None and np.array_equal(r[0], lo) and np.array_equal(r[1], hi)
    with pytest.raises(RuntimeError, match="lo <= hi"):
        m.set_feature_range(hi, lo)
    # the earlier layout: magic ...01, header, steps, mean, scale, params
    p, mean, scale = m.get_params()
    old = tmp_path / "old.bin"
    with open(old, "wb") as f:
        f.write(b"BTDMLP01" + struct.pack("4i", 8, 64, 10, len(p)) + struct.pack("i", 0))
        f.write(mean.tobytes() + scale.tobytes() + p.tobytes())
    o = bt.DegreeModel(path=old)
    assert o.feature_range() is None
    assert np.array_equal(o.get_params()[0], p)
