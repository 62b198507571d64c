"""The C-ABI library on a machine without a GPU: it loads, exports every
function include/BaryTreeInterface.h declares, keeps the reference's enum
numbering, and its host-side trees and interaction lists (BaryTree_Analyze,
no GPU work) agree with the oracle's."""
import os
import re

import numpy as np
import pytest

from tests.helpers import uniform_cube

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


def test_unsupported_combinations_are_errors(bt):
    t, s = uniform_cube(100, 100, seed=1)
    with pytest.raises(RuntimeError, match="HERMITE"):
        bt.analyze(t, s, 1, [0.5], 1, 2, 1, 0.7, 3, 10, 10)
    with pytest.raises(RuntimeError, match="REGULARIZED"):
        bt.analyze(t, s, 3, [0.3, 0.05], 2, 1, 2, 0.7, 3, 10, 10)
    with pytest.raises(RuntimeError, match="KERNEL"):
        bt.analyze(t, s, 9, [0.5], 1, 1, 1, 0.7, 3, 10, 10)
    with pytest.raises(RuntimeError, match="degree"):
        bt.analyze(t, s, 1, [0.5], 1, 1, 1, 0.7, 13, 10, 10)
    with pytest.raises(RuntimeError, match="positive"):
        bt.analyze([a[:0] for a in t], s, 1, [0.5], 1, 1, 1, 0.7, 3, 10, 10)
