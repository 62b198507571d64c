"""The device tree build's orchestration (barytree_amd/csrc/bt_devtree.cpp),
run on the CPU against emulations of its kernels (tests/devtree_emul.cpp),
gives the host builder's trees node for node: the same ranges, boxes,
centres, radii, children and levels, and the same particles in every leaf.
The host builder is the restatement of the reference's partition and split
rules (partition.c, tree_linked_list.c) that the oracle comparisons pin."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "barytree_amd", "csrc")


@pytest.fixture(scope="module")
def emul(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("devtree") / "libdevtree_emul.so")
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    cmd = ["g++", "-O1", "-shared", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
           "-D__HIP_PLATFORM_AMD__", f"-I{rocm}/include", "-fopenmp",
           os.path.join(ROOT, "tests", "devtree_emul.cpp"), os.path.join(SRC, "bt_devtree.cpp"),
           os.path.join(SRC, "bt_host.cpp"), os.path.join(SRC, "bt_devmem.cpp"), "-o", out]
    subprocess.run(cmd, check=True, capture_output=True)
    lib = ctypes.CDLL(out)
    lib.emul_compare.restype = ctypes.c_int
    return lib


def compare(lib, x, y, z, roots, maxpar, leaves_only):
    x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z))
    roots = np.ascontiguousarray(roots, dtype=np.int32)
    maxpar = np.ascontiguousarray(maxpar, dtype=np.int32)
    nn = ctypes.c_int(0)
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))  # noqa: E731
    ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))  # noqa: E731
    rc = lib.emul_compare(ctypes.c_int(len(x)), dp(x), dp(y), dp(z), ctypes.c_int(len(maxpar)), ip(roots),
                          ip(maxpar), ctypes.c_int(int(leaves_only)), ctypes.byref(nn))
    return rc, nn.value


@pytest.mark.parametrize("leaves_only", [False, True])
@pytest.mark.parametrize("n,leaf", [(1, 1), (2, 1), (7, 1), (100, 3), (5000, 60), (60000, 700), (200000, 8000)])
def test_uniform(emul, n, leaf, leaves_only):
    rng = np.random.default_rng(n + leaf)
    x, y, z = rng.uniform(-1, 1, (3, n))
    rc, nn = compare(emul, x, y, z, [0, n], [leaf], leaves_only)
    assert rc == 0 and nn >= 1


@pytest.mark.parametrize("leaves_only", [False, True])
@pytest.mark.parametrize("leaf", [1, 2, 3, 5, 17])
def test_lattice_ties_with_midpoints(emul, leaf, leaves_only):
    # lattice points: node midpoints fall on particle coordinates, so the
    # "<= mid" rule and the one-element "< mid" rule both decide
    g = np.arange(6, dtype=np.float64) * 0.25
    X, Y, Z = np.meshgrid(g, g[:4], g[:5], indexing="ij")
    x, y, z = X.ravel(), Y.ravel(), Z.ravel()
    perm = np.random.default_rng(leaf).permutation(len(x))
    rc, _ = compare(emul, x[perm], y[perm], z[perm], [0, len(x)], [leaf], leaves_only)
    assert rc == 0


@pytest.mark.parametrize("shape", [(3, 3, 3), (5, 5, 5), (5, 3, 7)])
@pytest.mark.parametrize("leaf", [1, 2])
def test_odd_lattices_one_element_ranges(emul, shape, leaf):
    # odd point counts put every midpoint on the lattice, so one-element
    # ranges meet particles at the midpoint: they go to the upper part
    axes = [np.linspace(0.0, 1.0, k) for k in shape]
    X, Y, Z = np.meshgrid(*axes, indexing="ij")
    x, y, z = X.ravel(), Y.ravel(), Z.ravel()
    for seed in range(3):
        perm = np.random.default_rng(seed).permutation(len(x))
        rc, _ = compare(emul, x[perm], y[perm], z[perm], [0, len(x)], [leaf], False)
        assert rc == 0, seed


def test_one_element_range_at_the_midpoint(emul):
    # after the x split the lower range holds one particle whose y is the
    # node's mid y: it belongs to the upper y part (partition.c's one-element
    # case compares with <), unlike a particle of a larger range
    x = np.array([0.0, 1.0, 1.0, 1.0])
    y = np.array([0.5, 0.0, 1.0, 0.0])
    z = np.array([0.0, 1.0, 0.0, 0.0])
    rc, nn = compare(emul, x, y, z, [0, 4], [1], False)
    assert rc == 0 and nn == 5


def test_duplicates_and_flat_boxes(emul):
    rng = np.random.default_rng(3)
    n = 3000
    x = np.round(rng.uniform(-1, 1, n), 1)          # many equal coordinates
    y = np.zeros(n)                                   # a flat dimension
    z = rng.choice([0.0, 0.5, -0.5], n)
    mult = np.unique(np.stack([x, z]), axis=1, return_counts=True)[1].max()
    for leaf in (mult, mult + 1, 2 * mult, 500):
        rc, _ = compare(emul, x, y, z, [0, n], [leaf], False)
        assert rc == 0, leaf
    # more coincident particles than the leaf size: no split separates them
    # (the reference recurses without end); both builders refuse the input
    for leaf in (1, 4, mult - 1):
        rc, _ = compare(emul, x, y, z, [0, n], [leaf], False)
        assert rc == 7, leaf


def test_clustered(emul):
    rng = np.random.default_rng(4)
    n = 80000
    r = 0.1 / np.sqrt(rng.uniform(0.01, 1, n) ** (-2 / 3) - 1)
    v = rng.normal(size=(3, n))
    v /= np.linalg.norm(v, axis=0)
    x, y, z = np.minimum(r, 1.0) * v
    for leaf in (500, 2000):
        rc, _ = compare(emul, x, y, z, [0, n], [leaf], False)
        assert rc == 0, leaf


@pytest.mark.parametrize("leaves_only", [False, True])
def test_many_systems(emul, leaves_only):
    # several systems of different sizes and leaf sizes, concatenated
    rng = np.random.default_rng(5)
    sizes = [1, 2, 37, 1000, 4097, 9000, 5]
    leaves = [1, 1, 4, 30, 100, 2000, 2]
    parts = [rng.normal(c, 0.05, (3, s)) for c, s in zip(rng.uniform(0, 1, len(sizes)), sizes)]
    x, y, z = np.concatenate(parts, axis=1)
    roots = np.concatenate([[0], np.cumsum(sizes)])
    rc, nn = compare(emul, x, y, z, roots, leaves, leaves_only)
    assert rc == 0 and nn >= len(sizes)


def test_device_memory_cache_on_the_emulated_allocator(emul):
    """bt_devmem.cpp: released blocks are reused for close sizes only,
    the limit trims the largest cached blocks first and frees a release
    beyond it, an allocation that runs out of memory empties the cache and
    tries again, and emptying reports the bytes freed."""
    emul.emul_devmem.restype = ctypes.c_int
    assert emul.emul_devmem() == 0
