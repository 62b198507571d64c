"""Pins of the oracle's Hermite approximation (CPU only).

- PC and CP results are bit-equal to the reference's own compiled Hermite
  kernels and downpass (oracle/_ref; the upward pass there is the oracle's);
- the upward pass (clusters.c:1224-1836) reproduces, per dimension,
  polynomials of degree <= 2n+1 from values and derivatives at the n+1
  Chebyshev nodes (the 8 planes weight p, p_x, p_y, p_z, p_xy, p_yz, p_xz,
  p_xyz at the grid points), and stops reproducing at degree 2n+2;
- particles on a node coordinate (the box faces always hold some) take the
  exact-hit branch without breaking that;
- the treecode converges to the direct sum with the degree;
- the combinations the reference rejects are rejected.
"""
import numpy as np
import pytest

from tests.helpers import KP, hermite_combos, rel_l2, uniform_cube


@pytest.mark.parametrize("ct,k,s", hermite_combos())
def test_hermite_bit_equal_to_reference_sources(oracle_mod, ct, k, s):
    if not oracle_mod.ref_available():
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    t, src = uniform_cube(2500, 2000, seed=70 + ct * 7 + k)
    args = (k, KP[k][s], s, 2, ct, 0.7, 3, 60, 50)
    a = oracle_mod.treecode(*t, *src, *args)
    st = oracle_mod.last_stats()
    assert st["num_pc"] + st["num_cp"] > 0
    assert np.array_equal(a, oracle_mod.ref_treecode(*t, *src, *args))


def _moment(Q, X, Y, Z, a, b, c):
    def d(e, v, k):
        if k == 0:
            return v ** e
        return e * v ** (e - 1) if e >= 1 else 0.0 * v
    val = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (0, 1, 1), (1, 0, 1), (1, 1, 1)]
    return sum(np.sum(Q[m] * d(a, X, i) * d(b, Y, j) * d(c, Z, k)) for m, (i, j, k) in enumerate(val))


@pytest.mark.parametrize("deg", [2, 3, 5])
def test_hermite_upward_reproduces_polynomials(oracle_mod, deg):
    rng = np.random.default_rng(deg)
    n = 400
    x, y, z = (rng.uniform(-1, 1, n) for _ in range(3))
    q = rng.uniform(-1, 1, n)
    w = rng.uniform(0.5, 1.5, n)
    # one particle on an interior node coordinate in x (exact hit)
    x0, x1 = x.min(), x.max()
    x[5] = x0 + (np.cos(2 * np.pi / deg) + 1.0) / 2.0 * (x1 - x0)
    for sing in (1, 2):
        cx, cy, cz, Q, W = oracle_mod.source_clusters(x, y, z, q, w, deg, 10 ** 6, sing, approximation=2)
        assert Q.shape == (1, 8, (deg + 1) ** 3)
        charge = q * w if sing == 2 else q
        X, Y, Z = cx[0], cy[0], cz[0]
        top = 2 * deg + 1
        for (a, b, c) in [(0, 0, 0), (1, 0, 0), (0, 1, 1), (1, 1, 1), (top, 0, 0), (0, top, 1),
                          (top, top, top)]:
            want = np.sum(charge * x ** a * y ** b * z ** c)
            got = _moment(Q[0], X, Y, Z, a, b, c)
            assert abs(got - want) <= 1e-11 * np.sum(np.abs(charge)), (sing, a, b, c, got, want)
            if sing == 2:
                gw = _moment(W[0], X, Y, Z, a, b, c)
                assert abs(gw - np.sum(w * x ** a * y ** b * z ** c)) <= 1e-11 * np.sum(w)
        # degree 2n+2 is beyond Hermite interpolation on n+1 nodes
        want = np.sum(charge * x ** (top + 1))
        assert abs(_moment(Q[0], X, Y, Z, top + 1, 0, 0) - want) > 1e-6 * np.sum(np.abs(charge))


@pytest.mark.parametrize("ct,k,s", hermite_combos())
def test_hermite_converges_to_direct(oracle_mod, ct, k, s):
    t, src = uniform_cube(2000, 2000, seed=90 + ct + 3 * k + s)
    direct = oracle_mod.direct(*t, *src, k, KP[k][s], s)
    errs = []
    for deg in (1, 3, 5):
        out = oracle_mod.treecode(*t, *src, k, KP[k][s], s, 2, ct, 0.7, deg, 50, 50)
        errs.append(rel_l2(direct, out))
    assert errs[2] < 0.05 * errs[0] and errs[2] < 1e-6, errs


def test_hermite_unsupported_combinations(oracle_mod):
    t, src = uniform_cube(300, 300, seed=1)
    for ct, k, s in [(3, 1, 1), (3, 2, 2), (2, 1, 2), (2, 2, 2)]:
        with pytest.raises(RuntimeError):
            oracle_mod.treecode(*t, *src, k, KP[k][s], s, 2, ct, 0.7, 3, 50, 50)
