"""Pins of the oracle's learned-degree parts (CPU only).

- Variable-degree treecode: with every far-field pair at the plan's degree it
  must equal the fixed-degree oracle (which is bit-identical to the
  reference's sources): charges from particles instead of child-to-parent
  and interpolation straight to particles instead of parent-to-child are
  exact rewrites, so only rounding differs.  Raising the degree of every pair
  must lower the error against the direct sum.
- Pair labels: the per-degree pair errors against an independent evaluation
  with scipy's BarycentricInterpolator (its own barycentric weights), the
  label rule, and convergence with the degree.
- Degree-model loss and gradient: central finite differences, and the closed
  form of a network whose weights are zero.
"""
import numpy as np
import pytest
from scipy.interpolate import BarycentricInterpolator

from tests.helpers import KP, rel_l2, supported_combos, uniform_cube


def n_far_pairs(st, ct):
    return {1: st["num_pc"], 2: st["num_cp"], 3: st["num_cc"] + st["num_pc"] + st["num_cp"]}[ct]


@pytest.mark.parametrize("ct,k,s", supported_combos())
def test_vardeg_uniform_degree_equals_fixed(oracle_mod, ct, k, s):
    t, src = uniform_cube(1500, 1200, seed=40 + 10 * ct + k)
    args = (k, KP[k][s], s, 1, ct, 0.7, 3, 40, 40)
    ref = oracle_mod.treecode(*t, *src, *args)
    npairs = n_far_pairs(oracle_mod.last_stats(), ct)
    assert npairs > 0
    got = oracle_mod.treecode_vardeg(*t, *src, *args, np.full(npairs, 3))
    assert rel_l2(ref, got) <= 1e-13


@pytest.mark.parametrize("p", [2, 6])
@pytest.mark.parametrize("ct,k,s", supported_combos())
def test_vardeg_uniform_other_degree_equals_fixed_at_that_degree(oracle_mod, ct, k, s, p):
    """Every far-field pair at degree p != the plan's degree 4 equals the
    fixed-degree oracle run at degree p.  size_check 0 makes the
    interaction lists independent of the degree (interaction_lists.c:484-486,
    557), so both runs see the same pairs; a per-pair degree that the
    variable-degree path dropped, or a grid it built at the plan's degree,
    would show here."""
    t, src = uniform_cube(1500, 1200, seed=60 + 10 * ct + k)
    args = (k, KP[k][s], s, 1, ct, 0.7)
    ref = oracle_mod.treecode(*t, *src, *args, p, 40, 40, 0.0)
    npairs_p = n_far_pairs(oracle_mod.last_stats(), ct)
    oracle_mod.treecode(*t, *src, *args, 4, 40, 40, 0.0)
    npairs = n_far_pairs(oracle_mod.last_stats(), ct)
    assert npairs == npairs_p and npairs > 0
    got = oracle_mod.treecode_vardeg(*t, *src, *args, 4, 40, 40, np.full(npairs, p), 0.0)
    assert rel_l2(ref, got) <= 1e-13


@pytest.mark.parametrize("ct", [1, 2, 3])
def test_vardeg_higher_degree_lowers_error(oracle_mod, ct):
    t, src = uniform_cube(2000, 2000, seed=7 + ct)
    args = (1, [0.5], 1, 1, ct, 0.8, 3, 40, 40)
    base = oracle_mod.treecode(*t, *src, *args)
    npairs = n_far_pairs(oracle_mod.last_stats(), ct)
    direct = oracle_mod.direct(*t, *src, 1, [0.5], 1)
    hi = oracle_mod.treecode_vardeg(*t, *src, *args, np.full(npairs, 8))
    rng = np.random.default_rng(ct)
    mixed = oracle_mod.treecode_vardeg(*t, *src, *args, rng.integers(5, 9, npairs))
    e3, e8, em = rel_l2(direct, base), rel_l2(direct, hi), rel_l2(direct, mixed)
    print(ct, e3, e8, em)
    assert e8 < 0.05 * e3
    assert e8 <= em < e3
    with pytest.raises(RuntimeError):
        oracle_mod.treecode_vardeg(*t, *src, *args, np.full(npairs + 1, 3))


# ---- pair labels -----------------------------------------------------------

def cheb(box, p):
    u = (np.cos(np.arange(p + 1) * np.pi / p) + 1.0) * 0.5
    return [box[2 * d] + u * (box[2 * d + 1] - box[2 * d]) for d in range(3)]


def basis_rows(nodes, x):
    """Lagrange basis values at x via scipy (identity data, own weights)."""
    return BarycentricInterpolator(nodes, np.eye(len(nodes)))(x)


def green(kernel, p0, d):
    r2 = np.sum(d * d, axis=-1)
    if kernel == 3:
        return 1.0 / np.sqrt(r2 + p0 * p0)
    r = np.sqrt(r2)
    return np.exp(-p0 * r) / r if kernel == 2 else 1.0 / r


def scipy_pair_error(kind, X, Y, q, tbox, sbox, p, kernel, p0):
    exact = np.array([np.sum(q * green(kernel, p0, x - Y)) for x in X])
    nl = p + 1
    gs = np.stack(np.meshgrid(*cheb(sbox, p), indexing="ij"), -1)  # [k1,k2,k3,3]
    gt = np.stack(np.meshgrid(*cheb(tbox, p), indexing="ij"), -1)
    if kind != 2:
        Bs = [basis_rows(cheb(sbox, p)[d], Y[:, d]) for d in range(3)]  # [ns, nl]
        Q = np.einsum("si,sj,sk,s->ijk", Bs[0], Bs[1], Bs[2], q)
    if kind == 1:
        approx = np.array([np.sum(Q * green(kernel, p0, x - gs)) for x in X])
    else:
        if kind == 2:
            V = np.array([np.sum(q * green(kernel, p0, g - Y)) for g in gt.reshape(-1, 3)])
        else:
            V = np.array([np.sum(Q.reshape(-1) * green(kernel, p0, g - gs.reshape(-1, 3)))
                          for g in gt.reshape(-1, 3)])
        V = V.reshape(nl, nl, nl)
        Bt = [basis_rows(cheb(tbox, p)[d], X[:, d]) for d in range(3)]
        approx = np.einsum("ti,tj,tk,ijk->t", Bt[0], Bt[1], Bt[2], V)
    return np.sqrt(np.sum((approx - exact) ** 2) / np.sum(exact ** 2))


def make_pairs(rng, n, spread=3.0):
    """Random pairs of clusters: a target box and a source box `spread` box
    widths apart; particles inside the boxes."""
    kinds, tb, tc, sb, sc, tbox, sbox = [], [], [], [], [], [], []
    T, S, Q = [], [], []
    nt_all = ns_all = 0
    for i in range(n):
        w = rng.uniform(0.2, 1.0)
        c0 = rng.uniform(-1, 1, 3)
        dirn = rng.normal(size=3)
        dirn /= np.linalg.norm(dirn)
        c1 = c0 + dirn * w * rng.uniform(spread, 2 * spread)
        nt, ns = int(rng.integers(5, 60)), int(rng.integers(5, 300))
        X = c0 + rng.uniform(-w / 2, w / 2, (nt, 3))
        Y = c1 + rng.uniform(-w / 2, w / 2, (ns, 3))
        T.append(X); S.append(Y); Q.append(rng.uniform(-1, 1, ns))
        bx = np.ravel([[X[:, d].min(), X[:, d].max()] for d in range(3)])
        by = np.ravel([[Y[:, d].min(), Y[:, d].max()] for d in range(3)])
        kinds.append(1 + i % 3); tb.append(nt_all); tc.append(nt); sb.append(ns_all); sc.append(ns)
        tbox.append(bx); sbox.append(by)
        nt_all += nt; ns_all += ns
    T, S, Q = np.concatenate(T), np.concatenate(S), np.concatenate(Q)
    return (np.array(kinds), np.array(tb), np.array(tc), np.array(sb), np.array(sc),
            np.concatenate(tbox), np.concatenate(sbox), T, S, Q)


def test_label_errors_match_scipy_interpolation(oracle_mod):
    rng = np.random.default_rng(3)
    kinds, tb, tc, sb, sc, tbox, sbox, T, S, Q = make_pairs(rng, 9)
    for kernel, p0 in ((1, 0.0), (2, 0.5), (3, 0.05)):
        labels, errors = oracle_mod.label_pairs(kinds, tb, tc, sb, sc, tbox, sbox, T[:, 0], T[:, 1],
                                                T[:, 2], S[:, 0], S[:, 1], S[:, 2], Q, kernel, p0,
                                                1e-30)
        assert np.all(labels == 13)  # unreachable tolerance: every degree computed
        for i in range(len(kinds)):
            # stride samples: up to 32 targets, up to 128 sources
            nt, ns = min(tc[i], 32), min(sc[i], 128)
            X = T[tb[i] + (np.arange(nt) * tc[i]) // nt]
            si = sb[i] + (np.arange(ns) * sc[i]) // ns
            for p in (2, 5, 9):
                e = scipy_pair_error(kinds[i], X, S[si], Q[si], tbox[6 * i:6 * i + 6],
                                     sbox[6 * i:6 * i + 6], p, kernel, p0)
                assert abs(e - errors[i, p - 2]) <= 1e-12 + 1e-7 * e, (kernel, i, p, e, errors[i, p - 2])


def green_w(kernel, p0, p1, d):
    """The subtraction kernels' second term (the weight's kernel)."""
    if kernel == 2:
        return green(kernel, p0, d)
    return np.exp(-np.sum(d * d, axis=-1) * p1) * green(kernel, p0, d)


def scipy_term(kind, X, Y, c, tbox, sbox, p, kern):
    """(exact, approximation) at X of the potential of charges c at Y with
    kernel kern(d), interpolated as the treecode does, basis from scipy."""
    exact = np.array([np.sum(c * kern(x - Y)) for x in X])
    nl = p + 1
    gs = np.stack(np.meshgrid(*cheb(sbox, p), indexing="ij"), -1)
    gt = np.stack(np.meshgrid(*cheb(tbox, p), indexing="ij"), -1)
    if kind != 2:
        Bs = [basis_rows(cheb(sbox, p)[d], Y[:, d]) for d in range(3)]
        Q = np.einsum("si,sj,sk,s->ijk", Bs[0], Bs[1], Bs[2], c)
    if kind == 1:
        return exact, np.array([np.sum(Q * kern(x - gs)) for x in X])
    if kind == 2:
        V = np.array([np.sum(c * kern(g - Y)) for g in gt.reshape(-1, 3)])
    else:
        V = np.array([np.sum(Q.reshape(-1) * kern(g - gs.reshape(-1, 3))) for g in gt.reshape(-1, 3)])
    Bt = [basis_rows(cheb(tbox, p)[d], X[:, d]) for d in range(3)]
    return exact, np.einsum("ti,tj,tk,ijk->t", Bt[0], Bt[1], Bt[2], V.reshape(nl, nl, nl))


def test_label_errors_ss_match_scipy_interpolation(oracle_mod):
    """Subtraction labels: the pair potential sum_s (q_s - q_t g_w/g) w_s g,
    each term interpolated on its own; errors against an independent
    evaluation through scipy's BarycentricInterpolator, and with zero target
    charges and unit weights the labels of the skipping kernel."""
    rng = np.random.default_rng(13)
    kinds, tb, tc, sb, sc, tbox, sbox, T, S, Q = make_pairs(rng, 9)
    tq = rng.uniform(-1, 1, len(T))
    w = rng.uniform(0.5, 1.5, len(S))
    for kernel, p0, p1 in ((1, 0.0, 4.0), (2, 0.5, 0.0), (3, 0.05, 1.0 / 0.09)):
        labels, errors = oracle_mod.label_pairs_ss(kinds, tb, tc, sb, sc, tbox, sbox, T[:, 0], T[:, 1], T[:, 2],
                                                   tq, S[:, 0], S[:, 1], S[:, 2], Q, w, kernel, p0, p1, 1e-30)
        assert np.all(labels == 13)
        for i in range(len(kinds)):
            nt, ns = min(tc[i], 32), min(sc[i], 128)
            ti = tb[i] + (np.arange(nt) * tc[i]) // nt
            si = sb[i] + (np.arange(ns) * sc[i]) // ns
            X, Y = T[ti], S[si]
            for p in (2, 5, 9):
                bx, by = tbox[6 * i:6 * i + 6], sbox[6 * i:6 * i + 6]
                e0, a0 = scipy_term(kinds[i], X, Y, Q[si] * w[si], bx, by, p, lambda d: green(kernel, p0, d))
                e1, a1 = scipy_term(kinds[i], X, Y, w[si], bx, by, p, lambda d: green_w(kernel, p0, p1, d))
                ex, ap = e0 - tq[ti] * e1, a0 - tq[ti] * a1
                e = np.sqrt(np.sum((ap - ex) ** 2) / np.sum(ex ** 2))
                assert abs(e - errors[i, p - 2]) <= 1e-12 + 1e-7 * e, (kernel, i, p, e, errors[i, p - 2])
        # the second term vanishes with zero target charges: the skipping kernel's labels
        l0, e0 = oracle_mod.label_pairs_ss(kinds, tb, tc, sb, sc, tbox, sbox, T[:, 0], T[:, 1], T[:, 2],
                                           np.zeros(len(T)), S[:, 0], S[:, 1], S[:, 2], Q, np.ones(len(S)),
                                           kernel, p0, p1, 1e-7)
        lk, ek = oracle_mod.label_pairs(kinds, tb, tc, sb, sc, tbox, sbox, T[:, 0], T[:, 1], T[:, 2],
                                        S[:, 0], S[:, 1], S[:, 2], Q, kernel, p0, 1e-7)
        assert np.array_equal(l0, lk)
        assert np.array_equal(np.isnan(e0), np.isnan(ek)) and np.allclose(e0[~np.isnan(e0)], ek[~np.isnan(ek)],
                                                                          rtol=1e-14, atol=0)


def test_label_rule_and_convergence(oracle_mod):
    rng = np.random.default_rng(5)
    kinds, tb, tc, sb, sc, tbox, sbox, T, S, Q = make_pairs(rng, 12, spread=4.0)
    args = (kinds, tb, tc, sb, sc, tbox, sbox, T[:, 0], T[:, 1], T[:, 2], S[:, 0], S[:, 1],
            S[:, 2], Q, 1, 0.0)
    _, full = oracle_mod.label_pairs(*args, 1e-30)
    # errors fall with the degree (well separated pairs converge geometrically)
    assert np.all(full[:, 10] < 1e-3 * full[:, 0])
    assert np.all(full[:, 10] < 1e-9)
    for tol in (1e-4, 1e-7):
        labels, errs = oracle_mod.label_pairs(*args, tol)
        for i, lab in enumerate(labels):
            assert 2 <= lab <= 12
            assert errs[i, lab - 2] <= tol
            assert np.all(full[i, :lab - 2] > tol)           # no smaller degree reaches tol
            assert np.all(np.isnan(errs[i, lab - 1:]))        # nothing computed beyond the label
    l4, _ = oracle_mod.label_pairs(*args, 1e-4)
    l7, _ = oracle_mod.label_pairs(*args, 1e-7)
    assert np.all(l7 >= l4) and np.any(l7 > l4)


# ---- degree-model loss and gradient ------------------------------------------

def random_model(rng, scale_w=0.3):
    n = 9546
    p = rng.normal(0.0, scale_w, n)
    return p, rng.normal(0, 1, 8), rng.uniform(0.5, 2.0, 8)


def test_mlp_gradient_matches_finite_differences(oracle_mod):
    rng = np.random.default_rng(11)
    assert oracle_mod.mlp_num_params() == 9546
    p, mu, sc = random_model(rng)
    x = rng.normal(0, 1, (40, 8)).astype(np.float32)
    y = rng.integers(2, 14, 40)
    loss, g = oracle_mod.mlp_loss_grad(p, mu, sc, x, y)
    idx = np.concatenate([rng.choice(512, 10, replace=False), 512 + rng.choice(64, 4, replace=False),
                          576 + rng.choice(4096, 10, replace=False), 8896 + rng.choice(650, 10, replace=False)])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (oracle_mod.mlp_loss_grad(pp, mu, sc, x, y, False)[0] -
              oracle_mod.mlp_loss_grad(pm, mu, sc, x, y, False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(g[i]), (i, fd, g[i])


def test_mlp_zero_weights_closed_form(oracle_mod):
    rng = np.random.default_rng(2)
    p = np.zeros(9546)
    b4 = rng.normal(0, 1, 10)
    p[9536:] = b4
    x = rng.normal(0, 1, (25, 8)).astype(np.float32)
    y = rng.integers(2, 14, 25)
    loss, g = oracle_mod.mlp_loss_grad(p, np.zeros(8), np.ones(8), x, y)
    Y = (y[:, None] > 2 + np.arange(10)[None, :]).astype(float)
    sp = np.maximum(b4, 0) + np.log1p(np.exp(-np.abs(b4)))
    assert np.isclose(loss, np.mean(np.sum(sp[None, :] - Y * b4[None, :], axis=1)), rtol=1e-14)
    sig = 1 / (1 + np.exp(-b4))
    assert np.allclose(g[9536:], np.mean(sig[None, :] - Y, axis=0), rtol=1e-13, atol=1e-16)
    assert np.all(g[:9536] == 0.0)
    z = oracle_mod.mlp_logits(p, np.zeros(8), np.ones(8), x)
    assert np.allclose(z, b4[None, :])
