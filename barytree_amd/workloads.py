"""Synthetic particle distributions of the benchmark configurations.

uniform_cube  : x, y, z ~ U[-1, 1], q ~ U[-1, 1], w = 1
gaussian_mixture : points = centre + N(0, sigma^2) per axis, 8 centres ~ U[0, 1]^3
plummer       : Plummer-sphere radii r = a / sqrt(u^(-2/3) - 1), resampled
                while r > rmax (examples/support_fns.c:382-400 uses the same
                radial law), isotropic directions
Every generator is numpy's PCG64 with an explicit seed.
"""
import numpy as np


def uniform_cube(n, seed):
    rng = np.random.default_rng(seed)
    x, y, z = (rng.uniform(-1.0, 1.0, n) for _ in range(3))
    q = rng.uniform(-1.0, 1.0, n)
    return x, y, z, q, np.ones(n)


def gaussian_mixture(n, seed, centres=8, sigma=0.05):
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.0, 1.0, (centres, 3))
    k = rng.integers(0, centres, n)
    p = c[k] + rng.normal(0.0, sigma, (n, 3))
    q = rng.uniform(-1.0, 1.0, n)
    return p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy(), q, np.ones(n)


def plummer(n, seed, a=0.1, rmax=1.0):
    rng = np.random.default_rng(seed)
    r = np.empty(n)
    filled = 0
    while filled < n:
        u = rng.uniform(0.0, 1.0, n - filled)
        u = u[u > 0]
        rr = a / np.sqrt(u ** (-2.0 / 3.0) - 1.0)
        rr = rr[rr <= rmax]
        r[filled:filled + len(rr)] = rr
        filled += len(rr)
    cth = rng.uniform(-1.0, 1.0, n)
    phi = rng.uniform(0.0, 2.0 * np.pi, n)
    sth = np.sqrt(1.0 - cth ** 2)
    x, y, z = r * sth * np.cos(phi), r * sth * np.sin(phi), r * cth
    q = rng.uniform(-1.0, 1.0, n)
    return x, y, z, q, np.ones(n)


def batched_sizes(count, seed, nmin=1000, nmax=64000):
    """System sizes log-uniform in [nmin, nmax]."""
    rng = np.random.default_rng(seed)
    return np.floor(np.exp(rng.uniform(np.log(nmin), np.log(nmax), count))).astype(np.int64)
