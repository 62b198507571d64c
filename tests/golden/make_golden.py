"""Writes tests/golden/ref_outputs.npz: potentials computed by the reference's
own compiled sources (oracle/_ref/libbt_refparts.so, built by
`make -C oracle ref` from /root/reference/src) for small seeded inputs, one
vector per supported (compute type, kernel, singularity) plus the two
known-answer runs.  Inputs are regenerated from the seeds by
tests/helpers.py, so only outputs are stored.

Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from tests.helpers import KP, kat_inputs, supported_combos, uniform_cube  # noqa: E402

N, M = 2000, 1500


def cases():
    for ct, k, s in supported_combos():
        seed = 1000 + 100 * ct + 10 * k + s
        yield f"ct{ct}_k{k}_s{s}", seed, (k, KP[k][s], s, 1, ct, 0.7, 3, 60, 50)


def main():
    if not O.ref_available():
        O.build(ref=True)
    out = {}
    for name, seed, args in cases():
        t, s = uniform_cube(N, M, seed)
        out[name] = O.ref_treecode(*t, *s, *args)
    x, y, z, rho, w = kat_inputs()
    out["kat_theta"] = O.ref_treecode(x, y, z, rho, x.copy(), y.copy(), z.copy(), rho.copy(), w,
                                      2, [0.5], 2, 1, 1, 0.8, 4, 50, 10, 1.0, -1.0)
    out["kat_beta"] = O.ref_treecode(x, y, z, rho, x.copy(), y.copy(), z.copy(), rho.copy(), w,
                                     2, [0.5], 2, 1, 1, -1.0, -1, 1, 1, 1.0, 0.1)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ref_outputs.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, len(out), "vectors")


if __name__ == "__main__":
    main()
