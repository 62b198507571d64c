"""The oracle against the committed reference outputs (tests/golden/
ref_outputs.npz, written by tests/golden/make_golden.py from the reference's
own compiled sources): bit-identical, so the oracle stays pinned on machines
without the reference tree."""
import os

import numpy as np
import pytest

from tests.golden.make_golden import M, N, cases
from tests.helpers import kat_inputs, uniform_cube

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_outputs.npz")


@pytest.fixture(scope="module")
def golden():
    return np.load(GOLDEN)


@pytest.mark.parametrize("name,seed,args", list(cases()))
def test_oracle_matches_reference_outputs(oracle_mod, golden, name, seed, args):
    t, s = uniform_cube(N, M, seed)
    assert np.array_equal(oracle_mod.treecode(*t, *s, *args), golden[name])


def test_oracle_matches_reference_known_answer_runs(oracle_mod, golden):
    x, y, z, rho, w = kat_inputs()
    a = oracle_mod.treecode(x, y, z, rho, x.copy(), y.copy(), z.copy(), rho.copy(), w,
                            2, [0.5], 2, 1, 1, 0.8, 4, 50, 10, 1.0, -1.0)
    b = oracle_mod.treecode(x, y, z, rho, x.copy(), y.copy(), z.copy(), rho.copy(), w,
                            2, [0.5], 2, 1, 1, -1.0, -1, 1, 1, 1.0, 0.1)
    assert np.array_equal(a, golden["kat_theta"]) and np.array_equal(b, golden["kat_beta"])
