"""GPU tests of the probe library (libbt_probe.so, barytree_amd.probe): the
VALU issue-cost probe, which prices the near-field kernel's instruction mix
in bench.py, and the accuracy of the FP64 reciprocal square root every FP64
pair kernel uses (csrc/bt_rsq.h)."""
import math

import numpy as np

import pytest

pytestmark = pytest.mark.gpu


def test_probe_valu_costs_are_sane():
    from barytree_amd import probe
    p = probe.probe_valu()
    assert set(p) == set(probe.VALU_KINDS)
    for k, v in p.items():
        # a wave instruction takes between 1 and 64 cycles at 0.5-3 GHz
        assert math.isfinite(v) and 0.3 < v < 130.0, (k, v)
    # transcendentals issue no faster than an FP32 fma
    assert p["exp_f32"] >= 0.9 * p["fma_f32"] and p["rsq_f32"] >= 0.9 * p["fma_f32"]


def test_rsq64_is_within_about_one_ulp():
    """rsq64 against 1/sqrt in extended precision (x87 long double, 64-bit
    significand) over 60 decades, powers of two, values near 1 and the
    squares of the near-field's typical distances."""
    from barytree_amd import probe
    rng = np.random.default_rng(5)
    x = np.concatenate([
        10.0 ** rng.uniform(-30, 30, 200_000),
        2.0 ** np.arange(-200, 201).astype(np.float64),
        1.0 + rng.uniform(-1e-6, 1e-6, 10_000),
        rng.uniform(0.0, 1.0, 100_000) ** 2 + 1e-300,
    ])
    y = probe.rsq64(x)
    ref = 1.0 / np.sqrt(x.astype(np.longdouble))
    assert np.finfo(np.longdouble).nmant >= 63, "needs an extended-precision long double"
    ulps = np.abs(y.astype(np.longdouble) - ref) / np.spacing(ref.astype(np.float64))
    print(f"rsq64: max error {float(ulps.max()):.3f} ulp, mean {float(ulps.mean()):.3f} ulp", flush=True)
    assert float(ulps.max()) <= 1.5
    # zero and infinity give NaN (callers skip coincident points by a select)
    assert np.all(np.isnan(probe.rsq64(np.array([0.0, np.inf]))))
