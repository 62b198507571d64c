"""GPU test of the VALU issue-cost probe (libbt_probe.so, barytree_amd.probe),
which prices the near-field kernel's instruction mix in bench.py."""
import math

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
