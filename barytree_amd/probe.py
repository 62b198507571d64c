"""Issue cost of the VALU instruction kinds on the running GPU
(libbt_probe.so, built from csrc/bt_probe.hip beside the drop-in library).

Measurement tooling for the benchmark's roofline: multiplied by a kernel's
per-kind instruction counts (rocprofv3 SQ_INSTS_VALU_* counters) the costs
give the time the VALU needs to issue that kernel's instruction mix.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PROBE_LIB = os.path.join(HERE, "libbt_probe.so")
VALU_KINDS = ("fma_f64", "add_f64", "fma_f32", "mul_f32", "exp_f32", "rsq_f32", "cvt_f64_f32")
_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(PROBE_LIB):
            raise RuntimeError(f"{PROBE_LIB} is not built (python -m barytree_amd.build)")
        _lib = ctypes.CDLL(PROBE_LIB)
        _lib.bt_probe_valu.restype = ctypes.c_int
        _lib.bt_probe_rsq64.restype = ctypes.c_int
    return _lib


def rsq64(x):
    """The pair kernels' FP64 1/sqrt (csrc/bt_rsq.h) evaluated on the GPU."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros_like(x)
    dp = ctypes.POINTER(ctypes.c_double)
    if _load().bt_probe_rsq64(x.ctypes.data_as(dp), y.ctypes.data_as(dp), ctypes.c_int(len(x))) != 0:
        raise RuntimeError("bt_probe_rsq64 failed")
    return y


def probe_valu():
    """{kind: ns per wave instruction and SIMD}, each kind measured alone
    (8 independent chains per lane, 8 workgroups of 256 lanes per CU)."""
    out = np.zeros(len(VALU_KINDS))
    n = _load().bt_probe_valu(out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.c_int(len(out)))
    if n != len(out):
        raise RuntimeError("bt_probe_valu failed")
    return dict(zip(VALU_KINDS, out.tolist()))
