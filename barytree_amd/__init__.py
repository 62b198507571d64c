"""barytree_amd -- MI355X-native BaryTree interaction path.

The product is the C ABI library libBaryTree_gpu.so (include/
BaryTreeInterface.h): host C++ builds trees and interaction lists, gfx950 HIP
kernels do the upward pass, the near and far field and the downpass.  This
package holds its sources (csrc/), its build script (build.py) and the
Python front end (interface.py).
"""
from .interface import (  # noqa: F401
    LIB_PATH, PRECISION_FP64, PRECISION_MIXED, Approximation, ComputeType, Kernel, Plan,
    DegreeModel, Schedule, Singularity, analyze, analyze_batched, analyze_degrees, autotune, callTreedriver,
    label_pairs, label_pairs_subtraction, last_error, last_stats, lib, solve, solve_batched, solve_learned,
    solve_device, DeviceBuffer, Batch,
    BETA_LEAVES_CPU, BETA_LEAVES_GPU, set_beta_leaf_sizes, empty_device_cache, set_device_cache_limit,
    cached_device_bytes)
