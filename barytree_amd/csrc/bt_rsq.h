// bt_rsq.h -- 1/sqrt(x) in FP64 for the gfx950 pair kernels.
#pragma once
#include <hip/hip_runtime.h>

namespace btk {

// The v_rsq_f64 estimate y (relative error below about 6e-8) refined by one
// third-order (Householder) step.  With d = x y^2 - 1, formed as
// fma(x y, y, -1),
//   1/sqrt(x) = y (1 + d)^(-1/2) = y (1 - d/2 + 3 d^2/8 - 5 d^3/16 + ...),
// so y + y d (3 d/8 - 1/2) leaves a truncation error near 5 |d|^3 / 16
// (below 1e-21) and about one ulp of rounding: 6 FP64 instructions, where
// two Newton steps y + y (1/2 - x y^2/2) take 9 for the same accuracy.
// x = 0 and x = inf give NaN (callers select the skip of coincident points).
__device__ __forceinline__ double rsq64(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    const double d = fma(x * y, y, -1.0);
    return fma(y, d * fma(d, 0.375, -0.5), y);
}

}  // namespace btk
