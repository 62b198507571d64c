// bt_noalloc.h -- device scratch that needs no allocation: one file-scope
// array per GPU (bt_noalloc.hip), for entry points whose callers cannot
// allocate device memory (code under HIP graph capture, serving loops).
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>

namespace btk {

// Size of the scratch array (bytes) and the largest direct sample it holds.
constexpr size_t NOALLOC_SCRATCH_BYTES = 5u << 20;
constexpr int NOALLOC_MAX_SAMPLE = 8192;

// The current GPU's scratch array (its device address, looked up once per
// GPU).  Every caller on a GPU shares it: calls that overlap on one GPU race.
hipError_t noalloc_scratch(void **ptr, size_t *bytes);

// pos[k] = -1 for k < m.
hipError_t launch_fill_neg(int m, int *pos, hipStream_t st);
// For the tree-order positions i in [b, e): the caller index c = order[i]
// is looked up in sel[0..m) (ascending, duplicates allowed) and every k with
// sel[k] == c gets pos[k] = i.
hipError_t launch_find_sel(int b, int e, const int *order, int m, const int *sel, int *pos, hipStream_t st);
// o*[k] = *[pos[k]] (NaN where pos[k] < 0).
hipError_t launch_gather_pos(int m, const int *pos, const double *x, const double *y, const double *z,
                             const double *q, double *ox, double *oy, double *oz, double *oq, hipStream_t st);

}  // namespace btk
