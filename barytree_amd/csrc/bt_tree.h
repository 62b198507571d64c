// bt_tree.h -- launch interface of the device tree-build kernels (bt_tree.hip).
//
// The trees are built level by level on the GPU: exact bounding boxes of the
// nodes of a level (chunked min/max), then for each dimension a node splits
// in, a stable binary partition of every range of the level around the
// node's midpoint.  The split decisions themselves are made on the host
// (bt::node_split), so node membership follows the reference's rules
// exactly.  Every pointer is a device pointer.
#pragma once

#include <hip/hip_runtime_api.h>

namespace btk {

constexpr int TREE_CHUNK = 4096;  // elements per block of the tree kernels

// A range of the particle arrays that one pass either splits (dim 0..2:
// x, y or z against mid) or passes through unchanged (dim -1).  A split sends
// an element to the lower part iff coord <= mid, or coord < mid when the
// range holds exactly one element (strict = 1): partition.c:94-175 (the
// one-element case returns "below" only for a key < value).  lowtot is the
// lower part's size (filled in after the count pass).
struct alignas(16) TreeSeg {
    double mid;
    int beg, cnt, dim, strict, lowtot, pad;
};
// A chunk of at most TREE_CHUNK elements of one segment; lowoff = lower-part
// elements of the segment's earlier chunks (after the count pass).
struct alignas(16) TreeChunk { int seg, beg, cnt, lowoff; };

// part[6c .. 6c+5] = min x, max x, min y, max y, min z, max z of chunk c.
hipError_t launch_tree_bbox(int nchunks, const TreeChunk *chunks, const double *x, const double *y,
                            const double *z, double *part, hipStream_t st);
// low[c] = elements of chunk c that go to the lower part.
hipError_t launch_tree_count(int nchunks, const TreeChunk *chunks, const TreeSeg *segs,
                             const double *x, const double *y, const double *z, int *low,
                             hipStream_t st);
// Stable partition of every segment from (x, y, z, idx) into (ox, oy, oz,
// oidx): lower part first, each part in its input order.
hipError_t launch_tree_scatter(int nchunks, const TreeChunk *chunks, const TreeSeg *segs,
                               const double *x, const double *y, const double *z, const int *idx,
                               double *ox, double *oy, double *oz, int *oidx, hipStream_t st);
// out_a[i] = a[idx[i]]; out_b[i] = b ? b[idx[i]] : fill_b (out_b may be null).
hipError_t launch_tree_gather(int n, const int *idx, const double *a, double *out_a, const double *b,
                              double fill_b, double *out_b, hipStream_t st);
// *flag = 1 if a[i] and b[i] differ in any bit for some i < n (else untouched).
hipError_t launch_differ(int n, const double *a, const double *b, int *flag, hipStream_t st);
// idx[i] = i (caller order before the build).
hipError_t launch_iota(int n, int *idx, hipStream_t st);
// inv[order[i]] = i.
hipError_t launch_invert(int n, const int *order, int *inv, hipStream_t st);
// For caller indices sel[k]: j = inv[sel[k]], o*[k] = *[j].
hipError_t launch_gather_sel(int m, const int *sel, const int *inv, const double *x, const double *y,
                             const double *z, const double *q, double *ox, double *oy, double *oz,
                             double *oq, hipStream_t st);

}  // namespace btk
