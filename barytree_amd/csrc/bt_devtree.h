// bt_devtree.h -- tree construction on the GPU (host orchestration).
#pragma once

#include <hip/hip_runtime_api.h>

#include <string>
#include <vector>

#include "bt_host.h"

namespace bt {

// Coordinates and caller indices of one particle set on the device, with a
// second set of buffers of the same size for the partition passes (they swap).
struct DevParticles {
    int n = 0;
    double *x = nullptr, *y = nullptr, *z = nullptr;
    int *idx = nullptr;
    double *x2 = nullptr, *y2 = nullptr, *z2 = nullptr;
    int *idx2 = nullptr;
};

// One tree of the set: particles [b, e) of the arrays (a system), split while
// a node holds more than maxpar particles.
struct DevRoot {
    int b = 0, e = 0, maxpar = 1;
};

// Builds the trees of `roots` (disjoint ranges that cover [0, p.n)) on the
// device, with the reference's node membership: tight boxes from the member
// particles (tree_linked_list.c:53-58), the split rule of node_split, and
// for each split dimension a partition at the midpoint that sends
// coord <= mid to the lower part (coord < mid for a one-element range,
// partition.c:94-175); children in the order of the host builder, empty
// children dropped (tree_linked_list.c:115-116).  On return p.x, p.y, p.z,
// p.idx hold the particles in tree order (within a leaf the order is the
// partition's stable order, not the reference's pivot-hole order; the
// numbers that depend on it change only by rounding).  nodes[r] gets root
// r's records in preorder (ranges relative to the root's b; leaves_only:
// only the leaves, in DFS order), depth[r] its depth.
std::string device_build_trees(DevParticles &p, const std::vector<DevRoot> &roots, bool leaves_only,
                               hipStream_t st, std::vector<std::vector<NodeRec>> &nodes,
                               std::vector<int> &depth);

// The leaves of a full tree's records, as batch records (DFS order).
std::vector<NodeRec> leaf_records(const std::vector<NodeRec> &full);

}  // namespace bt
