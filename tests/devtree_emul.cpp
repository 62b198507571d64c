// CPU check of the device tree build's host orchestration (bt_devtree.cpp).
//
// The five tree kernels of bt_tree.hip are emulated here with plain loops of
// the same semantics, and the few HIP runtime calls the orchestration makes
// (allocation, copies, synchronisation; the device-memory cache of
// bt_devmem.cpp on top of them) act on host memory.  emul_compare()
// builds the same trees with the host builder (bt_host.cpp) and with the
// device orchestration, and compares them node by node: ranges, boxes,
// centres, radii, children, levels, leaves, and the particles of every leaf.
// Test infrastructure only; compiled by tests/test_devtree_cpu.py.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <tuple>
#include <vector>

#include "../barytree_amd/csrc/bt_devmem.h"
#include "../barytree_amd/csrc/bt_devtree.h"
#include "../barytree_amd/csrc/bt_host.h"
#include "../barytree_amd/csrc/bt_tree.h"

// ---- the HIP runtime calls of the orchestration, on host memory ----------
static long long g_mallocs = 0, g_frees = 0;
static size_t g_oom_above = (size_t)-1;  // emulated device size limit (emul_devmem)
hipError_t hipMalloc(void **p, size_t n)
{
    if (n > g_oom_above) {
        *p = nullptr;
        return hipErrorOutOfMemory;
    }
    *p = std::malloc(n ? n : 1);
    if (*p) g_mallocs++;
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipFree(void *p)
{
    if (p) g_frees++;
    std::free(p);
    return hipSuccess;
}
hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind, hipStream_t)
{
    std::memmove(d, s, n);
    return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipDeviceSynchronize() { return hipSuccess; }
hipError_t hipGetDevice(int *d)
{
    *d = 0;
    return hipSuccess;
}
hipError_t hipGetLastError() { return hipSuccess; }
const char *hipGetErrorString(hipError_t) { return "emulated error"; }

// ---- the kernels ---------------------------------------------------------
namespace btk {

static bool goes_low(const TreeSeg &s, const double *x, const double *y, const double *z, int i)
{
    if (s.dim < 0) return true;
    const double a = s.dim == 0 ? x[i] : (s.dim == 1 ? y[i] : z[i]);
    return s.strict ? (a < s.mid) : (a <= s.mid);
}

hipError_t launch_tree_bbox(int nchunks, const TreeChunk *chunks, const double *x, const double *y,
                            const double *z, double *part, hipStream_t)
{
    for (int c = 0; c < nchunks; c++) {
        const TreeChunk k = chunks[c];
        if (k.cnt < 1 || k.cnt > TREE_CHUNK) return hipErrorInvalidValue;
        double v[6] = {x[k.beg], x[k.beg], y[k.beg], y[k.beg], z[k.beg], z[k.beg]};
        for (int i = k.beg; i < k.beg + k.cnt; i++) {
            v[0] = std::min(v[0], x[i]); v[1] = std::max(v[1], x[i]);
            v[2] = std::min(v[2], y[i]); v[3] = std::max(v[3], y[i]);
            v[4] = std::min(v[4], z[i]); v[5] = std::max(v[5], z[i]);
        }
        std::memcpy(part + 6 * c, v, sizeof v);
    }
    return hipSuccess;
}

hipError_t launch_tree_count(int nchunks, const TreeChunk *chunks, const TreeSeg *segs, const double *x,
                             const double *y, const double *z, int *low, hipStream_t)
{
    for (int c = 0; c < nchunks; c++) {
        const TreeChunk k = chunks[c];
        int n = 0;
        for (int i = k.beg; i < k.beg + k.cnt; i++) n += goes_low(segs[k.seg], x, y, z, i);
        low[c] = n;
    }
    return hipSuccess;
}

hipError_t launch_tree_scatter(int nchunks, const TreeChunk *chunks, const TreeSeg *segs, const double *x,
                               const double *y, const double *z, const int *idx, double *ox, double *oy,
                               double *oz, int *oidx, hipStream_t)
{
    for (int c = 0; c < nchunks; c++) {
        const TreeChunk k = chunks[c];
        const TreeSeg s = segs[k.seg];
        // the chunk must lie in its segment
        if (k.beg < s.beg || k.beg + k.cnt > s.beg + s.cnt) return hipErrorInvalidValue;
        int lo = s.beg + k.lowoff, hi = s.beg + s.lowtot + (k.beg - s.beg - k.lowoff);
        for (int i = k.beg; i < k.beg + k.cnt; i++) {
            const int d = goes_low(s, x, y, z, i) ? lo++ : hi++;
            if (d < s.beg || d >= s.beg + s.cnt) return hipErrorInvalidValue;  // would be out of bounds
            ox[d] = x[i]; oy[d] = y[i]; oz[d] = z[i]; oidx[d] = idx[i];
        }
    }
    return hipSuccess;
}

hipError_t launch_tree_gather(int n, const int *idx, const double *a, double *out_a, const double *b,
                              double fill_b, double *out_b, hipStream_t)
{
    for (int i = 0; i < n; i++) {
        out_a[i] = a[idx[i]];
        if (out_b) out_b[i] = b ? b[idx[i]] : fill_b;
    }
    return hipSuccess;
}

hipError_t launch_iota(int n, int *idx, hipStream_t)
{
    for (int i = 0; i < n; i++) idx[i] = i;
    return hipSuccess;
}

}  // namespace btk

// ---- the comparison ------------------------------------------------------
extern "C" int emul_compare(int n, const double *x, const double *y, const double *z, int nroots,
                            const int *roots, const int *maxpar, int leaves_only, int *nnodes_out)
{
    using namespace bt;
    // device orchestration
    DevParticles p;
    p.n = n;
    std::vector<double> X(x, x + n), Y(y, y + n), Z(z, z + n), X2(n), Y2(n), Z2(n);
    std::vector<int> I(n), I2(n);
    p.x = X.data(); p.y = Y.data(); p.z = Z.data(); p.idx = I.data();
    p.x2 = X2.data(); p.y2 = Y2.data(); p.z2 = Z2.data(); p.idx2 = I2.data();
    btk::launch_iota(n, p.idx, nullptr);
    std::vector<DevRoot> rts;
    for (int r = 0; r < nroots; r++) rts.push_back(DevRoot{roots[r], roots[r + 1], maxpar[r]});
    std::vector<std::vector<NodeRec>> recs;
    std::vector<int> depth;
    std::string e = device_build_trees(p, rts, leaves_only != 0, nullptr, recs, depth);
    if (!e.empty()) {
        // the host builder must refuse the same input (coincident particles)
        for (int r = 0; r < nroots; r++) {
            const int b = roots[r], m = roots[r + 1] - b;
            Particles P;
            P.n = m;
            P.x.assign(x + b, x + b + m); P.y.assign(y + b, y + b + m); P.z.assign(z + b, z + b + m);
            P.q.assign(m, 0.0);
            P.order.resize(m);
            Tree th;
            if (build_tree(th, P, maxpar[r], leaves_only != 0) == e) return 7;
        }
        std::printf("device build: %s\n", e.c_str());
        return 1;
    }
    int total = 0;
    for (int r = 0; r < nroots; r++) {
        const int b = roots[r], m = roots[r + 1] - b;
        // host builder on a private copy of the system
        Particles P;
        P.n = m;
        P.x.assign(x + b, x + b + m); P.y.assign(y + b, y + b + m); P.z.assign(z + b, z + b + m);
        P.q.assign(m, 0.0);
        P.order.resize(m);
        for (int i = 0; i < m; i++) P.order[i] = i;
        Tree th, td;
        if (!build_tree(th, P, maxpar[r], leaves_only != 0).empty()) return 8;
        finish_tree(td, recs[r], depth[r], leaves_only != 0);
        total += td.n;
        if (th.n != td.n || th.max_depth != td.max_depth) {
            std::printf("root %d: nodes %d vs %d, depth %d vs %d\n", r, th.n, td.n, th.max_depth, td.max_depth);
            return 2;
        }
        for (int i = 0; i < th.n; i++) {
            const bool same = th.ibeg[i] == td.ibeg[i] && th.iend[i] == td.iend[i] && th.level[i] == td.level[i] &&
                              th.nchild[i] == td.nchild[i] && th.parent[i] == td.parent[i] &&
                              th.xmin[i] == td.xmin[i] && th.xmax[i] == td.xmax[i] && th.ymin[i] == td.ymin[i] &&
                              th.ymax[i] == td.ymax[i] && th.zmin[i] == td.zmin[i] && th.zmax[i] == td.zmax[i] &&
                              th.xmid[i] == td.xmid[i] && th.ymid[i] == td.ymid[i] && th.zmid[i] == td.zmid[i] &&
                              th.radius[i] == td.radius[i] &&
                              std::equal(th.child.begin() + 8 * i, th.child.begin() + 8 * i + 8,
                                         td.child.begin() + 8 * i);
            if (!same) {
                std::printf("root %d node %d differs: [%d,%d) vs [%d,%d)\n", r, i, th.ibeg[i], th.iend[i], td.ibeg[i],
                            td.iend[i]);
                return 3;
            }
        }
        if (th.leaves != td.leaves || th.levels != td.levels) {
            std::printf("root %d: leaves or levels differ\n", r);
            return 4;
        }
        // the particles of every leaf: the same set (order may differ)
        for (int l : th.leaves) {
            std::vector<std::tuple<double, double, double, int>> a, c;
            for (int i = th.ibeg[l]; i < th.iend[l]; i++) {
                a.emplace_back(P.x[i], P.y[i], P.z[i], P.order[i] + b);
                const int j = b + i;
                c.emplace_back(p.x[j], p.y[j], p.z[j], p.idx[j]);
                // the device arrays carry the caller index of each particle
                if (x[p.idx[j]] != p.x[j] || y[p.idx[j]] != p.y[j] || z[p.idx[j]] != p.z[j]) {
                    std::printf("root %d: index does not follow its particle\n", r);
                    return 5;
                }
            }
            std::sort(a.begin(), a.end());
            std::sort(c.begin(), c.end());
            if (a != c) {
                std::printf("root %d leaf %d: particle sets differ\n", r, l);
                return 6;
            }
        }
    }
    *nnodes_out = total;
    return 0;
}

// The device-memory cache (bt_devmem.cpp) on the emulated allocator: reuse
// of a released block of a close size, no reuse of a much larger one, the
// limit, emptying, and the retry after an out-of-memory allocation.
// Returns 0, or the number of the failed check.
extern "C" int emul_devmem()
{
    btm::set_cache_limit(size_t(32) << 30);
    btm::empty_cache();
    const long long m0 = g_mallocs;
    void *a = btm::alloc(1000);
    if (!a || g_mallocs != m0 + 1) return 1;
    btm::release(a);
    if (btm::cached_bytes() != 1024) return 2;               // 512-byte steps below 1 MiB
    void *b = btm::alloc(900);                                // same rounded size: reused
    if (b != a || g_mallocs != m0 + 1 || btm::cached_bytes() != 0) return 3;
    btm::release(b);
    void *c = btm::alloc(100000);                             // far larger: a new block
    if (c == a || g_mallocs != m0 + 2) return 4;
    btm::release(c);
    void *d = btm::alloc(3u << 20);                           // 2 MiB steps above 1 MiB: 4 MiB
    btm::release(d);
    void *e = btm::alloc((5u << 20) / 2);                     // 2.5 MiB rounds to 4 MiB: reused
    if (e != d) return 5;
    void *f = btm::alloc(1u << 20);                           // 2 MiB: a 4 MiB block is within 2x
    btm::release(e);
    btm::release(f);
    const size_t cached = btm::cached_bytes();
    if (cached != 1024 + 100352 + (size_t(4) << 20) + (size_t(2) << 20)) return 6;
    const long long f0 = g_frees;
    btm::set_cache_limit(size_t(1) << 21);                    // frees the largest blocks first
    if (btm::cached_bytes() > (size_t(1) << 21) || g_frees <= f0) return 7;
    void *g = btm::alloc(size_t(3) << 21);                    // beyond the limit on release: freed
    const long long f1 = g_frees;
    btm::release(g);
    if (g_frees != f1 + 1) return 8;
    btm::set_cache_limit(size_t(32) << 30);
    void *h = btm::alloc(1u << 22);
    btm::release(h);
    g_oom_above = (size_t)1 << 23;                            // the next 12 MiB fails at first
    void *i = btm::alloc(size_t(12) << 20);
    if (i) return 9;                                          // still too large after emptying
    if (btm::cached_bytes() != 0) return 10;                  // the cache was emptied for it
    g_oom_above = (size_t)-1;
    if (btm::empty_cache() != 0) return 11;
    void *j = btm::alloc(4096);
    btm::release(j);
    if (btm::empty_cache() != 4096) return 12;
    return 0;
}
