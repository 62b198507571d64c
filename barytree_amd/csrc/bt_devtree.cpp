// bt_devtree.cpp -- builds the source and target trees on the GPU.
//
// Level by level: the exact bounding box of every node of the level (chunked
// min/max on the device, the chunks' results combined here), the split
// decision of each node on the host with the same code the host builder uses
// (node_split), then one stable partition pass per split dimension: in pass
// r every splitting node splits each of its current ranges along its r-th
// split dimension, exactly as the host builder's loop over x, y, z does, and
// every other particle is copied through.  Between passes the host reads the
// lower-part counts back to place the ranges; a level costs a few
// synchronisations, a pass a read and a write of 28 bytes per particle.
#include "bt_devtree.h"
#include "bt_devmem.h"

#include <algorithm>
#include <cstring>

#include "bt_tree.h"

namespace bt {

namespace {

#define HIPCHK(expr)                                                                 \
    do {                                                                             \
        hipError_t e_ = (expr);                                                      \
        if (e_ != hipSuccess) return std::string(#expr ": ") + hipGetErrorString(e_); \
    } while (0)

constexpr int CH = btk::TREE_CHUNK;

// Device buffer that grows and is freed with its owner.
template <class T> struct Grow {
    T *p = nullptr;
    size_t cap = 0;
    ~Grow() { drop(); }
    // the buffer's work on the device is finished before it goes back to
    // the device-memory cache
    void drop()
    {
        if (!p) return;
        (void)hipDeviceSynchronize();
        btm::release(p);
        p = nullptr;
        cap = 0;
    }
    hipError_t fit(size_t n)
    {
        if (n <= cap) return hipSuccess;
        drop();
        p = static_cast<T *>(btm::alloc(std::max<size_t>(n, 1) * sizeof(T)));
        if (!p) return hipErrorOutOfMemory;
        cap = n;
        return hipSuccess;
    }
};

// A node found by the level-by-level build.
struct BNode {
    int b, e, level, root;
    NodeRec rec;
    std::vector<int> kids;  // BNode ids, in child order
};

// A node that splits at the current level: its ranges as the host builder
// keeps them (rb/re, nr of them), and its split dimensions.
struct Split {
    int node;
    int dv[3];
    int nr;
    int rb[8], re[8];
};

void add_chunks(std::vector<btk::TreeChunk> &ch, int seg, int b, int e)
{
    for (int o = b; o < e; o += CH) ch.push_back(btk::TreeChunk{seg, o, std::min(CH, e - o), 0});
}

}  // namespace

std::vector<NodeRec> leaf_records(const std::vector<NodeRec> &full)
{
    std::vector<NodeRec> out;
    for (NodeRec r : full)
        if (r.nchild == 0) out.push_back(r);
    return out;
}

std::string device_build_trees(DevParticles &p, const std::vector<DevRoot> &roots, bool leaves_only,
                               hipStream_t st, std::vector<std::vector<NodeRec>> &nodes,
                               std::vector<int> &depth)
{
    std::vector<BNode> all;
    std::vector<int> cur;
    for (size_t r = 0; r < roots.size(); r++) {
        if (roots[r].e <= roots[r].b) return "device tree build: empty system";
        BNode n{};
        n.b = roots[r].b; n.e = roots[r].e; n.level = 0; n.root = (int)r;
        all.push_back(n);
        cur.push_back((int)all.size() - 1);
    }
    Grow<btk::TreeChunk> d_chunks;
    Grow<btk::TreeSeg> d_segs;
    Grow<double> d_part;
    Grow<int> d_low;
    std::vector<btk::TreeChunk> chunks;
    std::vector<btk::TreeSeg> segs;
    std::vector<double> part;
    std::vector<int> low;

    while (!cur.empty()) {
        // 1. exact bounding boxes of the level's nodes
        chunks.clear();
        for (size_t k = 0; k < cur.size(); k++) add_chunks(chunks, (int)k, all[cur[k]].b, all[cur[k]].e);
        HIPCHK(d_chunks.fit(chunks.size()));
        HIPCHK(d_part.fit(6 * chunks.size()));
        HIPCHK(hipMemcpyAsync(d_chunks.p, chunks.data(), chunks.size() * sizeof(btk::TreeChunk),
                              hipMemcpyHostToDevice, st));
        HIPCHK(btk::launch_tree_bbox((int)chunks.size(), d_chunks.p, p.x, p.y, p.z, d_part.p, st));
        part.resize(6 * chunks.size());
        HIPCHK(hipMemcpyAsync(part.data(), d_part.p, part.size() * sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        std::vector<int> first(cur.size(), -1);
        for (size_t c = 0; c < chunks.size(); c++) {
            NodeRec &r = all[cur[chunks[c].seg]].rec;
            const double *v = &part[6 * c];
            if (first[chunks[c].seg] < 0) {
                first[chunks[c].seg] = (int)c;
                r.lo[0] = v[0]; r.hi[0] = v[1]; r.lo[1] = v[2]; r.hi[1] = v[3]; r.lo[2] = v[4]; r.hi[2] = v[5];
                continue;
            }
            // min and max are exact, so the chunk order does not matter
            if (v[0] < r.lo[0]) r.lo[0] = v[0];
            if (v[1] > r.hi[0]) r.hi[0] = v[1];
            if (v[2] < r.lo[1]) r.lo[1] = v[2];
            if (v[3] > r.hi[1]) r.hi[1] = v[3];
            if (v[4] < r.lo[2]) r.lo[2] = v[4];
            if (v[5] > r.hi[2]) r.hi[2] = v[5];
        }

        // 2. split decisions (the host builder's rule)
        std::vector<Split> splits;
        for (int id : cur) {
            BNode &n = all[id];
            n.rec.b = n.b; n.rec.e = n.e; n.rec.level = n.level; n.rec.nchild = 0;
            Split s{};
            if (!node_split(n.e - n.b, roots[n.root].maxpar, n.rec, s.dv)) continue;
            s.node = id; s.nr = 1; s.rb[0] = n.b; s.re[0] = n.e;
            splits.push_back(s);
        }
        if (splits.empty()) break;

        // 3. one partition pass per split dimension (the r-th set dimension
        // of each node in pass r), copying every other particle through
        for (int pass = 0; pass < 3; pass++) {
            segs.clear();
            std::vector<std::pair<int, int>> owner;  // (split, range) of each split segment
            std::vector<int> sdim(splits.size(), -1);
            for (size_t k = 0; k < splits.size(); k++) {
                int seen = 0;
                for (int d = 0; d < 3; d++)
                    if (splits[k].dv[d] && seen++ == pass) sdim[k] = d;
                if (sdim[k] < 0) continue;
                const Split &s = splits[k];
                const double mid = all[s.node].rec.mid[sdim[k]];
                for (int i = 0; i < s.nr; i++) {
                    const int cnt = s.re[i] - s.rb[i];
                    if (cnt <= 0) continue;
                    segs.push_back(btk::TreeSeg{mid, s.rb[i], cnt, sdim[k], cnt == 1 ? 1 : 0, 0, 0});
                    owner.push_back({(int)k, i});
                }
            }
            // no node splits in this pass (every splitting node has fewer
            // dimensions); a splitting node always has a non-empty range
            if (owner.empty()) continue;
            // pass-through segments over the gaps between split segments
            std::vector<int> order(segs.size());
            for (size_t i = 0; i < order.size(); i++) order[i] = (int)i;
            std::sort(order.begin(), order.end(), [&](int a, int b) { return segs[a].beg < segs[b].beg; });
            const size_t nsplit = segs.size();
            int at = 0;
            for (int i : order) {
                if (segs[i].beg > at) segs.push_back(btk::TreeSeg{0.0, at, segs[i].beg - at, -1, 0, 0, 0});
                at = segs[i].beg + segs[i].cnt;
            }
            if (at < p.n) segs.push_back(btk::TreeSeg{0.0, at, p.n - at, -1, 0, 0, 0});
            chunks.clear();
            for (size_t i = 0; i < segs.size(); i++) add_chunks(chunks, (int)i, segs[i].beg, segs[i].beg + segs[i].cnt);
            HIPCHK(d_segs.fit(segs.size()));
            HIPCHK(d_chunks.fit(chunks.size()));
            HIPCHK(d_low.fit(chunks.size()));
            HIPCHK(hipMemcpyAsync(d_segs.p, segs.data(), segs.size() * sizeof(btk::TreeSeg), hipMemcpyHostToDevice,
                                  st));
            HIPCHK(hipMemcpyAsync(d_chunks.p, chunks.data(), chunks.size() * sizeof(btk::TreeChunk),
                                  hipMemcpyHostToDevice, st));
            HIPCHK(btk::launch_tree_count((int)chunks.size(), d_chunks.p, d_segs.p, p.x, p.y, p.z, d_low.p, st));
            low.resize(chunks.size());
            HIPCHK(hipMemcpyAsync(low.data(), d_low.p, low.size() * sizeof(int), hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            // each chunk's place in its segment's lower part, each segment's total
            for (auto &s : segs) s.lowtot = 0;
            for (size_t c = 0; c < chunks.size(); c++) {
                btk::TreeSeg &s = segs[chunks[c].seg];
                chunks[c].lowoff = s.lowtot;
                s.lowtot += low[c];
            }
            HIPCHK(hipMemcpyAsync(d_segs.p, segs.data(), segs.size() * sizeof(btk::TreeSeg), hipMemcpyHostToDevice,
                                  st));
            HIPCHK(hipMemcpyAsync(d_chunks.p, chunks.data(), chunks.size() * sizeof(btk::TreeChunk),
                                  hipMemcpyHostToDevice, st));
            HIPCHK(btk::launch_tree_scatter((int)chunks.size(), d_chunks.p, d_segs.p, p.x, p.y, p.z, p.idx, p.x2,
                                            p.y2, p.z2, p.idx2, st));
            // the host vectors are reused in the next pass: let the copies finish
            HIPCHK(hipStreamSynchronize(st));
            std::swap(p.x, p.x2); std::swap(p.y, p.y2); std::swap(p.z, p.z2); std::swap(p.idx, p.idx2);
            // the ranges after the pass, as the host builder arranges them:
            // range i keeps its lower part, range nr + i gets the upper part
            std::vector<int> mid_at(nsplit);
            for (size_t i = 0; i < nsplit; i++) mid_at[i] = segs[i].beg + segs[i].lowtot;
            std::vector<std::vector<int>> m(splits.size());
            for (size_t k = 0; k < splits.size(); k++)
                if (sdim[k] >= 0) {
                    m[k].resize(splits[k].nr);
                    for (int i = 0; i < splits[k].nr; i++) m[k][i] = splits[k].rb[i];  // empty ranges
                }
            for (size_t i = 0; i < nsplit; i++) m[owner[i].first][owner[i].second] = mid_at[i];
            for (size_t k = 0; k < splits.size(); k++) {
                if (sdim[k] < 0) continue;
                Split &s = splits[k];
                for (int i = 0; i < s.nr; i++) {
                    s.rb[s.nr + i] = m[k][i]; s.re[s.nr + i] = s.re[i]; s.re[i] = m[k][i];
                }
                s.nr *= 2;
            }
        }

        // 4. children: the non-empty ranges, in range order
        std::vector<int> next;
        for (const Split &s : splits) {
            int nonempty = 0;
            for (int i = 0; i < s.nr; i++) nonempty += s.rb[i] < s.re[i];
            if (nonempty == 1) return COINCIDENT_ERROR;
            for (int i = 0; i < s.nr; i++) {
                if (s.rb[i] >= s.re[i]) continue;
                BNode c{};
                c.b = s.rb[i]; c.e = s.re[i]; c.level = all[s.node].level + 1; c.root = all[s.node].root;
                all.push_back(c);
                const int cid = (int)all.size() - 1;
                all[s.node].kids.push_back(cid);
                next.push_back(cid);
            }
        }
        cur.swap(next);
    }

    // preorder records per root, ranges relative to the root
    nodes.assign(roots.size(), {});
    depth.assign(roots.size(), 0);
    // index of every node's record within its root's records
    std::vector<int> rec_of(all.size(), -1);
    const int nroots = (int)roots.size();
#pragma omp parallel for schedule(dynamic) if (nroots > 1)
    for (int r = 0; r < nroots; r++) {
        std::vector<NodeRec> &out = nodes[r];
        const int base = roots[r].b;
        // iterative DFS over this root's nodes (ids in preorder)
        std::vector<int> stack = {r}, order;
        while (!stack.empty()) {
            const int id = stack.back();
            stack.pop_back();
            const BNode &n = all[id];
            depth[r] = std::max(depth[r], n.level + 1);
            NodeRec rec = n.rec;
            rec.b = n.b - base; rec.e = n.e - base; rec.nchild = 0;
            out.push_back(rec);
            rec_of[id] = (int)out.size() - 1;
            order.push_back(id);
            for (size_t k = n.kids.size(); k-- > 0;) stack.push_back(n.kids[k]);
        }
        // child links: a node's children follow in preorder
        for (int id : order) {
            NodeRec &pr = out[rec_of[id]];
            for (int c : all[id].kids) pr.child[pr.nchild++] = rec_of[c];
        }
        if (leaves_only) out = leaf_records(out);
    }
    return "";
}

}  // namespace bt
