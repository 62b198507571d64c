// bt_solver.cpp -- plan (host setup + upload) and execute (GPU hot path).
#include "bt_solver.h"

#include "bt_devmem.h"
#include "bt_devtree.h"
#include "bt_noalloc.h"
#include "bt_tree.h"

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <parallel/algorithm>

namespace bt {

namespace {

double now()
{
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

#define HIPCHK(expr)                                                                 \
    do {                                                                             \
        hipError_t e_ = (expr);                                                      \
        if (e_ != hipSuccess) return std::string(#expr ": ") + hipGetErrorString(e_); \
    } while (0)

constexpr int NXCD = btk::TILE_QUEUES;  // XCDs of an MI355X, one tile queue each
// upward pass: leaves are split into parts when there are fewer than this
// many (four blocks per CU), parts of at least UP_PART_MIN particles
constexpr int UP_SPLIT_BLOCKS = 1024;
constexpr int UP_PART_MIN = 256;

}  // namespace

// Tiles.  All tiles of one key read the same sources, so they should run at
// the same time on the same XCD and share its L2.  Workgroups are dispatched
// round-robin over the 8 XCDs (block b runs on XCD b % 8), so keys (groups of
// tiles) are dealt to per-XCD lists, heaviest group to the least loaded XCD,
// and block 8*j + x takes the j-th tile of XCD x's list.  Lists are padded
// with empty tiles (count 0), which return at once.
namespace {
long long cluster_points(int id, int P, bool multi)
{
    if (!multi) return P;
    const long long nl = (id >> btk::DEG_SHIFT) + 1;
    return nl * nl * nl;
}
}  // namespace

void split_keys(const WorkCSR &in, long long max_src, int P, bool multi, WorkCSR &out)
{
    out = WorkCSR();
    out.seg_off.push_back(0);
    out.clu_off.push_back(0);
    out.pairs = in.pairs;
    auto open_key = [&](size_t k) {
        out.key.push_back(in.key[k]); out.tbeg.push_back(in.tbeg[k]); out.tcnt.push_back(in.tcnt[k]);
    };
    auto close_key = [&]() {
        out.seg_off.push_back((int)out.seg_beg.size());
        out.clu_off.push_back((int)out.clu_id.size());
    };
    for (size_t k = 0; k < in.key.size(); k++) {
        open_key(k);
        long long load = 0;
        bool empty = true;
        for (int c = in.clu_off[k]; c < in.clu_off[k + 1]; c++) {
            const long long w = cluster_points(in.clu_id[c], P, multi);
            if (!empty && load + w > max_src) { close_key(); open_key(k); load = 0; }
            out.clu_id.push_back(in.clu_id[c]);
            load += w;
            empty = false;
        }
        for (int s = in.seg_off[k]; s < in.seg_off[k + 1]; s++) {
            const long long w = in.seg_cnt[s];
            if (!empty && load + w > max_src) { close_key(); open_key(k); load = 0; }
            out.seg_beg.push_back(in.seg_beg[s]);
            out.seg_cnt.push_back(in.seg_cnt[s]);
            out.seg_chk.push_back(in.seg_chk[s]);
            load += w;
            empty = false;
        }
        close_key();
    }
}

void make_queue_tiles(const WorkCSR &w, int tile_t, bool grid, int P, std::vector<btk::Tile> &tiles,
                      int *qlen, bool multi)
{
    struct Group { long long work; std::vector<btk::Tile> t; };
    std::vector<Group> groups;
    for (size_t k = 0; k < w.key.size(); k++) {
        long long src = 0;
        for (int s = w.seg_off[k]; s < w.seg_off[k + 1]; s++) src += w.seg_cnt[s];
        for (int c = w.clu_off[k]; c < w.clu_off[k + 1]; c++) src += cluster_points(w.clu_id[c], P, multi);
        Group g{0, {}};
        for (int o = 0; o < w.tcnt[k]; o += tile_t) {
            btk::Tile t;
            t.x = (int)k;
            t.y = w.tbeg[k] + o;
            t.z = std::min(tile_t, w.tcnt[k] - o);
            t.w = grid ? w.key[k] : 0;
            g.work += (src + 1) * t.z;
            g.t.push_back(t);
        }
        if (!g.t.empty()) groups.push_back(std::move(g));
    }
    std::stable_sort(groups.begin(), groups.end(),
                     [](const Group &a, const Group &b) { return a.work > b.work; });
    std::vector<std::vector<btk::Tile>> lists(NXCD);
    std::vector<long long> load(NXCD, 0);
    for (const Group &g : groups) {
        int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        load[x] += g.work;
        lists[x].insert(lists[x].end(), g.t.begin(), g.t.end());
    }
    size_t L = 0;
    for (auto &l : lists) L = std::max(L, l.size());
    tiles.assign(L * NXCD, btk::Tile{0, 0, 0, 0});
    for (int x = 0; x < NXCD; x++)
        for (size_t j = 0; j < lists[x].size(); j++) tiles[j * NXCD + x] = lists[x][j];
    for (int x = 0; x < NXCD; x++) qlen[x] = (int)lists[x].size();
}

void make_pencil_tiles(const WorkCSR &w, int P, bool multi, std::vector<btk::Tile> &tiles)
{
    std::vector<std::pair<long long, btk::Tile>> ks;
    for (size_t k = 0; k < w.key.size(); k++) {
        long long src = 0;
        for (int s = w.seg_off[k]; s < w.seg_off[k + 1]; s++) src += w.seg_cnt[s];
        for (int c = w.clu_off[k]; c < w.clu_off[k + 1]; c++) src += cluster_points(w.clu_id[c], P, multi);
        ks.push_back({(src + 1) * w.tcnt[k], btk::Tile{(int)k, 0, w.tcnt[k], w.key[k]}});
    }
    std::stable_sort(ks.begin(), ks.end(), [](const auto &a, const auto &b) { return a.first > b.first; });
    tiles.clear();
    for (const auto &k : ks) tiles.push_back(k.second);
}

Plan::~Plan() { free_all(); }

void Plan::free_all()
{
    // the plan's work is on its stream: once it has finished, the buffers
    // go back to the device-memory cache
    if (swapped_) { std::swap(st_, aux_); swapped_ = false; }
    if (st_) (void)hipStreamSynchronize(st_);
    if (aux_) (void)hipStreamSynchronize(aux_);
    free_var();
    for (void *p : allocs_) btm::release(p);
    allocs_.clear();
    for (hipEvent_t *e : {&fork_ev_, &join_ev_, &red_ev_})
        if (*e) { (void)hipEventDestroy(*e); *e = nullptr; }
    if (aux_) {
        (void)hipStreamDestroy(aux_);
        aux_ = nullptr;
    }
    for (hipEvent_t &e : sym_ev_)
        if (e) { (void)hipEventDestroy(e); e = nullptr; }
    for (hipEvent_t &e : ev_)
        if (e) { (void)hipEventDestroy(e); e = nullptr; }
    if (st_) {
        (void)hipStreamDestroy(st_);
        st_ = nullptr;
    }
}

// The phase events of execute(), created at the first execution and kept
// until the plan is freed (an execution that fails part way leaks nothing).
std::string Plan::phase_events()
{
    for (hipEvent_t &e : ev_)
        if (!e) HIPCHK(hipEventCreate(&e));
    return "";
}

template <class T> T *Plan::dalloc(size_t n)
{
    void *p = btm::alloc(std::max<size_t>(n, 1) * sizeof(T));
    if (!p) return nullptr;
    (alloc_var_ ? var_allocs_ : allocs_).push_back(p);
    return static_cast<T *>(p);
}

void Plan::dfree(void *p)
{
    if (!p) return;
    for (auto *v : {&allocs_, &var_allocs_}) {
        auto it = std::find(v->begin(), v->end(), p);
        if (it != v->end()) {
            if (st_) (void)hipStreamSynchronize(st_);
            btm::release(p);
            v->erase(it);
            return;
        }
    }
}

template <class T, class A> std::string Plan::up(T **d, const std::vector<T, A> &h)
{
    *d = dalloc<T>(h.size());
    if (!*d) return "hipMalloc failed";
    return up_at(*d, h);
}

template <class T, class A> std::string Plan::up_at(T *d, const std::vector<T, A> &h)
{
    if (!h.empty()) HIPCHK(hipMemcpyAsync(d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, st_));
    return "";
}

bool Plan::pencil_grids() const
{
    const Params &p = prm_;
    // FP64 forms displacements; mixed uses the centred expansion, which needs
    // acceptance-separated sources (theta < 0.99)
    return (p.precision != PREC_MIXED || p.theta < 0.99) &&
           (p.kernel == K_COULOMB || p.kernel == K_YUKAWA ||
            (p.kernel == K_REG_COULOMB && p.singularity == S_SKIPPING));
}

// Symmetric near field of one system whose targets are its sources (the same
// particles in the same order after the tree builds).  Keys are target
// nodes (leaves or batches, disjoint particle ranges) with the source nodes
// of their near lists (segs_of, indexed by target node).  A source node whose
// particle range is another key's range, while that key lists this key's
// range, makes a symmetric pair: one two-sided item.  Every other source
// node (one that does not list the key back, and the key's own range) makes
// a one-sided item.  Every range is flagged SEG_SYM in sym_of (parallel to
// segs_of): the near launch leaves the particle pairs to the items.  The
// item's centre c is the centre of the two boxes' joint box and R the
// largest distance from c to either box: r^2 is formed as in the near keys
// (|t-c|^2 + |s-c|^2 - 2 (t-c).(s-c), rounding error below
// 6e-15 (|t-c|^2 + |s-c|^2)), so pairs below 2e-7 R^2 are recomputed from
// differences when the boxes come within that distance.
void Plan::build_sym(const System &sy, const std::vector<int> &keys,
                     const std::vector<std::vector<int>> &segs_of, std::vector<std::vector<char>> &sym_of,
                     bool symmetric)
{
    const Tree &st = sy.stree, &tt = sy.ttree;
    sym_items_.clear(); sym_ctr_.clear(); sym_red_tiles_.clear();
    sym_red_off_.clear(); sym_red_base_.clear();
    sym_slots_len_ = 0; sym_pairs_ = 0; sym_evals_ = 0;
    const int nk = (int)keys.size();
    // key index of every source node whose particle range is a key's range
    // (-1 otherwise); keys are disjoint leaves, so a range names one key
    std::vector<int> key_of_node(st.n, -1);
    if (symmetric) {
        std::vector<std::array<long long, 2>> key_range(nk);  // (beg << 32 | cnt, key index)
        for (int k = 0; k < nk; k++)
            key_range[k] = {((long long)tt.ibeg[keys[k]] << 32) | (unsigned)tt.numpar[keys[k]], (long long)k};
        std::sort(key_range.begin(), key_range.end());
#pragma omp parallel for schedule(static)
        for (int s = 0; s < st.n; s++) {
            const long long r = ((long long)st.ibeg[s] << 32) | (unsigned)st.numpar[s];
            auto it = std::lower_bound(key_range.begin(), key_range.end(), std::array<long long, 2>{r, -1});
            if (it != key_range.end() && (*it)[0] == r) key_of_node[s] = (int)(*it)[1];
        }
    }
    // per key: the keys among its source ranges, sorted (does ks list k back?)
    std::vector<std::vector<int>> nbr(nk);
    if (symmetric) {
#pragma omp parallel for schedule(dynamic, 64)
        for (int k = 0; k < nk; k++) {
            for (int s : segs_of[keys[k]])
                if (key_of_node[s] >= 0) nbr[k].push_back(key_of_node[s]);
            std::sort(nbr[k].begin(), nbr[k].end());
        }
    }
    auto has = [&](int a, int b) { return std::binary_search(nbr[a].begin(), nbr[a].end(), b); };
    // kb >= 0: two-sided (the pair of keys k < kb), kb < 0: one-sided;
    // candidates in key order, each key's ranges in list order
    struct Cand { int k, kb, s; };
    std::vector<int> first(nk + 1, 0);
#pragma omp parallel for schedule(dynamic, 64)
    for (int k = 0; k < nk; k++) {
        const int t = keys[k];
        sym_of[t].assign(segs_of[t].size(), 1);
        int c = 0;
        for (int s : segs_of[t]) {
            const int ks = key_of_node[s];
            const bool two = ks >= 0 && ks != k && has(ks, k);
            c += (!two || k < ks) ? 1 : 0;
        }
        first[k + 1] = c;
    }
    for (int k = 0; k < nk; k++) first[k + 1] += first[k];
    std::vector<Cand> cand(first[nk]);
#pragma omp parallel for schedule(dynamic, 64)
    for (int k = 0; k < nk; k++) {
        int c = first[k];
        for (int s : segs_of[keys[k]]) {
            const int ks = key_of_node[s];
            const bool two = ks >= 0 && ks != k && has(ks, k);
            if (!two) cand[c++] = {k, -1, s};
            else if (k < ks) cand[c++] = {k, ks, s};
        }
    }
    // heaviest first, ties in candidate order (a two-sided pair costs about
    // twice a one-sided one)
    const size_t nc = cand.size();
    {
        std::vector<std::pair<long long, int>> order(nc);
#pragma omp parallel for schedule(static)
        for (size_t i = 0; i < nc; i++) {
            const Cand &c = cand[i];
            order[i] = {-(long long)tt.numpar[keys[c.k]] * st.numpar[c.s] * (c.kb >= 0 ? 2 : 1), (int)i};
        }
        __gnu_parallel::sort(order.begin(), order.end());
        std::vector<Cand> sorted(nc);
#pragma omp parallel for schedule(static)
        for (size_t i = 0; i < nc; i++) sorted[i] = cand[order[i].second];
        cand.swap(sorted);
    }
    auto work = [&](const Cand &c) {
        return (long long)tt.numpar[keys[c.k]] * st.numpar[c.s] * (c.kb >= 0 ? 2 : 1);
    };
    auto far2 = [](double cx, double cy, double cz, const Tree &T, int n) {
        const double dx = std::max(std::fabs(T.xmin[n] - cx), std::fabs(T.xmax[n] - cx));
        const double dy = std::max(std::fabs(T.ymin[n] - cy), std::fabs(T.ymax[n] - cy));
        const double dz = std::max(std::fabs(T.zmin[n] - cz), std::fabs(T.zmax[n] - cz));
        return dx * dx + dy * dy + dz * dz;
    };
    // slots in candidate order: A's particles, then B's for a two-sided item
    std::vector<long long> slot0(nc + 1, 0);
    for (size_t i = 0; i < nc; i++) {
        const Cand &c = cand[i];
        slot0[i + 1] = slot0[i] + tt.numpar[keys[c.k]] + (c.kb >= 0 ? st.numpar[c.s] : 0);
    }
    sym_slots_len_ = slot0[nc];
    sym_items_.resize(nc);
    sym_ctr_.resize(4 * nc);
    long long pairs = 0, evals = 0;
#pragma omp parallel for schedule(static) reduction(+ : pairs, evals)
    for (size_t i = 0; i < nc; i++) {
        const Cand &c = cand[i];
        const int t = keys[c.k], s = c.s;
        btk::SymItem it;
        // A: targets, B: sources (symmetric plans: the same particles)
        it.a_beg = sy.toff + tt.ibeg[t]; it.a_cnt = tt.numpar[t];
        it.b_beg = sy.soff + st.ibeg[s]; it.b_cnt = st.numpar[s];
        it.rslot = slot0[i];
        it.cslot = c.kb >= 0 ? slot0[i] + it.a_cnt : -1;
        const double cx = 0.5 * (std::min(tt.xmin[t], st.xmin[s]) + std::max(tt.xmax[t], st.xmax[s]));
        const double cy = 0.5 * (std::min(tt.ymin[t], st.ymin[s]) + std::max(tt.ymax[t], st.ymax[s]));
        const double cz = 0.5 * (std::min(tt.zmin[t], st.zmin[s]) + std::max(tt.zmax[t], st.zmax[s]));
        const double R2 = std::max(far2(cx, cy, cz, tt, t), far2(cx, cy, cz, st, s));
        const double thresh = std::max(2e-7 * R2, 1e-300);
        const double g = std::max(1e-150, std::sqrt(thresh));
        const bool close = tt.xmin[t] - st.xmax[s] <= g && st.xmin[s] - tt.xmax[t] <= g &&
                           tt.ymin[t] - st.ymax[s] <= g && st.ymin[s] - tt.ymax[t] <= g &&
                           tt.zmin[t] - st.zmax[s] <= g && st.zmin[s] - tt.zmax[t] <= g;
        double *ctr = sym_ctr_.data() + 4 * i;
        ctr[0] = cx; ctr[1] = cy; ctr[2] = cz; ctr[3] = close ? -thresh : thresh;
        pairs += (c.kb >= 0 ? 2LL : 1LL) * it.a_cnt * it.b_cnt;
        evals += (long long)it.a_cnt * it.b_cnt;
        sym_items_[i] = it;
    }
    sym_pairs_ = pairs;
    sym_evals_ = evals;
    // per key, the bases of its slots in candidate order
    std::vector<std::vector<long long>> lists(nk);
    for (size_t i = 0; i < nc; i++) {
        const Cand &c = cand[i];
        const btk::SymItem &it = sym_items_[i];
        lists[c.k].push_back(it.rslot - it.a_beg);
        if (c.kb >= 0) lists[c.kb].push_back(it.cslot - it.b_beg);
    }
    // Launch order.  Workgroups are dealt round-robin over the 8 XCDs, so
    // block b runs on XCD b % 8.  A block streams its B leaf once per pass
    // over A; items that share a B leaf go to one XCD back to back, so their
    // passes re-read B from that XCD's L2 rather than from HBM.  Groups
    // (heaviest first; ties by first particle of B) go to the least loaded
    // XCD; the queues are interleaved (block 8j + x: item j of XCD x) and
    // padded with empty items.  Slots and sums do not depend on the order.
    {
        std::vector<std::pair<int, int>> by_b(nc);  // (first particle of B, item)
        for (size_t i = 0; i < nc; i++) by_b[i] = {sym_items_[i].b_beg, (int)i};
        __gnu_parallel::sort(by_b.begin(), by_b.end());
        struct Group { long long w; size_t beg, end; };
        std::vector<Group> groups;
        for (size_t i = 0; i < nc;) {
            size_t e = i;
            long long w = 0;
            while (e < nc && by_b[e].first == by_b[i].first) w += work(cand[by_b[e++].second]);
            groups.push_back({w, i, e});
            i = e;
        }
        std::stable_sort(groups.begin(), groups.end(), [](const Group &a, const Group &b) { return a.w > b.w; });
        std::vector<std::vector<int>> q(NXCD);
        long long load[NXCD] = {};
        for (const Group &g : groups) {
            const int x = (int)(std::min_element(load, load + NXCD) - load);
            load[x] += g.w;
            for (size_t e = g.beg; e < g.end; e++) q[x].push_back(by_b[e].second);
        }
        size_t qmax = 0;
        for (const auto &v : q) qmax = std::max(qmax, v.size());
        std::vector<btk::SymItem> items(qmax * NXCD, btk::SymItem{0, 0, 0, 0, 0, 0});
        std::vector<double> ctr(4 * qmax * NXCD, 0.0);
        for (int x = 0; x < NXCD; x++) {
            const int len = (int)q[x].size();
#pragma omp parallel for schedule(static)
            for (int j = 0; j < len; j++) {
                const size_t slot = (size_t)j * NXCD + x;
                items[slot] = sym_items_[q[x][j]];
                std::copy(sym_ctr_.begin() + 4 * (size_t)q[x][j], sym_ctr_.begin() + 4 * (size_t)q[x][j] + 4,
                          ctr.begin() + 4 * slot);
            }
        }
        sym_items_.swap(items);
        sym_ctr_.swap(ctr);
    }
    sym_red_off_.assign(1, 0);
    for (size_t k = 0; k < keys.size(); k++) {
        const int t = keys[k];
        sym_red_base_.insert(sym_red_base_.end(), lists[k].begin(), lists[k].end());
        sym_red_off_.push_back((int)sym_red_base_.size());
        if (lists[k].empty()) continue;
        for (int o = 0; o < tt.numpar[t]; o += btk::TPB_SYM_REDUCE)
            sym_red_tiles_.push_back(btk::Tile{(int)k, sy.toff + tt.ibeg[t] + o,
                                               std::min(btk::TPB_SYM_REDUCE, tt.numpar[t] - o), 0});
    }
    sym_ = !sym_items_.empty();
}

// Flattens the lists of every system into two global CSR work descriptions:
//   near_ : target leaves (particle targets) <- source particle ranges + source clusters
//   grid_ : target nodes (grid targets)      <- source particle ranges + source clusters
// with particle indices and node ids offset to the concatenated numbering.
void Plan::build_work()
{
    const Params &p = prm_;
    const int P = p.P;
    auto add_key = [](WorkCSR &w, int key, int tbeg, int tcnt) {
        w.key.push_back(key); w.tbeg.push_back(tbeg); w.tcnt.push_back(tcnt);
    };
    auto close_key = [](WorkCSR &w) {
        w.seg_off.push_back((int)w.seg_beg.size());
        w.clu_off.push_back((int)w.clu_id.size());
    };
    auto add_seg = [](WorkCSR &w, int b, int c, long long nt, int flags = 0) {
        w.seg_beg.push_back(b); w.seg_cnt.push_back(c); w.seg_chk.push_back((char)flags);
        w.pairs += nt * c;
    };
    auto add_clu = [P](WorkCSR &w, int id, long long nt) { w.clu_id.push_back(id); w.pairs += nt * P; };
    near_ = WorkCSR();
    grid_ = WorkCSR();
    near_ctr_.clear();
    sym_ = false;
    // items for the near field's particle pairs (one system)
    const bool items = sym_ok_ || items_ok_;
    // the work of one system into (near keys, grid keys, near centres)
    auto per_system = [&](const System &sy, WorkCSR &nw, WorkCSR &gw, std::vector<double> &nctr) {
        const Tree &st = sy.stree, &tt = sy.ttree;
        const Lists &L = sy.L;
        const int so = sy.soff, to = sy.toff, sn = sy.snode_off, tn = sy.tnode_off;
        // Mixed near field: r^2 is formed from coordinates relative to the
        // key's centre c (its target node's centre) as
        // |t-c|^2 + |s-c|^2 - 2 (t-c).(s-c), on the FP64 matrix cores or the
        // FP64 VALU, with a rounding error below 6e-15 (|t-c|^2 + |s-c|^2).
        // R is the largest distance from c to the key's target box and source
        // particle boxes: pairs with r^2 >= thresh = 2e-7 R^2 keep a relative
        // error below 6e-8 (one FP32 rounding), closer pairs are recomputed
        // from coordinate differences.  The acceptance criterion keeps source
        // clusters away from the targets; a cluster that is not far enough
        // (6e-15 (|t-c|^2 + |s-c|^2) above 6e-8 of its closest approach^2)
        // widens R to its far corner, and the key's clusters are then checked
        // too (the key's thresh is stored negated).
        auto far2 = [](double cx, double cy, double cz, const Tree &T, int n) {
            const double dx = std::max(std::fabs(T.xmin[n] - cx), std::fabs(T.xmax[n] - cx));
            const double dy = std::max(std::fabs(T.ymin[n] - cy), std::fabs(T.ymax[n] - cy));
            const double dz = std::max(std::fabs(T.zmin[n] - cz), std::fabs(T.zmax[n] - cz));
            return dx * dx + dy * dy + dz * dz;
        };
        auto gap2 = [&](int t, int s) {
            const double gx = std::max(0.0, std::max(tt.xmin[t] - st.xmax[s], st.xmin[s] - tt.xmax[t]));
            const double gy = std::max(0.0, std::max(tt.ymin[t] - st.ymax[s], st.ymin[s] - tt.ymax[t]));
            const double gz = std::max(0.0, std::max(tt.zmin[t] - st.zmax[s], st.zmin[s] - tt.zmax[t]));
            return gx * gx + gy * gy + gz * gz;
        };
        // Flags of a source range (node s) for a target node t: SEG_COINCIDE
        // if a source may coincide with a target (closed boxes within 1e-150:
        // the singularity skip can only trigger there), SEG_CLOSE if a pair may
        // come closer than sqrt(thresh).
        auto within = [&](int t, int s, double g) {
            return tt.xmin[t] - st.xmax[s] <= g && st.xmin[s] - tt.xmax[t] <= g &&
                   tt.ymin[t] - st.ymax[s] <= g && st.ymin[s] - tt.ymax[t] <= g &&
                   tt.zmin[t] - st.zmax[s] <= g && st.zmin[s] - tt.zmax[t] <= g;
        };
        // one near key: target node t (particles [tbeg, tbeg + tcnt) of the
        // concatenation) <- source clusters clus and source particle nodes segs
        auto near_key = [&](int t, int tbeg, int tcnt, const std::vector<int> &clus,
                            const std::vector<int> &segs, const std::vector<char> *sym = nullptr) {
            const double cx = tt.xmid[t], cy = tt.ymid[t], cz = tt.zmid[t];
            const double rt2 = far2(cx, cy, cz, tt, t);
            double R2 = rt2;
            for (int s : segs) R2 = std::max(R2, far2(cx, cy, cz, st, s));
            bool clu_check = false;
            for (int c : clus) {
                const double f2 = far2(cx, cy, cz, st, c);
                if (rt2 + f2 > 1e7 * gap2(t, c)) { clu_check = true; R2 = std::max(R2, f2); }
            }
            const double thresh = std::max(2e-7 * R2, 1e-300);
            const double gclose = std::max(1e-150, std::sqrt(thresh));
            const long long nt = tcnt;
            add_key(nw, tn + t, tbeg, tcnt);
            for (int c : clus) add_clu(nw, sn + c, nt);
            for (size_t i = 0; i < segs.size(); i++) {
                const int s = segs[i];
                add_seg(nw, so + st.ibeg[s], st.numpar[s], nt,
                        (within(t, s, 1e-150) ? SEG_COINCIDE : 0) | (within(t, s, gclose) ? SEG_CLOSE : 0) |
                            (sym && (*sym)[i] ? SEG_SYM : 0));
            }
            close_key(nw);
            nctr.push_back(cx); nctr.push_back(cy); nctr.push_back(cz);
            nctr.push_back(clu_check ? -thresh : thresh);
        };
        if (p.compute_type == C_PC) {
            // keys: target batches; pc list -> clusters, pp list -> source leaves
            std::vector<std::vector<char>> sym_of(tt.n);
            if (items) {
                std::vector<int> keys(tt.n);
                for (int b = 0; b < tt.n; b++) keys[b] = b;
                build_sym(sy, keys, L.pp, sym_of, sym_ok_);
            }
            for (int b = 0; b < tt.n; b++)
                near_key(b, to + tt.ibeg[b], tt.numpar[b], L.approx[b], L.pp[b], items ? &sym_of[b] : nullptr);
        } else if (p.compute_type == C_CP) {
            // lists keyed by source batch; invert to target-node keys
            std::vector<std::vector<int>> pp_of(tt.n), cp_of(tt.n);
            for (int b = 0; b < st.n; b++) {
                for (int t : L.pp[b]) pp_of[t].push_back(b);
                for (int t : L.approx[b]) cp_of[t].push_back(b);
            }
            const std::vector<int> none;
            std::vector<std::vector<char>> sym_of(tt.n);
            if (items) build_sym(sy, tt.leaves, pp_of, sym_of, sym_ok_);
            for (int leaf : tt.leaves)
                near_key(leaf, to + tt.ibeg[leaf], tt.numpar[leaf], none, pp_of[leaf],
                         items ? &sym_of[leaf] : nullptr);
            for (int t = 0; t < tt.n; t++) {
                if (cp_of[t].empty()) continue;
                add_key(gw, tn + t, 0, P);
                for (int b : cp_of[t]) add_seg(gw, so + st.ibeg[b], st.numpar[b], P);
                close_key(gw);
            }
        } else {
            // CC: a target leaf gets the PP and source-approx (PC) lists of
            // itself and of every ancestor; grids get CC and target-approx (CP)
            std::vector<std::vector<int>> clus_of(tt.n), segs_of(tt.n);
            for (int leaf : tt.leaves) {
                for (int a = leaf; a >= 0; a = tt.parent[a]) {
                    clus_of[leaf].insert(clus_of[leaf].end(), L.cc_pc[a].begin(), L.cc_pc[a].end());
                    segs_of[leaf].insert(segs_of[leaf].end(), L.pp[a].begin(), L.pp[a].end());
                }
            }
            std::vector<std::vector<char>> sym_of(tt.n);
            if (items) build_sym(sy, tt.leaves, segs_of, sym_of, sym_ok_);
            for (int leaf : tt.leaves)
                near_key(leaf, to + tt.ibeg[leaf], tt.numpar[leaf], clus_of[leaf], segs_of[leaf],
                         items ? &sym_of[leaf] : nullptr);
            for (int t = 0; t < tt.n; t++) {
                if (L.approx[t].empty() && L.cc_cp[t].empty()) continue;
                add_key(gw, tn + t, 0, P);
                for (int c : L.approx[t]) add_clu(gw, sn + c, P);
                for (int s : L.cc_cp[t]) add_seg(gw, so + st.ibeg[s], st.numpar[s], P);
                close_key(gw);
            }
        }
    };
    auto csr_start = [](WorkCSR &w) { w.seg_off.assign(1, 0); w.clu_off.assign(1, 0); };
    csr_start(near_);
    csr_start(grid_);
    const int ns = (int)sys_.size();
    if (ns == 1) {
        per_system(sys_[0], near_, grid_, near_ctr_);
    } else {
        // systems in parallel (the symmetric near field needs a single
        // system), then concatenated in system order
        std::vector<WorkCSR> nws(ns), gws(ns);
        std::vector<std::vector<double>> ctrs(ns);
#pragma omp parallel for schedule(dynamic)
        for (int s = 0; s < ns; s++) {
            csr_start(nws[s]);
            csr_start(gws[s]);
            per_system(sys_[s], nws[s], gws[s], ctrs[s]);
        }
        auto append = [](WorkCSR &w, const WorkCSR &f) {
            const int so = (int)w.seg_beg.size(), co = (int)w.clu_id.size();
            w.key.insert(w.key.end(), f.key.begin(), f.key.end());
            w.tbeg.insert(w.tbeg.end(), f.tbeg.begin(), f.tbeg.end());
            w.tcnt.insert(w.tcnt.end(), f.tcnt.begin(), f.tcnt.end());
            for (size_t k = 1; k < f.seg_off.size(); k++) w.seg_off.push_back(so + f.seg_off[k]);
            for (size_t k = 1; k < f.clu_off.size(); k++) w.clu_off.push_back(co + f.clu_off[k]);
            w.seg_beg.insert(w.seg_beg.end(), f.seg_beg.begin(), f.seg_beg.end());
            w.seg_cnt.insert(w.seg_cnt.end(), f.seg_cnt.begin(), f.seg_cnt.end());
            w.seg_chk.insert(w.seg_chk.end(), f.seg_chk.begin(), f.seg_chk.end());
            w.clu_id.insert(w.clu_id.end(), f.clu_id.begin(), f.clu_id.end());
            w.pairs += f.pairs;
        };
        for (int s = 0; s < ns; s++) {
            append(near_, nws[s]);
            append(grid_, gws[s]);
            near_ctr_.insert(near_ctr_.end(), ctrs[s].begin(), ctrs[s].end());
        }
    }

    // heavy grid keys are split by sources; every key writes its own row
    {
        WorkCSR raw = std::move(grid_);
        split_keys(raw, MAX_KEY_SOURCES, P, false, grid_);
    }
    grid_row_.assign(grid_.key.size(), 0);
    grid_rows_len_ = 0;
    for (size_t k = 0; k < grid_.key.size(); k++) {
        grid_row_[k] = grid_rows_len_;
        grid_rows_len_ += grid_.tcnt[k];
    }
    node_rows_off_.assign(ntnodes_ + 1, 0);
    for (int node : grid_.key) node_rows_off_[node + 1]++;
    for (int n = 0; n < ntnodes_; n++) node_rows_off_[n + 1] += node_rows_off_[n];
    make_queue_tiles(near_, tile_targets_, false, P, near_tiles_, near_qlen_);
    make_queue_tiles(grid_, GRID_TILE, true, P, grid_tiles_, grid_qlen_);
    make_pencil_tiles(grid_, P, false, pencil_tiles_);

    // upward pass (source trees, PC and CC): non-leaf nodes per level,
    // deepest first, every system in the same launch
    up_levels_off_.assign(1, 0);
    up_nodes_.clear(); up_child_off_.assign(1, 0); up_child_id_.clear();
    g_sleaves_.clear();
    if (p.compute_type != C_CP) {
        for (int lv = sdepth_ - 1; lv >= 0; --lv) {
            for (const System &sy : sys_) {
                if (lv >= sy.stree.max_depth) continue;
                for (int par : sy.stree.levels[lv]) {
                    up_nodes_.push_back(sy.snode_off + par);
                    for (int k = 0; k < sy.stree.nchild[par]; k++)
                        up_child_id_.push_back(sy.snode_off + sy.stree.child[8 * par + k]);
                    up_child_off_.push_back((int)up_child_id_.size());
                }
            }
            up_levels_off_.push_back((int)up_nodes_.size());
        }
        for (const System &sy : sys_)
            for (int leaf : sy.stree.leaves) g_sleaves_.push_back(sy.snode_off + leaf);
    }
    // downward pass (target trees, CP and CC): children per level, top first
    down_levels_off_.assign(1, 0);
    down_nodes_.clear();
    leaf_tiles_.clear();
    if (p.compute_type != C_PC) {
        for (int lv = 0; lv < tdepth_; ++lv) {
            for (const System &sy : sys_) {
                if (lv >= sy.ttree.max_depth) continue;
                for (int par : sy.ttree.levels[lv])
                    for (int k = 0; k < sy.ttree.nchild[par]; k++)
                        down_nodes_.push_back(sy.tnode_off + sy.ttree.child[8 * par + k]);
            }
            down_levels_off_.push_back((int)down_nodes_.size());
        }
        for (const System &sy : sys_) {
            for (int leaf : sy.ttree.leaves) {
                for (int o = 0; o < sy.ttree.numpar[leaf]; o += DOWN_TILE) {
                    btk::Tile t;
                    t.x = sy.tnode_off + leaf;
                    t.y = sy.toff + sy.ttree.ibeg[leaf] + o;
                    t.z = std::min(DOWN_TILE, sy.ttree.numpar[leaf] - o);
                    t.w = 0;
                    leaf_tiles_.push_back(t);
                }
            }
        }
    }
    // global node ranges and parents
    g_sibeg_.clear(); g_siend_.clear(); g_tparent_.clear();
    for (const System &sy : sys_) {
        for (int i = 0; i < sy.stree.n; i++) {
            g_sibeg_.push_back(sy.soff + sy.stree.ibeg[i]);
            g_siend_.push_back(sy.soff + sy.stree.iend[i]);
        }
        for (int i = 0; i < sy.ttree.n; i++)
            g_tparent_.push_back(sy.ttree.parent[i] < 0 ? -1 : sy.tnode_off + sy.ttree.parent[i]);
    }
    // leaves' modified charges: one block per leaf, or, when the leaves are
    // too few to fill the GPU (fewer than UP_SPLIT_BLOCKS), parts of at least
    // UP_PART_MIN particles (multiples of the 64-particle chunk) whose rows
    // are then added per leaf in part order
    up_part_node_.clear(); up_part_beg_.clear(); up_part_end_.clear(); up_part_off_.clear();
    if (p.compute_type != C_CP && p.approximation == A_LAGRANGE && (int)g_sleaves_.size() < UP_SPLIT_BLOCKS) {
        long long np = 0;
        for (int l : g_sleaves_) np += g_siend_[l] - g_sibeg_[l];
        int part = (int)std::max<long long>(UP_PART_MIN, (np + UP_SPLIT_BLOCKS - 1) / UP_SPLIT_BLOCKS);
        part = (part + 63) / 64 * 64;
        std::vector<int> first(nsnodes_ + 1, 0);
        std::vector<std::vector<std::array<int, 2>>> of(nsnodes_);
        for (int l : g_sleaves_)
            for (int b = g_sibeg_[l]; b < g_siend_[l]; b += part) of[l].push_back({b, std::min(b + part, g_siend_[l])});
        up_part_off_.assign(1, 0);
        for (int n = 0; n < nsnodes_; n++) {
            for (const auto &r : of[n]) {
                up_part_node_.push_back(n);
                up_part_beg_.push_back(r[0]);
                up_part_end_.push_back(r[1]);
            }
            up_part_off_.push_back((int)up_part_node_.size());
        }
    }
}

// Every index a kernel will dereference is checked here, on the host,
// before anything is launched.
std::string Plan::validate() const
{
    const int P = prm_.P;
    auto chk_csr = [&](const WorkCSR &w, const std::vector<btk::Tile> &tiles, bool grid) -> std::string {
        const int nk = (int)w.key.size();
        if ((int)w.seg_off.size() != nk + 1 || (int)w.clu_off.size() != nk + 1) return "csr offsets";
        if (w.seg_off[nk] != (int)w.seg_beg.size() || w.clu_off[nk] != (int)w.clu_id.size()) return "csr ends";
        for (int k = 0; k < nk; k++)
            if (w.seg_off[k] > w.seg_off[k + 1] || w.clu_off[k] > w.clu_off[k + 1]) return "csr order";
        for (size_t i = 0; i < w.seg_beg.size(); i++)
            if (w.seg_beg[i] < 0 || w.seg_cnt[i] < 0 || (long long)w.seg_beg[i] + w.seg_cnt[i] > N_ ||
                (unsigned)w.seg_cnt[i] > btk::SEG_COUNT_MASK)
                return "source range";
        for (int c : w.clu_id)
            if (c < 0 || c >= nsnodes_) return "cluster id";
        for (const btk::Tile &t : tiles) {
            if (t.z == 0) continue;  // padding
            if (t.x < 0 || t.x >= nk || t.z < 0 || t.z > (grid ? GRID_TILE : tile_targets_) || t.y < 0)
                return "tile";
            if (grid) {
                if (t.w < 0 || t.w >= ntnodes_ || t.y + t.z > P) return "grid tile";
            } else if ((long long)t.y + t.z > M_) {
                return "target tile";
            }
        }
        return "";
    };
    auto chk_q = [&](const std::vector<btk::Tile> &tiles, const int *qlen) {
        for (int x = 0; x < NXCD; x++)
            for (int j = 0; j < qlen[x]; j++) {
                const size_t t = (size_t)j * NXCD + x;
                if (t >= tiles.size() || tiles[t].z <= 0) return false;
            }
        return true;
    };
    if (!chk_q(near_tiles_, near_qlen_) || !chk_q(grid_tiles_, grid_qlen_)) return "tile queues";
    std::string e = chk_csr(near_, near_tiles_, false);
    if (!e.empty()) return "near " + e;
    if (near_ctr_.size() != 4 * near_.key.size()) return "near centres";
    if (sym_) {
        if (sym_ctr_.size() != 4 * sym_items_.size() || sym_items_.size() > (size_t)INT32_MAX) return "sym items";
        for (const btk::SymItem &it : sym_items_) {
            if (it.a_cnt == 0 && it.b_cnt == 0 && it.rslot == 0 && it.cslot == 0) continue;  // padding
            if (it.a_beg < 0 || it.a_cnt <= 0 || (long long)it.a_beg + it.a_cnt > M_ || it.b_beg < 0 ||
                it.b_cnt <= 0 || (long long)it.b_beg + it.b_cnt > N_ || (it.cslot >= 0 && (!sym_ok_ || N_ != M_)))
                return "sym item range";
            if (it.rslot < 0 || it.rslot + it.a_cnt > sym_slots_len_ || it.cslot < -1 ||
                (it.cslot >= 0 && it.cslot + it.b_cnt > sym_slots_len_))
                return "sym item slots";
        }
        const int nl = (int)sym_red_off_.size() - 1;
        if (nl < 0 || sym_red_off_[nl] != (int)sym_red_base_.size()) return "sym lists";
        for (int l = 0; l < nl; l++)
            if (sym_red_off_[l] > sym_red_off_[l + 1]) return "sym lists";
        for (const btk::Tile &t : sym_red_tiles_) {
            if (t.x < 0 || t.x >= nl || t.y < 0 || t.z <= 0 || t.z > btk::TPB_SYM_REDUCE ||
                (long long)t.y + t.z > M_)
                return "sym tile";
            for (int k = sym_red_off_[t.x]; k < sym_red_off_[t.x + 1]; k++)
                if (sym_red_base_[k] + t.y < 0 || sym_red_base_[k] + t.y + t.z > sym_slots_len_)
                    return "sym slot base";
        }
    }
    if (!(e = chk_csr(grid_, grid_tiles_, true)).empty()) return "far " + e;
    if (pencil_tiles_.size() != grid_.key.size()) return "pencil tiles";
    for (const btk::Tile &t : pencil_tiles_)
        if (t.x < 0 || t.x >= (int)grid_.key.size() || t.y != 0 || t.z != P || t.w != grid_.key[t.x])
            return "pencil tile";
    if (grid_row_.size() != grid_.key.size() || (int)node_rows_off_.size() != ntnodes_ + 1 ||
        node_rows_off_[ntnodes_] != (int)grid_.key.size())
        return "grid rows";
    for (size_t k = 0; k < grid_.key.size(); k++) {
        const int node = grid_.key[k];
        if (grid_row_[k] < 0 || grid_row_[k] + grid_.tcnt[k] > grid_rows_len_ || grid_.tcnt[k] != P ||
            (int)k < node_rows_off_[node] || (int)k >= node_rows_off_[node + 1])
            return "grid row";
    }
    for (size_t i = 0; i < g_sibeg_.size(); i++)
        if (g_sibeg_[i] < 0 || g_siend_[i] > N_ || g_sibeg_[i] >= g_siend_[i]) return "source node range";
    for (int l : g_sleaves_)
        if (l < 0 || l >= nsnodes_) return "source leaf";
    if (!up_part_node_.empty()) {
        if ((int)up_part_off_.size() != nsnodes_ + 1 || up_part_off_[nsnodes_] != (int)up_part_node_.size())
            return "upward parts";
        for (int n = 0; n < nsnodes_; n++)
            for (int k = up_part_off_[n]; k < up_part_off_[n + 1]; k++)
                if (up_part_node_[k] != n || up_part_beg_[k] < g_sibeg_[n] || up_part_end_[k] > g_siend_[n] ||
                    up_part_beg_[k] >= up_part_end_[k])
                    return "upward part range";
    }
    for (int c : up_child_id_)
        if (c < 0 || c >= nsnodes_) return "upward child";
    for (int n : up_nodes_)
        if (n < 0 || n >= nsnodes_) return "upward parent";
    for (int c : down_nodes_)
        if (c < 0 || c >= ntnodes_ || g_tparent_[c] < 0 || g_tparent_[c] >= ntnodes_) return "downward child";
    for (const btk::Tile &t : leaf_tiles_)
        if (t.x < 0 || t.x >= ntnodes_ || t.y < 0 || (long long)t.y + t.z > M_ || t.z <= 0) return "leaf tile";
    // (a device-built order is a stable partition of 0..M-1: a permutation)
    if (!dev_tree_) {
        if ((int)g_torder_.size() != M_) return "target order size";
        for (int i = 0; i < M_; i++)
            if (g_torder_[i] < 0 || g_torder_[i] >= M_) return "target order";
    }
    return "";
}

std::string Plan::create(const Params &prm, int M, int N, const double *tx, const double *ty,
                         const double *tz, const double *tq, const double *sx, const double *sy,
                         const double *sz, const double *sq, const double *sw, bool host_only,
                         bool device_inputs)
{
    const int toffs[2] = {0, M}, soffs[2] = {0, N};
    PlanInput in;
    // below SMALL_TARGETS targets, tiles of 256 targets (one per thread) give
    // the near launch twice the blocks: about 2 per CU at 128k targets
    in.nsys = 1; in.toffs = toffs; in.soffs = soffs; in.tpt = M < SMALL_TARGETS ? 1 : 2;
    in.tx = tx; in.ty = ty; in.tz = tz; in.tq = tq;
    in.sx = sx; in.sy = sy; in.sz = sz; in.sq = sq; in.sw = sw;
    in.device = device_inputs;
    return create(prm, in, host_only);
}

std::string Plan::create(const Params &prm, const PlanInput &in, bool host_only)
{
    free_all();
    prm_ = prm;
    std::string e = params_check(prm);
    if (!e.empty()) return e;
    if (in.nsys < 1 || !in.toffs || !in.soffs) return "need at least one system";
    if (in.tpt != 1 && in.tpt != 2) return "targets per thread must be 1 or 2";
    tpt_ = in.tpt;
    tile_targets_ = 256 * tpt_;
    persistent_ = in.persistent != 0;
    for (int s = 0; s < in.nsys; s++)
        if (in.toffs[s + 1] - in.toffs[s] <= 0 || in.soffs[s + 1] - in.soffs[s] <= 0)
            return "numTargets and numSources must be positive (every system)";
    if (in.toffs[0] != 0 || in.soffs[0] != 0) return "system offsets must start at 0";
    if (in.device && host_only) return "device-resident particles need a GPU plan";
    M_ = in.toffs[in.nsys];
    N_ = in.soffs[in.nsys];
    if (!in.tx || !in.ty || !in.tz || !in.sx || !in.sy || !in.sz || !in.sq)
        return "One or more required particle arrays is unset";
    if (prm.singularity == S_SUBTRACTION && !in.sw && !host_only)
        return "Singularity subtraction requires weights for the sources";
    for (int s = 0; s < in.nsys; s++) {
        if (in.msl && in.msl[s] < 1) return "leaf sizes must be positive";
        if (in.mtl && in.mtl[s] < 1) return "leaf sizes must be positive";
    }
    if (!host_only) HIPCHK(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));

    tm_ = Timings();
    sys_.assign(in.nsys, System());
    for (int s = 0; s < in.nsys; s++) {
        System &y = sys_[s];
        y.toff = in.toffs[s]; y.soff = in.soffs[s];
        y.M = in.toffs[s + 1] - y.toff; y.N = in.soffs[s + 1] - y.soff;
        y.msl = in.msl ? in.msl[s] : prm.max_source_leaf;
        y.mtl = in.mtl ? in.mtl[s] : prm.max_target_leaf;
    }
    dev_tree_ = !host_only;
    shared_ = false;
    if (dev_tree_) {
        // GPU plans: the particles go to the device as they are and the
        // trees are built there
        if (!(e = device_trees(in)).empty()) return e;
    } else {
        if (!(e = host_trees(in)).empty()) return e;
    }
    double t2 = now();
    return finish_create(in, host_only, t2);
}

// Host-only plans (BaryTree_Analyze): private tree-ordered copies of the
// particles, trees built on the host.
std::string Plan::host_trees(const PlanInput &in)
{
    const Params &prm = prm_;
    double t0 = now();
    for (System &y : sys_) {
        y.S.n = y.N; y.T.n = y.M;
        for (auto *v : {&y.S.x, &y.S.y, &y.S.z, &y.S.q, &y.S.w}) v->resize(y.N);
        for (auto *v : {&y.T.x, &y.T.y, &y.T.z, &y.T.q}) v->resize(y.M);
        y.S.order.resize(y.N);
        y.T.order.resize(y.M);
    }
    // private copies (the caller's arrays are never permuted), in parallel:
    // over systems when there are several, else over particles
    const int ns = (int)sys_.size();
#pragma omp parallel for schedule(dynamic) if (ns > 1)
    for (int s = 0; s < ns; s++) {
        System &y = sys_[s];
        const double *sx = in.sx + y.soff, *sy = in.sy + y.soff, *sz = in.sz + y.soff;
        const double *sq = in.sq + y.soff, *sw = in.sw ? in.sw + y.soff : nullptr;
        const double *tx = in.tx + y.toff, *ty = in.ty + y.toff, *tz = in.tz + y.toff;
        const double *tq = in.tq ? in.tq + y.toff : nullptr;
        const int n = y.N, m = y.M;
#pragma omp parallel for schedule(static) if (n > 65536)
        for (int i = 0; i < n; i++) {
            y.S.x[i] = sx[i]; y.S.y[i] = sy[i]; y.S.z[i] = sz[i]; y.S.q[i] = sq[i];
            y.S.w[i] = sw ? sw[i] : 1.0;
            y.S.order[i] = i;
        }
#pragma omp parallel for schedule(static) if (m > 65536)
        for (int i = 0; i < m; i++) {
            y.T.x[i] = tx[i]; y.T.y[i] = ty[i]; y.T.z[i] = tz[i]; y.T.q[i] = tq ? tq[i] : 0.0;
            y.T.order[i] = i;
        }
    }
    double t1 = now();
    tm_.host_copy = t1 - t0;

    // trees (treedriver.c:128/146, 364/375, 654/665), all systems at once
    std::vector<TreeJob> jobs;
    for (System &y : sys_) {
        if (prm.compute_type == C_PC) {
            jobs.push_back({&y.stree, &y.S, y.msl, false});
            jobs.push_back({&y.ttree, &y.T, y.mtl, true});
        } else if (prm.compute_type == C_CP) {
            jobs.push_back({&y.ttree, &y.T, y.mtl, false});
            jobs.push_back({&y.stree, &y.S, y.msl, true});
        } else {
            jobs.push_back({&y.stree, &y.S, y.msl, false});
            jobs.push_back({&y.ttree, &y.T, y.mtl, false});
        }
    }
    std::string e = build_trees(jobs);
    tm_.host_tree = now() - t1;
    return e;
}

// GPU plans: the caller's arrays go to the device unchanged, the trees are
// built there (device_build_trees) and the charges, weights and the target
// order are gathered into tree order.  When every system's targets are its
// sources, bit for bit, and its leaf sizes are equal, one tree and one set
// of arrays serve both roles (the target batches of PC and the source
// batches of CP are the leaves of that tree).
std::string Plan::device_trees(const PlanInput &in)
{
    const Params &prm = prm_;
    const bool ss = prm.singularity == S_SUBTRACTION;
    double t0 = now();
    bool same = M_ == N_ && in.tq != nullptr;
    for (int s = 0; same && s < (int)sys_.size(); s++)
        same = in.toffs[s] == in.soffs[s] && sys_[s].msl == sys_[s].mtl;
    const bool alias = in.tx == in.sx && in.ty == in.sy && in.tz == in.sz && in.tq == in.sq;
    if (same && !alias && in.device) {
        // bitwise equality on the device
        int *flag = dalloc<int>(1);
        if (!flag) return "hipMalloc failed (comparison)";
        HIPCHK(hipMemsetAsync(flag, 0, sizeof(int), st_));
        HIPCHK(btk::launch_differ(M_, in.tx, in.sx, flag, st_));
        HIPCHK(btk::launch_differ(M_, in.ty, in.sy, flag, st_));
        HIPCHK(btk::launch_differ(M_, in.tz, in.sz, flag, st_));
        HIPCHK(btk::launch_differ(M_, in.tq, in.sq, flag, st_));
        int differ = 0;
        HIPCHK(hipMemcpyAsync(&differ, flag, sizeof(int), hipMemcpyDeviceToHost, st_));
        HIPCHK(hipStreamSynchronize(st_));
        dfree(flag);
        same = differ == 0;
    } else if (same && !alias) {
        // bitwise equality, in parallel blocks
        const long long n = M_;
        const long long blk = 1 << 18, nb = (n + blk - 1) / blk;
        int diff = 0;
#pragma omp parallel for reduction(| : diff) schedule(static) if (nb > 1)
        for (long long b = 0; b < nb; b++) {
            const size_t o = (size_t)(b * blk), len = (size_t)std::min(blk, n - b * blk) * sizeof(double);
            diff |= std::memcmp(in.tx + o, in.sx + o, len) | std::memcmp(in.ty + o, in.sy + o, len) |
                    std::memcmp(in.tz + o, in.sz + o, len) | std::memcmp(in.tq + o, in.sq + o, len);
        }
        same = diff == 0;
    }
    shared_ = same;
    const double t1 = now();
    tm_.host_copy = t1 - t0;
    const hipMemcpyKind kind = in.device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;

    // one particle set: coordinates uploaded, caller indices 0..n-1
    auto upload_set = [&](DevParticles &p, int n, const double *x, const double *y, const double *z) -> std::string {
        p.n = n;
        p.x = dalloc<double>(n); p.y = dalloc<double>(n); p.z = dalloc<double>(n); p.idx = dalloc<int>(n);
        p.x2 = dalloc<double>(n); p.y2 = dalloc<double>(n); p.z2 = dalloc<double>(n); p.idx2 = dalloc<int>(n);
        if (!p.x || !p.y || !p.z || !p.idx || !p.x2 || !p.y2 || !p.z2 || !p.idx2) return "hipMalloc failed (particles)";
        const size_t b = (size_t)n * sizeof(double);
        HIPCHK(hipMemcpyAsync(p.x, x, b, kind, st_));
        HIPCHK(hipMemcpyAsync(p.y, y, b, kind, st_));
        HIPCHK(hipMemcpyAsync(p.z, z, b, kind, st_));
        HIPCHK(btk::launch_iota(n, p.idx, st_));
        return "";
    };
    // the trees of one set: per system its range and leaf size
    auto build = [&](DevParticles &p, bool src, bool leaves_only, std::vector<std::vector<NodeRec>> &recs,
                     std::vector<int> &depth) -> std::string {
        std::vector<DevRoot> roots;
        for (const System &y : sys_) {
            DevRoot r;
            r.b = src ? y.soff : y.toff;
            r.e = r.b + (src ? y.N : y.M);
            r.maxpar = src ? y.msl : y.mtl;
            roots.push_back(r);
        }
        return device_build_trees(p, roots, leaves_only, st_, recs, depth);
    };
    // after the build: the coordinates stay, the second buffers are scratch
    auto release_scratch = [&](DevParticles &p) {
        dfree(p.x2); dfree(p.y2); dfree(p.z2); dfree(p.idx2);
        p.x2 = p.y2 = p.z2 = nullptr;
        p.idx2 = nullptr;
    };
    std::string e;
    const bool src_batches = prm.compute_type == C_CP, tgt_batches = prm.compute_type == C_PC;
    DevParticles S, T;
    if (!(e = upload_set(S, N_, in.sx, in.sy, in.sz)).empty()) return e;
    std::vector<std::vector<NodeRec>> srec, trec;
    std::vector<int> sdep, tdep;
    double tb = now();
    if (!(e = build(S, true, shared_ ? false : src_batches, srec, sdep)).empty()) return e;
    double tree_s = now() - tb;
    // charges and weights in tree order: the raw arrays go to the scratch
    // buffers, then a gather
    d_sx_ = S.x; d_sy_ = S.y; d_sz_ = S.z;
    d_sq_ = dalloc<double>(N_);
    if (ss) d_sw_ = dalloc<double>(N_);
    if (!d_sq_ || (ss && !d_sw_)) return "hipMalloc failed (charges)";
    const size_t nb = (size_t)N_ * sizeof(double);
    HIPCHK(hipMemcpyAsync(S.x2, in.sq, nb, kind, st_));
    if (ss) HIPCHK(hipMemcpyAsync(S.y2, in.sw, nb, kind, st_));
    HIPCHK(btk::launch_tree_gather(N_, S.idx, S.x2, d_sq_, ss ? S.y2 : nullptr, 1.0, ss ? d_sw_ : nullptr, st_));
    if (shared_) {
        d_tx_ = d_sx_; d_ty_ = d_sy_; d_tz_ = d_sz_; d_tq_ = d_sq_;
        d_torder_ = S.idx;
        S.idx = nullptr;
        HIPCHK(hipStreamSynchronize(st_));
        release_scratch(S);
    } else {
        HIPCHK(hipStreamSynchronize(st_));
        release_scratch(S);
        dfree(S.idx);
        if (!(e = upload_set(T, M_, in.tx, in.ty, in.tz)).empty()) return e;
        tb = now();
        if (!(e = build(T, false, tgt_batches, trec, tdep)).empty()) return e;
        tree_s += now() - tb;
        d_tx_ = T.x; d_ty_ = T.y; d_tz_ = T.z;
        d_tq_ = dalloc<double>(M_);
        if (!d_tq_) return "hipMalloc failed (target values)";
        if (in.tq) {
            HIPCHK(hipMemcpyAsync(T.x2, in.tq, (size_t)M_ * sizeof(double), kind, st_));
            HIPCHK(btk::launch_tree_gather(M_, T.idx, T.x2, d_tq_, nullptr, 0.0, nullptr, st_));
        } else {
            HIPCHK(hipMemsetAsync(d_tq_, 0, (size_t)M_ * sizeof(double), st_));
        }
        d_torder_ = T.idx;
        T.idx = nullptr;
        HIPCHK(hipStreamSynchronize(st_));
        release_scratch(T);
    }
    // the host trees from the device records
    double tf = now();
    const int nsys = (int)sys_.size();
#pragma omp parallel for schedule(dynamic) if (nsys > 1)
    for (int k = 0; k < nsys; k++) {
        System &y = sys_[k];
        if (shared_) {
            const std::vector<NodeRec> leaves = leaf_records(srec[k]);
            // PC: source tree + target batches; CP: target tree + source
            // batches; CC: both trees
            finish_tree(y.stree, src_batches ? leaves : srec[k], sdep[k], src_batches);
            finish_tree(y.ttree, tgt_batches ? leaves : srec[k], sdep[k], tgt_batches);
        } else {
            finish_tree(y.stree, srec[k], sdep[k], src_batches);
            finish_tree(y.ttree, trec[k], tdep[k], tgt_batches);
        }
    }
    tree_s += now() - tf;
    tm_.host_tree = tree_s;
    tm_.h2d = now() - t1 - tree_s;
    return "";
}

std::string Plan::finish_create(const PlanInput &in, bool host_only, double t2)
{
    const Params &prm = prm_;
    const int ns = (int)sys_.size();
    std::string e;

    // interaction lists (treedriver.c:181, 462, 753)
#pragma omp parallel for schedule(dynamic) if (ns > 1)
    for (int s = 0; s < ns; s++) {
        System &y = sys_[s];
        if (prm.compute_type == C_PC) make_pc_lists(y.L, y.stree, y.ttree, prm);
        else if (prm.compute_type == C_CP) make_pc_lists(y.L, y.ttree, y.stree, prm);
        else make_cc_lists(y.L, y.stree, y.ttree, prm);
    }
    double t3 = now();
    tm_.host_lists = t3 - t2;

    // global node numbering
    nsnodes_ = ntnodes_ = sdepth_ = tdepth_ = 0;
    for (System &y : sys_) {
        y.snode_off = nsnodes_; y.tnode_off = ntnodes_;
        nsnodes_ += y.stree.n; ntnodes_ += y.ttree.n;
        sdepth_ = std::max(sdepth_, y.stree.max_depth);
        tdepth_ = std::max(tdepth_, y.ttree.max_depth);
    }

    // reference diagnostics (treedriver.c:191-200, 472-483, 763-792), summed
    counts_ = Counts();
    const long long P = prm.P;
    for (const System &y : sys_) {
        const Tree &key = prm.compute_type == C_CP ? y.stree : y.ttree;
        const Tree &ent = prm.compute_type == C_CP ? y.ttree : y.stree;
        const Lists &L = y.L;
        for (int i = 0; i < key.n; i++) {
            const long long nt = key.numpar[i];
            counts_.num_pp += (long long)L.pp[i].size();
            for (int s : L.pp[i]) counts_.pp_ptwise += nt * ent.numpar[s];
            if (prm.compute_type == C_PC) {
                counts_.num_pc += (long long)L.approx[i].size();
                for (int s : L.approx[i]) { counts_.pc_ptwise += nt * P; counts_.pc_replaced += nt * ent.numpar[s]; }
            } else if (prm.compute_type == C_CP) {
                counts_.num_cp += (long long)L.approx[i].size();
                for (int s : L.approx[i]) { counts_.cp_ptwise += nt * P; counts_.cp_replaced += nt * ent.numpar[s]; }
            } else {
                counts_.num_cc += (long long)L.approx[i].size();
                counts_.num_pc += (long long)L.cc_pc[i].size();
                counts_.num_cp += (long long)L.cc_cp[i].size();
                for (int s : L.approx[i]) { counts_.cc_ptwise += P * P; counts_.cc_replaced += nt * ent.numpar[s]; }
                for (int s : L.cc_pc[i]) { counts_.pc_ptwise += nt * P; counts_.pc_replaced += nt * ent.numpar[s]; }
                for (int s : L.cc_cp[i]) { counts_.cp_ptwise += P * ent.numpar[s]; counts_.cp_replaced += nt * ent.numpar[s]; }
            }
        }
    }

    herm_ = prm.approximation == A_HERMITE;
    // symmetric near field: one system (Coulomb or Yukawa, or regularized
    // Coulomb with skipping; either approximation, which only concerns the
    // far field) whose targets are its sources, particle for particle in
    // tree order (the same order after both tree builds)
    sym_ok_ = false;
    if (sys_.size() == 1 &&
        (prm.kernel == K_COULOMB || prm.kernel == K_YUKAWA ||
         (prm.kernel == K_REG_COULOMB && prm.singularity == S_SKIPPING))) {
        const System &y = sys_[0];
        if (dev_tree_) {
            // one tree and one set of arrays serve as sources and targets
            sym_ok_ = shared_;
        } else {
            auto same = [&](const HostVec<double> &u, const HostVec<double> &v) {
                return std::memcmp(u.data(), v.data(), (size_t)y.M * sizeof(double)) == 0;
            };
            sym_ok_ = y.M == y.N && same(y.S.x, y.T.x) && same(y.S.y, y.T.y) && same(y.S.z, y.T.z) &&
                      same(y.S.q, y.T.q);
        }
    }
    // one system with other targets, skipping kernels: one-sided items
    // (run in mixed precision, in place of the one-directional particle
    // pairs of the near launch)
    items_ok_ = !sym_ok_ && sys_.size() == 1 && prm.singularity == S_SKIPPING &&
                (prm.kernel == K_COULOMB || prm.kernel == K_YUKAWA || prm.kernel == K_REG_COULOMB);
    build_work();
    enumerate_pairs();
    if (herm_) build_hermite_work();
    const bool need_src_clusters = prm.compute_type != C_CP;
    const bool need_tgt_clusters = prm.compute_type != C_PC;
    std::vector<double> stab, ttab, tab;
    for (const System &y : sys_) {
        if (need_src_clusters) {
            node_table(y.stree, prm.degree, tab);
            stab.insert(stab.end(), tab.begin(), tab.end());
        }
        if (need_tgt_clusters) {
            node_table(y.ttree, prm.degree, tab);
            ttab.insert(ttab.end(), tab.begin(), tab.end());
        }
    }
    if (!dev_tree_) {
        g_torder_.resize(M_);
#pragma omp parallel for schedule(dynamic) if (ns > 1)
        for (int s = 0; s < ns; s++) {
            const System &y = sys_[s];
            int *o = g_torder_.data() + y.toff;
            const int *lo = y.T.order.data();
            const int off = y.toff, m = y.M;
#pragma omp parallel for schedule(static) if (m > 65536)
            for (int i = 0; i < m; i++) o[i] = off + lo[i];
        }
    }
    double t4 = now();
    tm_.host_work = t4 - t3;
    if (!(e = validate()).empty()) return "internal work-list check failed: " + e;
    if (herm_ && !(e = validate_hermite()).empty()) return "internal work-list check failed: " + e;
    if (host_only) return "";

    // kernel parameters (src/kernels/*/ *_ss_correction.c and pair kernels)
    kp_ = btk::KParams{};
    kp_.p0 = prm.kp[0];
    kp_.p1 = prm.kp[1];
    if (prm.kernel == K_COULOMB) {
        kp_.inv_a2 = prm.singularity == S_SUBTRACTION ? 1.0 / (prm.kp[0] * prm.kp[0]) : 0.0;
        kp_.ss_param = 2.0 * M_PI * prm.kp[0] * prm.kp[0];
    } else if (prm.kernel == K_YUKAWA) {
        kp_.ss_param = 4.0 * M_PI / prm.kp[0] / prm.kp[0];
    } else if (prm.kernel == K_REG_YUKAWA) {
        // (kappa, epsilon); regularized-yukawa_ss_correction.c:12-14
        kp_.eps2 = prm.kp[1] * prm.kp[1];
        kp_.ss_param = 4.0 * M_PI / prm.kp[0] / prm.kp[0];
    } else if (prm.kernel == K_ATAN) {
        // (domain length, delta): atan_pp.c:24-28
        const double delta = prm.kp[1];
        kp_.c1 = std::sqrt(1. + 1.0 / (delta * delta));
        kp_.c2 = 1. / (1. - delta / std::sqrt(1. + delta * delta));
    } else if (prm.kernel == K_MQ) {
        // (domain length, delta): mq_pp.c:24-27
        kp_.c1 = prm.kp[1] * prm.kp[1] / prm.kp[0] / prm.kp[0];
        kp_.c2 = std::sqrt(1 + 4 * kp_.c1);
    } else if (prm.kernel == K_RBS_U || prm.kernel == K_RBS_V) {
        kp_.c1 = prm.kp[0] * prm.kp[0];  // delta^2
    } else if (prm.kernel == K_SIN_OVER_R || prm.kernel == K_USER) {
        // k = kernel_params[0] (p0)
    } else {
        if (prm.singularity == S_SUBTRACTION) {
            kp_.inv_a2 = 1.0 / (prm.kp[0] * prm.kp[0]);
            kp_.eps2 = prm.kp[1] * prm.kp[1];
        } else {
            kp_.eps2 = prm.kp[0] * prm.kp[0];
        }
        kp_.ss_param = 2.0 * M_PI * prm.kp[0] * prm.kp[0];
    }
    if (prm.singularity != S_SUBTRACTION) kp_.ss_param = 0.0;

    // the particles and the target order are on the device already
    // (device_trees); the work lists and tables follow
    const bool ss = prm.singularity == S_SUBTRACTION;
    std::string r;
    const size_t Pz = prm.P;
    const size_t planes = herm_ ? 8 : 1;  // Hermite: 8 planes per grid point
    if (need_src_clusters) {
        if (!(r = up(&d_stab_, stab)).empty()) return r;
        const size_t np = (size_t)nsnodes_ * Pz;
        d_cx_ = dalloc<double>(np); d_cy_ = dalloc<double>(np); d_cz_ = dalloc<double>(np);
        d_cq_ = dalloc<double>(np * planes);
        if (ss) d_cw_ = dalloc<double>(np * planes);
        if (!d_cx_ || !d_cy_ || !d_cz_ || !d_cq_ || (ss && !d_cw_)) return "hipMalloc failed (source clusters)";
        if (!(r = up(&d_sibeg_, g_sibeg_)).empty()) return r;
        if (!(r = up(&d_siend_, g_siend_)).empty()) return r;
        if (!(r = up(&d_sleaves_, g_sleaves_)).empty()) return r;
        if (!up_part_node_.empty()) {
            if (!(r = up(&d_up_part_node_, up_part_node_)).empty()) return r;
            if (!(r = up(&d_up_part_beg_, up_part_beg_)).empty()) return r;
            if (!(r = up(&d_up_part_end_, up_part_end_)).empty()) return r;
            if (!(r = up(&d_up_part_off_, up_part_off_)).empty()) return r;
            const size_t npp = up_part_node_.size() * Pz;
            d_up_Qp_ = dalloc<double>(npp);
            if (ss) d_up_Wp_ = dalloc<double>(npp);
            if (!d_up_Qp_ || (ss && !d_up_Wp_)) return "hipMalloc failed (upward parts)";
        }
        if (!(r = up(&d_up_nodes_, up_nodes_)).empty()) return r;
        if (!(r = up(&d_up_child_off_, up_child_off_)).empty()) return r;
        if (!(r = up(&d_up_child_id_, up_child_id_)).empty()) return r;
        if (!up_child_id_.empty()) {
            // per parent-child entry: its parent, and the identity offsets
            std::vector<int> ent_parent(up_child_id_.size()), ent_off(up_child_id_.size() + 1);
            for (size_t k = 0; k + 1 < up_child_off_.size(); k++)
                for (int m = up_child_off_[k]; m < up_child_off_[k + 1]; m++) ent_parent[m] = up_nodes_[k];
            for (size_t m = 0; m < ent_off.size(); m++) ent_off[m] = (int)m;
            if (!(r = up(&d_up_entry_parent_, ent_parent)).empty()) return r;
            if (!(r = up(&d_up_entry_off_, ent_off)).empty()) return r;
            const size_t ne = up_child_id_.size() * Pz;
            d_up_Qe_ = dalloc<double>(ne);
            if (ss) d_up_We_ = dalloc<double>(ne);
            if (!d_up_Qe_ || (ss && !d_up_We_)) return "hipMalloc failed (upward entries)";
        }
    }
    if (need_tgt_clusters) {
        if (!(r = up(&d_ttab_, ttab)).empty()) return r;
        const size_t np = (size_t)ntnodes_ * Pz;
        d_Qt_ = dalloc<double>(np * planes);
        if (ss) d_Wt_ = dalloc<double>(np * planes);
        if (!d_Qt_ || (ss && !d_Wt_)) return "hipMalloc failed (target clusters)";
        if (!(r = up(&d_tparent_, g_tparent_)).empty()) return r;
        if (!(r = up(&d_down_nodes_, down_nodes_)).empty()) return r;
        if (!(r = up(&d_leaf_tiles_, leaf_tiles_)).empty()) return r;
    }
    auto up_csr = [&](const WorkCSR &w, int **so, btk::Seg **sg, int **co, int **ci) -> std::string {
        std::vector<btk::Seg> segs(w.seg_beg.size());
        for (size_t i = 0; i < segs.size(); i++) {
            segs[i].x = w.seg_beg[i];
            segs[i].y = (int)((unsigned)w.seg_cnt[i] | ((w.seg_chk[i] & SEG_COINCIDE) ? btk::SEG_COINCIDE_BIT : 0u) |
                              ((w.seg_chk[i] & SEG_CLOSE) ? btk::SEG_CLOSE_BIT : 0u) |
                              ((w.seg_chk[i] & SEG_SYM) ? btk::SEG_SYM_BIT : 0u));
        }
        std::string rr;
        if (!(rr = up(so, w.seg_off)).empty()) return rr;
        if (!(rr = up(sg, segs)).empty()) return rr;
        if (!(rr = up(co, w.clu_off)).empty()) return rr;
        return up(ci, w.clu_id);
    };
    if (!(r = up_csr(near_, &d_nseg_off_, &d_nsegs_, &d_nclu_off_, &d_nclu_id_)).empty()) return r;
    if (!(r = up(&d_near_tiles_, near_tiles_)).empty()) return r;
    if (!(r = up(&d_near_ctr_, near_ctr_)).empty()) return r;
    if (sym_) {
        if (!(r = up(&d_sym_items_, sym_items_)).empty()) return r;
        if (!(r = up(&d_sym_ctr_, sym_ctr_)).empty()) return r;
        if (!(r = up(&d_sym_red_tiles_, sym_red_tiles_)).empty()) return r;
        if (!(r = up(&d_sym_red_off_, sym_red_off_)).empty()) return r;
        if (!(r = up(&d_sym_red_base_, sym_red_base_)).empty()) return r;
    }
    if (!grid_tiles_.empty()) {
        if (!(r = up_csr(grid_, &d_gseg_off_, &d_gsegs_, &d_gclu_off_, &d_gclu_id_)).empty()) return r;
        if (!(r = up(&d_grid_tiles_, grid_tiles_)).empty()) return r;
        if (!(r = up(&d_pencil_tiles_, pencil_tiles_)).empty()) return r;
    }
    if (need_tgt_clusters) {
        if (!(r = up(&d_grid_row_, grid_row_)).empty()) return r;
        if (!(r = up(&d_node_rows_off_, node_rows_off_)).empty()) return r;
        d_Qrows_ = dalloc<double>(grid_rows_len_);
        if (ss) d_Wrows_ = dalloc<double>(grid_rows_len_);
        if (!d_Qrows_ || (ss && !d_Wrows_)) return "hipMalloc failed (grid rows)";
    }
    d_pot_ = dalloc<double>(M_);
    d_out_ = dalloc<double>(M_);
    d_cursor_ = dalloc<int>(2 * NXCD);
    if (!d_pot_ || !d_out_ || !d_cursor_) return "hipMalloc failed (potential)";
    if (herm_ && !(r = upload_hermite()).empty()) return r;
    HIPCHK(hipStreamSynchronize(st_));
    tm_.h2d += now() - t4;
    return "";
}

// Upward pass at the plan's degree over the source trees
// (clusters.c:50-200): grid points, leaves from their particles, then every
// level's parents from their children's grids, deepest first.
std::string Plan::upward_fixed()
{
    const Params &p = prm_;
    const int nl = p.degree + 1;
    HIPCHK(btk::launch_cluster_points(nsnodes_, nl, d_stab_, d_cx_, d_cy_, d_cz_, st_));
    btk::AnterpArgs a{};
    a.degree = p.degree; a.subtraction = p.singularity == S_SUBTRACTION;
    a.ibeg = d_sibeg_; a.iend = d_siend_; a.tab = d_stab_;
    a.px = d_sx_; a.py = d_sy_; a.pz = d_sz_; a.pq = d_sq_; a.pw = d_sw_;
    a.Q = d_cq_; a.W = d_cw_;
    if (up_part_node_.empty()) {
        a.mode = 0; a.nodes = d_sleaves_; a.nblocks = (int)g_sleaves_.size();
        HIPCHK(btk::launch_anterp(a, st_));
    } else {
        // the leaves in parts, then each leaf's rows in part order (every
        // other node's row is cleared here and written by its level below)
        btk::AnterpArgs pa = a;
        pa.mode = 2; pa.nodes = d_up_part_node_; pa.ibeg = d_up_part_beg_; pa.iend = d_up_part_end_;
        pa.Q = d_up_Qp_; pa.W = d_up_Wp_; pa.nblocks = (int)up_part_node_.size();
        HIPCHK(btk::launch_anterp(pa, st_));
        HIPCHK(btk::launch_sum_parts(nsnodes_, p.P, d_up_part_off_, d_up_Qp_, d_up_Wp_, d_cq_,
                                     a.subtraction ? d_cw_ : nullptr, st_));
    }
    // parents level by level: one block per (parent, child) entry into the
    // entry's row, then each parent's rows in child order
    for (size_t lv = 0; lv + 1 < up_levels_off_.size(); lv++) {
        const int b = up_levels_off_[lv], e = up_levels_off_[lv + 1];
        if (e == b) continue;
        const int e0 = up_child_off_[b], e1 = up_child_off_[e];
        btk::AnterpArgs ca = a;
        ca.mode = 3; ca.nodes = d_up_entry_parent_ + e0; ca.nblocks = e1 - e0;
        ca.child_off = d_up_entry_off_ + e0; ca.child_id = d_up_child_id_;
        ca.Qo = d_up_Qe_; ca.Wo = d_up_We_;
        HIPCHK(btk::launch_anterp(ca, st_));
        HIPCHK(btk::launch_sum_parts(e - b, p.P, d_up_child_off_ + b, d_up_Qe_, d_up_We_, d_cq_,
                                     a.subtraction ? d_cw_ : nullptr, st_, d_up_nodes_ + b));
    }
    return "";
}

// Downpass at the plan's degree over the target trees
// (interaction_compute_downpass.c:37-142): parents' interpolants onto the
// children's grids level by level, then the leaves' onto their targets.
std::string Plan::down_fixed()
{
    const Params &p = prm_;
    btk::DownArgs d{};
    d.degree = p.degree; d.subtraction = p.singularity == S_SUBTRACTION; d.tab = d_ttab_; d.parent = d_tparent_;
    d.tx = d_tx_; d.ty = d_ty_; d.tz = d_tz_; d.tq = d_tq_; d.Q = d_Qt_; d.W = d_Wt_; d.pot = d_pot_;
    for (size_t lv = 0; lv + 1 < down_levels_off_.size(); lv++) {
        const int b = down_levels_off_[lv], e = down_levels_off_[lv + 1];
        if (e == b) continue;
        d.mode = 0; d.nodes = d_down_nodes_ + b; d.nblocks = e - b;
        HIPCHK(btk::launch_down(d, st_));
    }
    d.mode = 1; d.tiles = d_leaf_tiles_; d.nblocks = (int)leaf_tiles_.size();
    HIPCHK(btk::launch_down(d, st_));
    return "";
}

// The symmetric near field's items into their slots, then the slots into
// the potential (after the near launch has stored every target's sum).
bool Plan::sym_ready()
{
    if (!items_run()) return false;
    if (d_sym_slots_) return true;
    // the slots at the first execution that uses them; without the memory
    // for them the plan keeps the one-directional near field
    if (hipEventCreate(&sym_ev_[0]) != hipSuccess || hipEventCreate(&sym_ev_[1]) != hipSuccess) {
        sym_ = false;
        return false;
    }
    d_sym_slots_ = dalloc<double>((size_t)sym_slots_len_);
    if (!d_sym_slots_) {
        (void)hipGetLastError();  // clear the allocation error
        sym_ = false;
        return false;
    }
    return true;
}

std::string Plan::launch_items()
{
    btk::SymArgs a{};
    a.kernel = prm_.kernel;
    a.singularity = prm_.singularity;
    a.fp64 = prm_.precision == PREC_MIXED ? 0 : 1;
    a.nitems = (int)sym_items_.size();
    a.items = d_sym_items_;
    a.ctr = reinterpret_cast<const double4 *>(d_sym_ctr_);
    a.x = d_sx_; a.y = d_sy_; a.z = d_sz_; a.q = d_sq_; a.w = d_sw_;
    a.tx = d_tx_; a.ty = d_ty_; a.tz = d_tz_; a.tq = d_tq_;
    a.kp = kp_;
    a.slots = d_sym_slots_;
    forked_ = false;
    if (fork_ev_) {
        HIPCHK(hipEventRecord(fork_ev_, st_));
        forked_ = true;
    }
    HIPCHK(hipEventRecord(sym_ev_[0], st_));
    HIPCHK(btk::launch_near_sym(a, st_));
    HIPCHK(hipEventRecord(sym_ev_[1], st_));
    items_started_ = true;
    return "";
}

std::string Plan::sym_reduce()
{
    btk::SymReduceArgs r{};
    r.ntiles = (int)sym_red_tiles_.size();
    r.tiles = d_sym_red_tiles_;
    r.off = d_sym_red_off_;
    r.base = d_sym_red_base_;
    r.slots = d_sym_slots_;
    r.pot = d_pot_;
    HIPCHK(btk::launch_sym_reduce(r, st_));
    return "";
}

std::string Plan::near_sym(bool reduce)
{
    if (!items_started_) {
        std::string r = launch_items();
        if (!r.empty()) return r;
    }
    items_started_ = false;
    sym_ran_ = true;
    return reduce ? sym_reduce() : "";
}

std::string Plan::start_items()
{
    if (items_started_ || !sym_ready()) return "";
    (void)overlap_ok();  // the fork event, if the second stream is available
    return launch_items();
}

void Plan::side_begin()
{
    if (items_started_ && forked_ && !swapped_ && aux_) {
        std::swap(st_, aux_);
        swapped_ = true;
    }
}

void Plan::side_end()
{
    if (swapped_) {
        std::swap(st_, aux_);
        swapped_ = false;
    }
}

bool Plan::overlap_ok()
{
    if (herm_ || !sym_ready()) return false;
    if (items_started_ && !forked_) return false;  // started before the second stream existed
    if (!aux_) {
        if (hipStreamCreateWithFlags(&aux_, hipStreamNonBlocking) != hipSuccess) {
            (void)hipGetLastError();
            aux_ = nullptr;
            return false;
        }
        for (hipEvent_t *e : {&fork_ev_, &join_ev_, &red_ev_})
            if (hipEventCreate(e) != hipSuccess) {
                (void)hipGetLastError();
                *e = nullptr;
            }
    }
    return fork_ev_ && join_ev_ && red_ev_;
}

std::string Plan::fork_phases(bool ov)
{
    if (!ov) return "";
    if (!items_started_) {
        std::string r = launch_items();
        if (!r.empty()) return r;
    }
    HIPCHK(hipStreamWaitEvent(aux_, fork_ev_, 0));
    std::swap(st_, aux_);
    swapped_ = true;
    return "";
}

std::string Plan::join_phases(bool ov)
{
    if (!ov) return "";
    HIPCHK(hipEventRecord(join_ev_, st_));
    std::swap(st_, aux_);
    swapped_ = false;
    HIPCHK(hipStreamWaitEvent(st_, join_ev_, 0));
    HIPCHK(hipEventRecord(red_ev_, st_));
    return sym_reduce();
}

std::string Plan::execute(bool timed)
{
    if (var_) return execute_var(timed);
    const Params &p = prm_;
    const bool ss = p.singularity == S_SUBTRACTION;
    const int nl = p.degree + 1;
    std::string er = phase_events();
    if (!er.empty()) return er;
    hipEvent_t *ev = ev_;
    sym_ran_ = false;
    const bool ahead = items_started_;
    // with the items running on st_, the phases below run on the second stream
    const bool ov = overlap_ok();
    AuxGuard guard{*this};
    if (!(er = fork_phases(ov)).empty()) return er;
    HIPCHK(hipEventRecord(ev[0], st_));

    // 1. upward pass over the source trees (clusters.c:50-200)
    if (herm_ && p.compute_type == C_PC) {
        std::string r = hermite_upward();
        if (!r.empty()) return r;
    } else if (p.compute_type != C_CP) {
        std::string r = upward_fixed();
        if (!r.empty()) return r;
    }
    HIPCHK(hipEventRecord(ev[1], st_));

    // 2. near field and particle-cluster far field into the potential.  The
    // potential starts as NaN (all bytes 0xFF): the near-field tiles write
    // every target once, so a target that no tile covered would show NaN,
    // never a value left from an earlier execution.
    HIPCHK(hipMemsetAsync(d_pot_, 0xFF, (size_t)M_ * sizeof(double), st_));
    btk::InteractArgs ia{};
    ia.kernel = p.kernel; ia.singularity = p.singularity; ia.precision = p.precision;
    ia.tpt = tpt_;
    ia.persistent = persistent_ ? 1 : 0;
    ia.P = p.P; ia.nl = nl; ia.kp = kp_;
    ia.tx = d_tx_; ia.ty = d_ty_; ia.tz = d_tz_; ia.tq = d_tq_;
    ia.sx = d_sx_; ia.sy = d_sy_; ia.sz = d_sz_; ia.sq = d_sq_; ia.sw = d_sw_;
    ia.cx = d_cx_; ia.cy = d_cy_; ia.cz = d_cz_; ia.cq = d_cq_; ia.cw = d_cw_;
    {
        btk::InteractArgs a = ia;
        a.grid_targets = 0; a.ntiles = (int)near_tiles_.size(); a.tiles = d_near_tiles_;
        a.seg_off = d_nseg_off_; a.segs = d_nsegs_; a.clu_off = d_nclu_off_; a.clu_id = d_nclu_id_;
        a.out1 = d_pot_;
        if (p.precision == PREC_MIXED) a.key_ctr = reinterpret_cast<const double4 *>(d_near_ctr_);
        a.skip_sym = sym_ready() ? 1 : 0;
        std::copy(near_qlen_, near_qlen_ + NXCD, a.qlen);
        a.cursor = d_cursor_;
        HIPCHK(btk::launch_interact(a, st_));
        if (a.skip_sym) {
            std::string r = near_sym(!ov);
            if (!r.empty()) return r;
        }
    }
    HIPCHK(hipEventRecord(ev[2], st_));

    // 3. far field: Hermite PC clusters into the potential (after the near
    // launch has stored it); grids of target clusters (CP, CC)
    if (herm_ && p.compute_type == C_PC) {
        std::string r = hermite_pc();
        if (!r.empty()) return r;
    } else if (herm_ && p.compute_type == C_CP) {
        std::string r = hermite_cp();
        if (!r.empty()) return r;
    } else if (p.compute_type != C_PC) {
        if (!grid_tiles_.empty()) {
            btk::InteractArgs a = ia;
            // acceptance-separated sources: the centred expansion of r^2 (mixed)
            a.grid_scaled = p.precision == PREC_MIXED && p.theta < 0.99;
            a.grid_targets = 1; a.ntiles = (int)grid_tiles_.size(); a.tiles = d_grid_tiles_;
            a.seg_off = d_gseg_off_; a.segs = d_gsegs_; a.clu_off = d_gclu_off_; a.clu_id = d_gclu_id_;
            a.tnode_tab = d_ttab_;
            a.out1 = d_Qrows_; a.out2 = d_Wrows_; a.out_row = d_grid_row_;
            std::copy(grid_qlen_, grid_qlen_ + NXCD, a.qlen);
            a.cursor = d_cursor_ + NXCD;
            if (pencil_grids()) {
                a.ntiles = (int)pencil_tiles_.size(); a.tiles = d_pencil_tiles_;
                HIPCHK(btk::launch_grid_pencil(a, st_));
            } else {
                HIPCHK(btk::launch_interact(a, st_));
            }
        }
        // every target grid: the sum of its rows, in row order (zero without rows)
        HIPCHK(btk::launch_sum_rows(ntnodes_, p.P, d_node_rows_off_, d_grid_row_, d_Qrows_, d_Wrows_, d_Qt_,
                                    ss ? d_Wt_ : nullptr, st_));
    }
    HIPCHK(hipEventRecord(ev[3], st_));

    // 4. downpass over the target trees (interaction_compute_downpass.c:37-142)
    if (herm_ && p.compute_type == C_CP) {
        std::string r = hermite_down();
        if (!r.empty()) return r;
    } else if (p.compute_type != C_PC) {
        std::string r = down_fixed();
        if (!r.empty()) return r;
    }
    HIPCHK(hipEventRecord(ev[4], st_));
    if (!(er = join_phases(ov)).empty()) return er;

    // 5. subtraction correction and the caller's order (treedriver.c:589-592)
    HIPCHK(btk::launch_finish(M_, d_pot_, d_tq_, kp_.ss_param, d_torder_, d_out_, st_));
    HIPCHK(hipEventRecord(ev[5], st_));
    HIPCHK(hipEventSynchronize(ev[5]));
    if (timed) {
        float ms[5];
        for (int i = 0; i < 5; i++) HIPCHK(hipEventElapsedTime(&ms[i], ev[i], ev[i + 1]));
        tm_.gpu_upward = ms[0] * 1e-3; tm_.gpu_near = ms[1] * 1e-3; tm_.gpu_far = ms[2] * 1e-3;
        tm_.gpu_down = ms[3] * 1e-3; tm_.gpu_finish = ms[4] * 1e-3;
        tm_.gpu_total = (ms[0] + ms[1] + ms[2] + ms[3] + ms[4]) * 1e-3;
        float msym = 0.0f;
        if (sym_ran_) HIPCHK(hipEventElapsedTime(&msym, sym_ev_[0], sym_ev_[1]));
        tm_.gpu_near_sym = msym * 1e-3;
        if (ov) {
            // the items beside the other phases: the near phase counts them,
            // the total is the elapsed time from their start
            float mt = 0.0f, mf = 0.0f;
            HIPCHK(hipEventElapsedTime(&mt, sym_ev_[0], ev[5]));
            HIPCHK(hipEventElapsedTime(&mf, red_ev_, ev[5]));
            tm_.gpu_near += tm_.gpu_near_sym;
            tm_.gpu_finish = mf * 1e-3;
            tm_.gpu_total = mt * 1e-3;
        } else if (ahead) {
            // the item launch ran before ev[0]: count it in the near phase
            tm_.gpu_near += tm_.gpu_near_sym;
            tm_.gpu_total += tm_.gpu_near_sym;
        }
    }
    return "";
}

std::string Plan::download(double *out, bool device)
{
    HIPCHK(hipMemcpyAsync(out, d_out_, (size_t)M_ * sizeof(double),
                          device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, st_));
    HIPCHK(hipStreamSynchronize(st_));
    return "";
}

std::string Plan::direct(const std::vector<int> &idx, std::vector<double> &out)
{
    const int m = (int)idx.size();
    for (int k = 0; k < m; k++)
        if (idx[k] < 0 || idx[k] >= M_) return "target index out of range";
    out.assign(m, 0.0);
    if (m == 0) return "";
    if (!st_) return "the plan has no GPU work (host-only plan)";
    // position of each caller target in the sorted arrays, on the device
    int *dinv = dalloc<int>(M_);
    if (!dinv) return "hipMalloc failed";
    HIPCHK(btk::launch_invert(M_, d_torder_, dinv, st_));
    double *dw = d_sw_;
    std::string r;
    if (!dw) {
        std::vector<double> ones(N_, 1.0);
        if (!(r = up(&dw, ones)).empty()) return r;
    }
    // each target against the sources of its own system
    for (const System &y : sys_) {
        std::vector<int> which, sel;
        for (int k = 0; k < m; k++)
            if (idx[k] >= y.toff && idx[k] < y.toff + y.M) { which.push_back(k); sel.push_back(idx[k]); }
        if (which.empty()) continue;
        const int mm = (int)which.size();
        int *dsel = nullptr;
        if (!(r = up(&dsel, sel)).empty()) return r;
        double *dx = dalloc<double>(mm), *dy = dalloc<double>(mm), *dz = dalloc<double>(mm), *dq = dalloc<double>(mm);
        if (!dx || !dy || !dz || !dq) return "hipMalloc failed";
        HIPCHK(btk::launch_gather_sel(mm, dsel, dinv, d_tx_, d_ty_, d_tz_, d_tq_, dx, dy, dz, dq, st_));
        double *dout = dalloc<double>(mm);
        // enough source chunks for about 2048 blocks, each chunk at least 4096 sources
        const int tblocks = (mm + 255) / 256;
        const int nchunks = std::max(1, std::min((2048 + tblocks - 1) / tblocks, (y.N + 4095) / 4096));
        double *dpart = dalloc<double>((size_t)nchunks * mm);
        if (!dout || !dpart) return "hipMalloc failed";
        HIPCHK(btk::launch_direct(prm_.kernel, prm_.singularity, kp_, mm, dx, dy, dz, dq, y.N,
                                  d_sx_ + y.soff, d_sy_ + y.soff, d_sz_ + y.soff, d_sq_ + y.soff,
                                  dw + y.soff, dpart, nchunks, dout, st_));
        std::vector<double> res(mm);
        HIPCHK(hipMemcpyAsync(res.data(), dout, (size_t)mm * sizeof(double), hipMemcpyDeviceToHost, st_));
        HIPCHK(hipStreamSynchronize(st_));
        for (int k = 0; k < mm; k++) out[which[k]] = res[k];
        for (void *t : {(void *)dsel, (void *)dx, (void *)dy, (void *)dz, (void *)dq, (void *)dout, (void *)dpart})
            dfree(t);
    }
    dfree(dinv);
    if (dw != d_sw_) dfree(dw);
    return "";
}

// Caller indices sorted (with their slots) on the host; per system, its
// sorted subrange is found in tree order on the device (k_find_sel over the
// system's positions), gathered, and summed by the kernels direct() uses with
// the same chunking, so the results are the same bit for bit.  Scratch
// layout (each part 256-byte aligned): sel (m ints), pos (m ints), target
// x, y, z, q and the sums (m doubles each), the partial sums.
std::string Plan::direct_noalloc(int m, const int *idx, double *out, void *scratch, size_t bytes)
{
    if (m < 0 || m > btk::NOALLOC_MAX_SAMPLE) return "at most 8192 targets (the scratch's size)";
    for (int k = 0; k < m; k++)
        if (idx[k] < 0 || idx[k] >= M_) return "target index out of range";
    if (m == 0) return "";
    if (!st_) return "the plan has no GPU work (host-only plan)";
    // host work in fixed per-thread buffers: (index, slot) pairs sorted by index
    static thread_local std::array<std::pair<int, int>, btk::NOALLOC_MAX_SAMPLE> srt;
    static thread_local std::array<int, btk::NOALLOC_MAX_SAMPLE> sel;
    static thread_local std::array<double, btk::NOALLOC_MAX_SAMPLE> res;
    for (int k = 0; k < m; k++) srt[k] = {idx[k], k};
    std::sort(srt.begin(), srt.begin() + m);
    for (int k = 0; k < m; k++) sel[k] = srt[k].first;
    auto carve = [&](size_t &at, size_t n) {
        char *p = static_cast<char *>(scratch) + at;
        at += (n + 255) / 256 * 256;
        return p;
    };
    size_t at = 0;
    int *d_sel = reinterpret_cast<int *>(carve(at, (size_t)m * sizeof(int)));
    int *d_pos = reinterpret_cast<int *>(carve(at, (size_t)m * sizeof(int)));
    double *dx = reinterpret_cast<double *>(carve(at, (size_t)m * sizeof(double)));
    double *dy = reinterpret_cast<double *>(carve(at, (size_t)m * sizeof(double)));
    double *dz = reinterpret_cast<double *>(carve(at, (size_t)m * sizeof(double)));
    double *dq = reinterpret_cast<double *>(carve(at, (size_t)m * sizeof(double)));
    double *dout = reinterpret_cast<double *>(carve(at, (size_t)m * sizeof(double)));
    const size_t part_at = at;
    HIPCHK(hipMemcpyAsync(d_sel, sel.data(), (size_t)m * sizeof(int), hipMemcpyHostToDevice, st_));
    HIPCHK(btk::launch_fill_neg(m, d_pos, st_));
    int k0 = 0;
    for (const System &y : sys_) {
        // this system's part of the sorted indices
        int k1 = k0;
        while (k1 < m && sel[k1] < y.toff + y.M) k1++;
        const int mm = k1 - k0;
        if (mm > 0) {
            const int tblocks = (mm + 255) / 256;
            const int nchunks = std::max(1, std::min((2048 + tblocks - 1) / tblocks, (y.N + 4095) / 4096));
            if (part_at + (size_t)nchunks * mm * sizeof(double) > bytes) return "scratch too small";
            double *dpart = reinterpret_cast<double *>(static_cast<char *>(scratch) + part_at);
            HIPCHK(btk::launch_find_sel(y.toff, y.toff + y.M, d_torder_, mm, d_sel + k0, d_pos + k0, st_));
            HIPCHK(btk::launch_gather_pos(mm, d_pos + k0, d_tx_, d_ty_, d_tz_, d_tq_, dx + k0, dy + k0, dz + k0,
                                          dq + k0, st_));
            // without subtraction the kernels read no weights
            HIPCHK(btk::launch_direct(prm_.kernel, prm_.singularity, kp_, mm, dx + k0, dy + k0, dz + k0, dq + k0,
                                      y.N, d_sx_ + y.soff, d_sy_ + y.soff, d_sz_ + y.soff, d_sq_ + y.soff,
                                      d_sw_ ? d_sw_ + y.soff : nullptr, dpart, nchunks, dout + k0, st_));
        }
        k0 = k1;
    }
    HIPCHK(hipMemcpyAsync(res.data(), dout, (size_t)m * sizeof(double), hipMemcpyDeviceToHost, st_));
    HIPCHK(hipStreamSynchronize(st_));
    for (int k = 0; k < m; k++) out[srt[k].second] = res[k];
    return "";
}

}  // namespace bt
