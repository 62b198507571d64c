// bt_solver.cpp -- plan (host setup + upload) and execute (GPU hot path).
#include "bt_solver.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>

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

constexpr int TILE_TARGETS = 512;  // TPB*TPT of k_interact
constexpr int DOWN_TILE = 1024;    // targets per block of the leaf downpass

}  // namespace

Plan::~Plan() { free_all(); }

void Plan::free_all()
{
    for (void *p : allocs_) (void)hipFree(p);
    allocs_.clear();
    if (st_) {
        (void)hipStreamDestroy(st_);
        st_ = nullptr;
    }
}

template <class T> T *Plan::dalloc(size_t n)
{
    void *p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) return nullptr;
    allocs_.push_back(p);
    return static_cast<T *>(p);
}

template <class T, class A> std::string Plan::up(T **d, const std::vector<T, A> &h)
{
    *d = dalloc<T>(h.size());
    if (!*d) return "hipMalloc failed";
    if (!h.empty()) HIPCHK(hipMemcpyAsync(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, st_));
    return "";
}

// Flattens the mode-specific lists into two CSR work descriptions:
//   near_ : target leaves (particle targets) <- source particle ranges + source clusters
//   grid_ : target nodes (grid targets)      <- source particle ranges + source clusters
void Plan::build_work()
{
    const Params &p = prm_;
    const int P = p.P;
    auto add_key = [](WorkCSR &w, int key, int tbeg, int tcnt) {
        w.key.push_back(key); w.tbeg.push_back(tbeg); w.tcnt.push_back(tcnt);
        if (w.seg_off.empty()) { w.seg_off.push_back(0); w.clu_off.push_back(0); }
    };
    auto close_key = [](WorkCSR &w) {
        w.seg_off.push_back((int)w.seg_beg.size());
        w.clu_off.push_back((int)w.clu_id.size());
    };
    auto add_seg = [](WorkCSR &w, int b, int c, long long nt, bool check = false) {
        w.seg_beg.push_back(b); w.seg_cnt.push_back(c); w.seg_chk.push_back(check ? 1 : 0);
        w.pairs += nt * c;
    };
    // May a source of node s coincide with a target of node t?  The
    // singularity skip can only trigger when their closed boxes overlap (up to
    // a 1e-150 gap: any pair across a wider gap has r^2 > DBL_MIN).
    auto touch = [this](int t, int s) {
        const double g = 1e-150;
        return ttree_.xmin[t] - stree_.xmax[s] <= g && stree_.xmin[s] - ttree_.xmax[t] <= g &&
               ttree_.ymin[t] - stree_.ymax[s] <= g && stree_.ymin[s] - ttree_.ymax[t] <= g &&
               ttree_.zmin[t] - stree_.zmax[s] <= g && stree_.zmin[s] - ttree_.zmax[t] <= g;
    };
    auto add_clu = [P](WorkCSR &w, int id, long long nt) { w.clu_id.push_back(id); w.pairs += nt * P; };
    near_ = WorkCSR();
    grid_ = WorkCSR();
    near_.seg_off.push_back(0); near_.clu_off.push_back(0);
    grid_.seg_off.push_back(0); grid_.clu_off.push_back(0);

    if (p.compute_type == C_PC) {
        // keys: target batches; pc list -> clusters, pp list -> source leaves
        for (int b = 0; b < ttree_.n; b++) {
            const long long nt = ttree_.numpar[b];
            add_key(near_, b, ttree_.ibeg[b], ttree_.numpar[b]);
            for (int c : L_.approx[b]) add_clu(near_, c, nt);
            for (int s : L_.pp[b]) add_seg(near_, stree_.ibeg[s], stree_.numpar[s], nt, touch(b, s));
            close_key(near_);
        }
    } else if (p.compute_type == C_CP) {
        // lists keyed by source batch; invert to target-node keys
        std::vector<std::vector<int>> pp_of(ttree_.n), cp_of(ttree_.n);
        for (int b = 0; b < stree_.n; b++) {
            for (int t : L_.pp[b]) pp_of[t].push_back(b);
            for (int t : L_.approx[b]) cp_of[t].push_back(b);
        }
        for (int leaf : ttree_.leaves) {
            const long long nt = ttree_.numpar[leaf];
            add_key(near_, leaf, ttree_.ibeg[leaf], ttree_.numpar[leaf]);
            for (int b : pp_of[leaf]) add_seg(near_, stree_.ibeg[b], stree_.numpar[b], nt, touch(leaf, b));
            close_key(near_);
        }
        for (int t = 0; t < ttree_.n; t++) {
            if (cp_of[t].empty()) continue;
            add_key(grid_, t, 0, P);
            for (int b : cp_of[t]) add_seg(grid_, stree_.ibeg[b], stree_.numpar[b], P);
            close_key(grid_);
        }
    } else {
        // CC: a target leaf gets the PP and source-approx (PC) lists of
        // itself and of every ancestor; grids get CC and target-approx (CP)
        for (int leaf : ttree_.leaves) {
            const long long nt = ttree_.numpar[leaf];
            add_key(near_, leaf, ttree_.ibeg[leaf], ttree_.numpar[leaf]);
            for (int a = leaf; a >= 0; a = ttree_.parent[a]) {
                for (int c : L_.cc_pc[a]) add_clu(near_, c, nt);
                for (int s : L_.pp[a]) add_seg(near_, stree_.ibeg[s], stree_.numpar[s], nt, touch(leaf, s));
            }
            close_key(near_);
        }
        for (int t = 0; t < ttree_.n; t++) {
            if (L_.approx[t].empty() && L_.cc_cp[t].empty()) continue;
            add_key(grid_, t, 0, P);
            for (int c : L_.approx[t]) add_clu(grid_, c, P);
            for (int s : L_.cc_cp[t]) add_seg(grid_, stree_.ibeg[s], stree_.numpar[s], P);
            close_key(grid_);
        }
    }

    // Tiles.  All tiles of one key read the same sources, so they should run
    // at the same time on the same XCD and share its L2.  Workgroups are
    // dispatched round-robin over the 8 XCDs (block b runs on XCD b % 8), so
    // keys (groups of tiles) are dealt to per-XCD lists, heaviest group to the
    // least loaded XCD, and block 8*j + x takes the j-th tile of XCD x's list.
    // Lists are padded with empty tiles (count 0), which return at once.
    auto make_tiles = [&](const WorkCSR &w, bool grid, std::vector<btk::Tile> &tiles) {
        constexpr int NXCD = 8;
        struct Group { long long work; std::vector<btk::Tile> t; };
        std::vector<Group> groups;
        for (size_t k = 0; k < w.key.size(); k++) {
            long long src = 0;
            for (int s = w.seg_off[k]; s < w.seg_off[k + 1]; s++) src += w.seg_cnt[s];
            src += (long long)(w.clu_off[k + 1] - w.clu_off[k]) * P;
            Group g{0, {}};
            for (int o = 0; o < w.tcnt[k]; o += TILE_TARGETS) {
                btk::Tile t;
                t.x = (int)k;
                t.y = w.tbeg[k] + o;
                t.z = std::min(TILE_TARGETS, w.tcnt[k] - o);
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
    };
    make_tiles(near_, false, near_tiles_);
    make_tiles(grid_, true, grid_tiles_);

    // upward pass (source tree, PC and CC): non-leaf nodes per level, deepest first
    up_levels_off_.assign(1, 0);
    up_nodes_.clear(); up_child_off_.assign(1, 0); up_child_id_.clear();
    if (p.compute_type != C_CP) {
        for (int lv = stree_.max_depth - 1; lv >= 0; --lv) {
            for (int par : stree_.levels[lv]) {
                up_nodes_.push_back(par);
                for (int k = 0; k < stree_.nchild[par]; k++) up_child_id_.push_back(stree_.child[8 * par + k]);
                up_child_off_.push_back((int)up_child_id_.size());
            }
            up_levels_off_.push_back((int)up_nodes_.size());
        }
    }
    // downward pass (target tree, CP and CC): children per level, top first
    down_levels_off_.assign(1, 0);
    down_nodes_.clear();
    leaf_tiles_.clear();
    if (p.compute_type != C_PC) {
        for (int lv = 0; lv < ttree_.max_depth; ++lv) {
            for (int par : ttree_.levels[lv])
                for (int k = 0; k < ttree_.nchild[par]; k++) down_nodes_.push_back(ttree_.child[8 * par + k]);
            down_levels_off_.push_back((int)down_nodes_.size());
        }
        for (int leaf : ttree_.leaves) {
            for (int o = 0; o < ttree_.numpar[leaf]; o += DOWN_TILE) {
                btk::Tile t;
                t.x = leaf;
                t.y = ttree_.ibeg[leaf] + o;
                t.z = std::min(DOWN_TILE, ttree_.numpar[leaf] - o);
                t.w = 0;
                leaf_tiles_.push_back(t);
            }
        }
    }
}

// Every index a kernel will dereference is checked here, on the host,
// before anything is launched.
std::string Plan::validate() const
{
    const int P = prm_.P;
    auto chk_csr = [&](const WorkCSR &w, const std::vector<btk::Tile> &tiles, bool grid,
                       const Tree &src_tree) -> std::string {
        const int nk = (int)w.key.size();
        if ((int)w.seg_off.size() != nk + 1 || (int)w.clu_off.size() != nk + 1) return "csr offsets";
        if (w.seg_off[nk] != (int)w.seg_beg.size() || w.clu_off[nk] != (int)w.clu_id.size()) return "csr ends";
        for (int k = 0; k < nk; k++)
            if (w.seg_off[k] > w.seg_off[k + 1] || w.clu_off[k] > w.clu_off[k + 1]) return "csr order";
        for (size_t i = 0; i < w.seg_beg.size(); i++)
            if (w.seg_beg[i] < 0 || w.seg_cnt[i] < 0 || (long long)w.seg_beg[i] + w.seg_cnt[i] > N_)
                return "source range";
        for (int c : w.clu_id)
            if (c < 0 || c >= src_tree.n) return "cluster id";
        for (const btk::Tile &t : tiles) {
            if (t.z == 0) continue;  // padding
            if (t.x < 0 || t.x >= nk || t.z < 0 || t.z > TILE_TARGETS || t.y < 0) return "tile";
            if (grid) {
                if (t.w < 0 || t.w >= ttree_.n || t.y + t.z > P) return "grid tile";
            } else if ((long long)t.y + t.z > M_) {
                return "target tile";
            }
        }
        return "";
    };
    std::string e = chk_csr(near_, near_tiles_, false, stree_);
    if (!e.empty()) return "near " + e;
    if (!(e = chk_csr(grid_, grid_tiles_, true, stree_)).empty()) return "far " + e;
    for (int i = 0; i < stree_.n; i++)
        if (stree_.ibeg[i] < 0 || stree_.iend[i] > N_ || stree_.ibeg[i] >= stree_.iend[i]) return "source node range";
    for (int i = 0; i < ttree_.n; i++)
        if (ttree_.ibeg[i] < 0 || ttree_.iend[i] > M_ || ttree_.ibeg[i] >= ttree_.iend[i]) return "target node range";
    for (int c : up_child_id_)
        if (c < 0 || c >= stree_.n) return "upward child";
    for (int c : down_nodes_)
        if (c < 0 || c >= ttree_.n || ttree_.parent[c] < 0) return "downward child";
    for (const btk::Tile &t : leaf_tiles_)
        if (t.x < 0 || t.x >= ttree_.n || t.y < 0 || (long long)t.y + t.z > M_ || t.z <= 0) return "leaf tile";
    for (int i = 0; i < M_; i++)
        if (T_.order[i] < 0 || T_.order[i] >= M_) return "target order";
    return "";
}

std::string Plan::create(const Params &prm, int M, int N, const double *tx, const double *ty,
                         const double *tz, const double *tq, const double *sx, const double *sy,
                         const double *sz, const double *sq, const double *sw, bool host_only)
{
    free_all();
    prm_ = prm;
    M_ = M; N_ = N;
    std::string e = params_check(prm);
    if (!e.empty()) return e;
    if (M <= 0 || N <= 0) return "numTargets and numSources must be positive";
    if (!tx || !ty || !tz || !sx || !sy || !sz || !sq) return "One or more required particle arrays is unset";
    if (prm.singularity == S_SUBTRACTION && !sw && !host_only)
        return "Singularity subtraction requires weights for the sources";
    if (!host_only) HIPCHK(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking));

    double t0 = now();
    // private copies (the caller's arrays are never permuted), in parallel
    S_ = Particles(); T_ = Particles();
    S_.n = N; T_.n = M;
    for (auto *v : {&S_.x, &S_.y, &S_.z, &S_.q, &S_.w}) v->resize(N);
    for (auto *v : {&T_.x, &T_.y, &T_.z, &T_.q}) v->resize(M);
    S_.order.resize(N);
    T_.order.resize(M);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; i++) {
        S_.x[i] = sx[i]; S_.y[i] = sy[i]; S_.z[i] = sz[i]; S_.q[i] = sq[i];
        S_.w[i] = sw ? sw[i] : 1.0;
        S_.order[i] = i;
    }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < M; i++) {
        T_.x[i] = tx[i]; T_.y[i] = ty[i]; T_.z[i] = tz[i]; T_.q[i] = tq ? tq[i] : 0.0;
        T_.order[i] = i;
    }
    double t1 = now();
    tm_ = Timings();
    tm_.host_copy = t1 - t0;

    // trees (treedriver.c:128/146, 364/375, 654/665)
    if (prm.compute_type == C_PC) {
        build_trees({{&stree_, &S_, prm.max_source_leaf, false}, {&ttree_, &T_, prm.max_target_leaf, true}});
    } else if (prm.compute_type == C_CP) {
        build_trees({{&ttree_, &T_, prm.max_target_leaf, false}, {&stree_, &S_, prm.max_source_leaf, true}});
    } else {
        build_trees({{&stree_, &S_, prm.max_source_leaf, false}, {&ttree_, &T_, prm.max_target_leaf, false}});
    }
    double t2 = now();
    tm_.host_tree = t2 - t1;

    // interaction lists (treedriver.c:181, 462, 753)
    if (prm.compute_type == C_PC) make_pc_lists(L_, stree_, ttree_, prm);
    else if (prm.compute_type == C_CP) make_pc_lists(L_, ttree_, stree_, prm);
    else make_cc_lists(L_, stree_, ttree_, prm);
    double t3 = now();
    tm_.host_lists = t3 - t2;

    // reference diagnostics (treedriver.c:191-200, 472-483, 763-792)
    counts_ = Counts();
    {
        const long long P = prm.P;
        const Tree &key = prm.compute_type == C_CP ? stree_ : ttree_;
        const Tree &ent = prm.compute_type == C_CP ? ttree_ : stree_;
        for (int i = 0; i < key.n; i++) {
            const long long nt = key.numpar[i];
            counts_.num_pp += (long long)L_.pp[i].size();
            for (int s : L_.pp[i]) counts_.pp_ptwise += nt * ent.numpar[s];
            if (prm.compute_type == C_PC) {
                counts_.num_pc += (long long)L_.approx[i].size();
                for (int s : L_.approx[i]) { counts_.pc_ptwise += nt * P; counts_.pc_replaced += nt * ent.numpar[s]; }
            } else if (prm.compute_type == C_CP) {
                counts_.num_cp += (long long)L_.approx[i].size();
                for (int s : L_.approx[i]) { counts_.cp_ptwise += nt * P; counts_.cp_replaced += nt * ent.numpar[s]; }
            } else {
                counts_.num_cc += (long long)L_.approx[i].size();
                counts_.num_pc += (long long)L_.cc_pc[i].size();
                counts_.num_cp += (long long)L_.cc_cp[i].size();
                for (int s : L_.approx[i]) { counts_.cc_ptwise += P * P; counts_.cc_replaced += nt * ent.numpar[s]; }
                for (int s : L_.cc_pc[i]) { counts_.pc_ptwise += nt * P; counts_.pc_replaced += nt * ent.numpar[s]; }
                for (int s : L_.cc_cp[i]) { counts_.cp_ptwise += P * ent.numpar[s]; counts_.cp_replaced += nt * ent.numpar[s]; }
            }
        }
    }

    build_work();
    std::vector<double> stab, ttab;
    const bool need_src_clusters = prm.compute_type != C_CP;
    const bool need_tgt_clusters = prm.compute_type != C_PC;
    if (need_src_clusters) node_table(stree_, prm.degree, stab);
    if (need_tgt_clusters) node_table(ttree_, prm.degree, ttab);
    double t4 = now();
    tm_.host_work = t4 - t3;
    if (!(e = validate()).empty()) return "internal work-list check failed: " + e;
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

    // upload
    const bool ss = prm.singularity == S_SUBTRACTION;
    std::string r;
    if (!(r = up(&d_tx_, T_.x)).empty()) return r;
    if (!(r = up(&d_ty_, T_.y)).empty()) return r;
    if (!(r = up(&d_tz_, T_.z)).empty()) return r;
    if (!(r = up(&d_tq_, T_.q)).empty()) return r;
    if (!(r = up(&d_sx_, S_.x)).empty()) return r;
    if (!(r = up(&d_sy_, S_.y)).empty()) return r;
    if (!(r = up(&d_sz_, S_.z)).empty()) return r;
    if (!(r = up(&d_sq_, S_.q)).empty()) return r;
    if (ss && !(r = up(&d_sw_, S_.w)).empty()) return r;
    if (!(r = up(&d_torder_, T_.order)).empty()) return r;
    const size_t P = prm.P;
    if (need_src_clusters) {
        if (!(r = up(&d_stab_, stab)).empty()) return r;
        const size_t np = (size_t)stree_.n * P;
        d_cx_ = dalloc<double>(np); d_cy_ = dalloc<double>(np); d_cz_ = dalloc<double>(np);
        d_cq_ = dalloc<double>(np);
        if (ss) d_cw_ = dalloc<double>(np);
        if (!d_cx_ || !d_cy_ || !d_cz_ || !d_cq_ || (ss && !d_cw_)) return "hipMalloc failed (source clusters)";
        if (!(r = up(&d_sibeg_, stree_.ibeg)).empty()) return r;
        if (!(r = up(&d_siend_, stree_.iend)).empty()) return r;
        if (!(r = up(&d_sleaves_, stree_.leaves)).empty()) return r;
        if (!(r = up(&d_up_nodes_, up_nodes_)).empty()) return r;
        if (!(r = up(&d_up_child_off_, up_child_off_)).empty()) return r;
        if (!(r = up(&d_up_child_id_, up_child_id_)).empty()) return r;
    }
    if (need_tgt_clusters) {
        if (!(r = up(&d_ttab_, ttab)).empty()) return r;
        const size_t np = (size_t)ttree_.n * P;
        d_Qt_ = dalloc<double>(np);
        if (ss) d_Wt_ = dalloc<double>(np);
        if (!d_Qt_ || (ss && !d_Wt_)) return "hipMalloc failed (target clusters)";
        if (!(r = up(&d_tparent_, ttree_.parent)).empty()) return r;
        if (!(r = up(&d_down_nodes_, down_nodes_)).empty()) return r;
        if (!(r = up(&d_leaf_tiles_, leaf_tiles_)).empty()) return r;
    }
    auto up_csr = [&](const WorkCSR &w, int **so, btk::Seg **sg, int **co, int **ci) -> std::string {
        std::vector<btk::Seg> segs(w.seg_beg.size());
        for (size_t i = 0; i < segs.size(); i++) {
            segs[i].x = w.seg_beg[i];
            segs[i].y = (int)((unsigned)w.seg_cnt[i] | (w.seg_chk[i] ? 0x80000000u : 0u));
        }
        std::string rr;
        if (!(rr = up(so, w.seg_off)).empty()) return rr;
        if (!(rr = up(sg, segs)).empty()) return rr;
        if (!(rr = up(co, w.clu_off)).empty()) return rr;
        return up(ci, w.clu_id);
    };
    if (!(r = up_csr(near_, &d_nseg_off_, &d_nsegs_, &d_nclu_off_, &d_nclu_id_)).empty()) return r;
    if (!(r = up(&d_near_tiles_, near_tiles_)).empty()) return r;
    if (!grid_tiles_.empty()) {
        if (!(r = up_csr(grid_, &d_gseg_off_, &d_gsegs_, &d_gclu_off_, &d_gclu_id_)).empty()) return r;
        if (!(r = up(&d_grid_tiles_, grid_tiles_)).empty()) return r;
    }
    d_pot_ = dalloc<double>(M);
    d_out_ = dalloc<double>(M);
    if (!d_pot_ || !d_out_) return "hipMalloc failed (potential)";
    HIPCHK(hipStreamSynchronize(st_));
    tm_.h2d = now() - t4;
    return "";
}

std::string Plan::execute(bool timed)
{
    const Params &p = prm_;
    const bool ss = p.singularity == S_SUBTRACTION;
    const int nl = p.degree + 1;
    hipEvent_t ev[6];
    for (auto &e : ev) HIPCHK(hipEventCreate(&e));
    HIPCHK(hipEventRecord(ev[0], st_));

    // 1. upward pass over the source tree (clusters.c:50-200)
    if (p.compute_type != C_CP) {
        HIPCHK(btk::launch_cluster_points(stree_.n, nl, d_stab_, d_cx_, d_cy_, d_cz_, st_));
        btk::AnterpArgs a{};
        a.degree = p.degree; a.subtraction = ss;
        a.ibeg = d_sibeg_; a.iend = d_siend_; a.tab = d_stab_;
        a.px = d_sx_; a.py = d_sy_; a.pz = d_sz_; a.pq = d_sq_; a.pw = d_sw_;
        a.Q = d_cq_; a.W = d_cw_;
        a.mode = 0; a.nodes = d_sleaves_; a.nblocks = (int)stree_.leaves.size();
        HIPCHK(btk::launch_anterp(a, st_));
        for (size_t lv = 0; lv + 1 < up_levels_off_.size(); lv++) {
            const int b = up_levels_off_[lv], e = up_levels_off_[lv + 1];
            if (e == b) continue;
            a.mode = 1; a.nodes = d_up_nodes_ + b; a.nblocks = e - b;
            a.child_off = d_up_child_off_ + b; a.child_id = d_up_child_id_;
            HIPCHK(btk::launch_anterp(a, st_));
        }
    }
    HIPCHK(hipEventRecord(ev[1], st_));

    // 2. near field and particle-cluster far field into the potential
    btk::InteractArgs ia{};
    ia.kernel = p.kernel; ia.singularity = p.singularity; ia.precision = p.precision;
    ia.P = p.P; ia.nl = nl; ia.kp = kp_;
    ia.tx = d_tx_; ia.ty = d_ty_; ia.tz = d_tz_; ia.tq = d_tq_;
    ia.sx = d_sx_; ia.sy = d_sy_; ia.sz = d_sz_; ia.sq = d_sq_; ia.sw = d_sw_;
    ia.cx = d_cx_; ia.cy = d_cy_; ia.cz = d_cz_; ia.cq = d_cq_; ia.cw = d_cw_;
    {
        btk::InteractArgs a = ia;
        a.grid_targets = 0; a.ntiles = (int)near_tiles_.size(); a.tiles = d_near_tiles_;
        a.seg_off = d_nseg_off_; a.segs = d_nsegs_; a.clu_off = d_nclu_off_; a.clu_id = d_nclu_id_;
        a.out1 = d_pot_;
        HIPCHK(btk::launch_interact(a, st_));
    }
    HIPCHK(hipEventRecord(ev[2], st_));

    // 3. far field into target-cluster grids (CP, CC)
    if (p.compute_type != C_PC) {
        const size_t np = (size_t)ttree_.n * p.P;
        HIPCHK(hipMemsetAsync(d_Qt_, 0, np * sizeof(double), st_));
        if (ss) HIPCHK(hipMemsetAsync(d_Wt_, 0, np * sizeof(double), st_));
        if (!grid_tiles_.empty()) {
            btk::InteractArgs a = ia;
            a.grid_targets = 1; a.ntiles = (int)grid_tiles_.size(); a.tiles = d_grid_tiles_;
            a.seg_off = d_gseg_off_; a.segs = d_gsegs_; a.clu_off = d_gclu_off_; a.clu_id = d_gclu_id_;
            a.tnode_tab = d_ttab_;
            a.out1 = d_Qt_; a.out2 = d_Wt_;
            HIPCHK(btk::launch_interact(a, st_));
        }
    }
    HIPCHK(hipEventRecord(ev[3], st_));

    // 4. downpass over the target tree (interaction_compute_downpass.c:37-142)
    if (p.compute_type != C_PC) {
        btk::DownArgs d{};
        d.degree = p.degree; d.subtraction = ss; d.tab = d_ttab_; d.parent = d_tparent_;
        d.tx = d_tx_; d.ty = d_ty_; d.tz = d_tz_; d.tq = d_tq_; d.Q = d_Qt_; d.W = d_Wt_; d.pot = d_pot_;
        for (size_t lv = 0; lv + 1 < down_levels_off_.size(); lv++) {
            const int b = down_levels_off_[lv], e = down_levels_off_[lv + 1];
            if (e == b) continue;
            d.mode = 0; d.nodes = d_down_nodes_ + b; d.nblocks = e - b;
            HIPCHK(btk::launch_down(d, st_));
        }
        d.mode = 1; d.tiles = d_leaf_tiles_; d.nblocks = (int)leaf_tiles_.size();
        HIPCHK(btk::launch_down(d, st_));
    }
    HIPCHK(hipEventRecord(ev[4], st_));

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
    }
    for (auto &e : ev) (void)hipEventDestroy(e);
    return "";
}

std::string Plan::download(double *out)
{
    HIPCHK(hipMemcpyAsync(out, d_out_, (size_t)M_ * sizeof(double), hipMemcpyDeviceToHost, st_));
    HIPCHK(hipStreamSynchronize(st_));
    return "";
}

std::string Plan::direct(const std::vector<int> &idx, std::vector<double> &out)
{
    // targets in caller order: gather from the sorted copy through order
    const int m = (int)idx.size();
    for (int k = 0; k < m; k++)
        if (idx[k] < 0 || idx[k] >= M_) return "target index out of range";
    if (m == 0) { out.clear(); return ""; }
    std::vector<int> inv(M_);
    for (int i = 0; i < M_; i++) inv[T_.order[i]] = i;
    std::vector<double> hx(m), hy(m), hz(m), hq(m);
    for (int k = 0; k < m; k++) {
        const int i = inv[idx[k]];
        hx[k] = T_.x[i]; hy[k] = T_.y[i]; hz[k] = T_.z[i]; hq[k] = T_.q[i];
    }
    double *dx = nullptr, *dy = nullptr, *dz = nullptr, *dq = nullptr, *dout = nullptr, *dw = d_sw_;
    std::string r;
    if (!(r = up(&dx, hx)).empty()) return r;
    if (!(r = up(&dy, hy)).empty()) return r;
    if (!(r = up(&dz, hz)).empty()) return r;
    if (!(r = up(&dq, hq)).empty()) return r;
    dout = dalloc<double>(m);
    if (!dout) return "hipMalloc failed";
    if (!dw) {
        std::vector<double> ones(N_, 1.0);
        if (!(r = up(&dw, ones)).empty()) return r;
    }
    // enough source chunks for about 2048 blocks, each chunk at least 4096 sources
    const int tblocks = (m + 255) / 256;
    int nchunks = std::max(1, std::min((2048 + tblocks - 1) / tblocks, (N_ + 4095) / 4096));
    double *dpart = dalloc<double>((size_t)nchunks * m);
    if (!dpart) return "hipMalloc failed";
    HIPCHK(btk::launch_direct(prm_.kernel, prm_.singularity, kp_, m, dx, dy, dz, dq, N_, d_sx_, d_sy_,
                              d_sz_, d_sq_, dw, dpart, nchunks, dout, st_));
    out.resize(m);
    HIPCHK(hipMemcpyAsync(out.data(), dout, (size_t)m * sizeof(double), hipMemcpyDeviceToHost, st_));
    HIPCHK(hipStreamSynchronize(st_));
    return "";
}

}  // namespace bt
