// bt_vardeg.cpp -- far-field pairs with their own interpolation degree.
//
// The fixed-degree path follows the reference: one degree for every cluster,
// charges passed child to parent (src/clusters/clusters.c:50-200), grid
// values passed parent to child (interaction_compute_downpass.c:37-142).
// With a degree per pair that chain breaks (a parent and its child may be
// used at different degrees), so here
//   - a source node used at degree p gets modified charges of degree p
//     computed straight from its particles (source slot (node, p)),
//   - a target node receiving at degree p gets a grid of degree p (target
//     slot (node, p)) whose interpolant is evaluated straight at the node's
//     particles.
// Both are exact rewrites of the chains: a degree-p interpolant reproduces
// every polynomial of degree <= p, so the Lagrange basis of the parent grid
// is reproduced by the child grid and the two orders give the same sums
// (up to rounding).  With every pair at the plan's degree the result is the
// fixed-degree result.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#include "bt_learn.h"
#include "bt_devmem.h"
#include "bt_solver.h"

namespace bt {

namespace {

#define HIPCHK(expr)                                                                 \
    do {                                                                             \
        hipError_t e_ = (expr);                                                      \
        if (e_ != hipSuccess) return std::string(#expr ": ") + hipGetErrorString(e_); \
    } while (0)

constexpr int NQ = btk::TILE_QUEUES;
constexpr int MAXDEG = 12;

}  // namespace

void Plan::enumerate_pairs()
{
    pairs_.clear();
    for (int sy = 0; sy < (int)sys_.size(); sy++) {
        const System &y = sys_[sy];
        const Lists &L = y.L;
        if (prm_.compute_type == C_PC) {
            for (int b = 0; b < y.ttree.n; b++)
                for (int c : L.approx[b]) pairs_.push_back({sy, FAR_PC, b, c});
        } else if (prm_.compute_type == C_CP) {
            for (int b = 0; b < y.stree.n; b++)
                for (int t : L.approx[b]) pairs_.push_back({sy, FAR_CP, t, b});
        } else {
            for (int t = 0; t < y.ttree.n; t++) {
                for (int s : L.approx[t]) pairs_.push_back({sy, FAR_CC, t, s});
                for (int s : L.cc_pc[t]) pairs_.push_back({sy, FAR_PC, t, s});
                for (int s : L.cc_cp[t]) pairs_.push_back({sy, FAR_CP, t, s});
            }
        }
    }
}

// Features of pair (t, s), all scale free:
//   0 r_t/d   1 r_s/d      (node radii over the distance of the node centres,
//                           the quantities of the acceptance criterion)
//   2 log2(n_t)/16  3 log2(n_s)/16   (particle counts)
//   4 [kind == PC]  5 [kind == CP]   (which side is interpolated; CC: both)
//   6 kernel parameter at the pair's scale: Yukawa kappa*d, regularized
//     Coulomb epsilon/d, Coulomb 0
//   7 largest half-width of the interpolated box(es) over d
void Plan::pair_features(std::vector<float> &f) const
{
    const int nf = NUM_PAIR_FEATURES;
    f.assign(pairs_.size() * nf, 0.0f);
    const bool ss = prm_.singularity == S_SUBTRACTION;
    double par = 0.0;
    int pk = 0;  // 0 none, 1 multiply by d, 2 divide by d
    if (prm_.kernel == K_YUKAWA) { par = prm_.kp[0]; pk = 1; }
    if (prm_.kernel == K_REG_COULOMB) { par = ss ? prm_.kp[1] : prm_.kp[0]; pk = 2; }
    const long long np = (long long)pairs_.size();
#pragma omp parallel for schedule(static) if (np > 20000)
    for (long long i = 0; i < np; i++) {
        const FarPair &q = pairs_[i];
        const Tree &tt = sys_[q.sys].ttree, &st = sys_[q.sys].stree;
        const int t = q.t, s = q.s;
        const double dx = tt.xmid[t] - st.xmid[s], dy = tt.ymid[t] - st.ymid[s],
                     dz = tt.zmid[t] - st.zmid[s];
        const double d = std::max(std::sqrt(dx * dx + dy * dy + dz * dz), 1e-300);
        auto halfw = [](const Tree &T, int n) {
            return 0.5 * std::max({T.xmax[n] - T.xmin[n], T.ymax[n] - T.ymin[n], T.zmax[n] - T.zmin[n]});
        };
        double hw = 0.0;
        if (q.kind != FAR_CP) hw = std::max(hw, halfw(st, s));
        if (q.kind != FAR_PC) hw = std::max(hw, halfw(tt, t));
        float *o = f.data() + i * nf;
        o[0] = (float)(tt.radius[t] / d);
        o[1] = (float)(st.radius[s] / d);
        o[2] = (float)(std::log2((double)std::max(1, tt.numpar[t])) / 16.0);
        o[3] = (float)(std::log2((double)std::max(1, st.numpar[s])) / 16.0);
        o[4] = q.kind == FAR_PC ? 1.0f : 0.0f;
        o[5] = q.kind == FAR_CP ? 1.0f : 0.0f;
        o[6] = (float)(pk == 1 ? par * d : (pk == 2 ? par / d : 0.0));
        o[7] = (float)(hw / d);
    }
}

void Plan::free_var()
{
    if (st_ && !var_allocs_.empty()) (void)hipStreamSynchronize(st_);
    for (void *p : var_allocs_) btm::release(p);
    var_allocs_.clear();
    dw_.clear();
    vnear_ = WorkCSR();
    vfar_ = WorkCSR();
    vnear_tiles_.clear();
    vfar_tiles_.clear();
    vpencil_tiles_.clear();
    d_vp_tiles_ = nullptr;
    vfar_row_.clear();
    vfar_rows_len_ = 0;
    var_ = false;
    var_planned_ = false;
    var_mean_deg_ = 0;
    var_evals_ = 0;
    d_vn_clu_off_ = d_vn_clu_id_ = nullptr;
    d_vn_tiles_ = nullptr;
    d_vf_seg_off_ = d_vf_clu_off_ = d_vf_clu_id_ = nullptr;
    d_vf_segs_ = nullptr;
    d_vf_tiles_ = nullptr;
    d_vf_row_ = nullptr;
    d_vf_R_ = d_vf_RW_ = nullptr;
    fold_up_ = fold_down_ = false;
    fold_node_.clear(); fold_off_.clear(); fold_item_.clear();
    emat_.clear(); emat_off_.clear();
    d_fold_node_ = d_fold_off_ = d_fold_item_ = nullptr;
    d_emat_ = nullptr;
}

namespace {
inline int multi_id(int deg, int slot) { return (deg << btk::DEG_SHIFT) | slot; }
}  // namespace

std::string Plan::set_pair_degrees(const int *deg)
{
    free_var();
    if (!deg) return "";
    if (herm_) return "per-pair degrees need the Lagrange approximation";
    if (!coulomb_family(prm_.kernel))
        return "per-pair degrees are set up for the Coulomb, Yukawa and regularized Coulomb kernels";
    const long long np = (long long)pairs_.size();
    for (long long i = 0; i < np; i++)
        if (deg[i] < 1 || deg[i] > MAXDEG) return "pair degree outside [1, 12]";
    const bool ss = prm_.singularity == S_SUBTRACTION;

    std::vector<std::vector<int>> sslot(MAXDEG + 1), tslot(MAXDEG + 1);
    std::vector<int> used(MAXDEG + 1, 0);
    for (long long i = 0; i < np; i++) used[deg[i]] = 1;
    dw_.assign(MAXDEG + 1, DegreeWork());
    for (int p = 1; p <= MAXDEG; p++) {
        dw_[p].degree = p; dw_[p].nl = p + 1; dw_[p].P = (p + 1) * (p + 1) * (p + 1);
        dw_[p].used = used[p] != 0;
        if (used[p]) { sslot[p].assign(nsnodes_, -1); tslot[p].assign(ntnodes_, -1); }
    }
    if (nsnodes_ > btk::SLOT_MASK || ntnodes_ > btk::SLOT_MASK) return "too many tree nodes for degree slots";
    // node tables per system and degree, computed on demand
    std::vector<std::vector<std::vector<double>>> stab(sys_.size()), ttab(sys_.size());
    for (size_t y = 0; y < sys_.size(); y++) {
        stab[y].resize(MAXDEG + 1);
        ttab[y].resize(MAXDEG + 1);
    }
    auto src_slot = [&](int sy, int s, int p) {
        const System &y = sys_[sy];
        int &sl = sslot[p][y.snode_off + s];
        if (sl < 0) {
            DegreeWork &w = dw_[p];
            sl = (int)w.s_ibeg.size();
            w.s_node.push_back(y.snode_off + s);
            w.s_ibeg.push_back(y.soff + y.stree.ibeg[s]);
            w.s_iend.push_back(y.soff + y.stree.iend[s]);
            if (stab[sy][p].empty()) node_table(y.stree, p, stab[sy][p]);
            const double *row = stab[sy][p].data() + (size_t)s * 3 * (p + 1);
            w.s_tab.insert(w.s_tab.end(), row, row + 3 * (p + 1));
        }
        return sl;
    };
    std::vector<std::vector<int>> slot_node(MAXDEG + 1), slot_sys(MAXDEG + 1);
    auto tgt_slot = [&](int sy, int t, int p) {
        const System &y = sys_[sy];
        int &sl = tslot[p][y.tnode_off + t];
        if (sl < 0) {
            DegreeWork &w = dw_[p];
            sl = (int)slot_node[p].size();
            slot_node[p].push_back(t);
            slot_sys[p].push_back(sy);
            if (ttab[sy][p].empty()) node_table(y.ttree, p, ttab[sy][p]);
            const double *row = ttab[sy][p].data() + (size_t)t * 3 * (p + 1);
            w.t_tab.insert(w.t_tab.end(), row, row + 3 * (p + 1));
        }
        return sl;
    };
    // PC entries per global target node; CC clusters and CP ranges per
    // target slot; all in pair order
    std::vector<std::vector<int>> pc_at(ntnodes_);
    std::vector<std::vector<std::vector<int>>> tclu(MAXDEG + 1), tseg(MAXDEG + 1), tsegc(MAXDEG + 1);
    double dsum = 0.0;
    var_evals_ = 0;
    for (long long i = 0; i < np; i++) {
        const FarPair &q = pairs_[i];
        const int p = deg[i];
        const System &y = sys_[q.sys];
        DegreeWork &w = dw_[p];
        const long long P = w.P;
        dsum += p;
        if (q.kind == FAR_PC) {
            pc_at[y.tnode_off + q.t].push_back(multi_id(p, src_slot(q.sys, q.s, p)));
            w.evals += (long long)y.ttree.numpar[q.t] * P;
            continue;
        }
        const int sl = tgt_slot(q.sys, q.t, p);
        if ((int)tclu[p].size() <= sl) { tclu[p].resize(sl + 1); tseg[p].resize(sl + 1); tsegc[p].resize(sl + 1); }
        if (q.kind == FAR_CP) {
            tseg[p][sl].push_back(y.soff + y.stree.ibeg[q.s]);
            tsegc[p][sl].push_back(y.stree.numpar[q.s]);
            w.evals += P * y.stree.numpar[q.s];
        } else {
            tclu[p][sl].push_back(multi_id(p, src_slot(q.sys, q.s, p)));
            w.evals += P * P;
        }
    }
    var_mean_deg_ = np ? dsum / np : 0.0;
    for (int p = 1; p <= MAXDEG; p++) var_evals_ += dw_[p].evals;

    // near: the fixed plan's keys and particle ranges, with the source slots
    // of the key (PC) or of the key and its ancestors (CC)
    vnear_ = near_;
    vnear_.clu_id.clear();
    vnear_.clu_off.assign(1, 0);
    for (size_t k = 0; k < near_.key.size(); k++) {
        for (int a = near_.key[k]; a >= 0; a = prm_.compute_type == C_CC ? g_tparent_[a] : -1)
            for (int id : pc_at[a]) vnear_.clu_id.push_back(id);
        vnear_.clu_off.push_back((int)vnear_.clu_id.size());
    }
    make_queue_tiles(vnear_, tile_targets_, false, 0, vnear_tiles_, vnear_qlen_, true);

    // far: every target slot, degree by degree, split by sources; rows of
    // partial grid values, summed per slot after the launch
    WorkCSR raw;
    raw.seg_off.push_back(0); raw.clu_off.push_back(0);
    for (int p = 1; p <= MAXDEG; p++) {
        const int T = (int)slot_node[p].size();
        for (int sl = 0; sl < T; sl++) {
            raw.key.push_back(multi_id(p, sl));
            raw.tbeg.push_back(0);
            raw.tcnt.push_back(dw_[p].P);
            for (size_t j = 0; j < tseg[p][sl].size(); j++) {
                raw.seg_beg.push_back(tseg[p][sl][j]);
                raw.seg_cnt.push_back(tsegc[p][sl][j]);
                raw.seg_chk.push_back(0);
            }
            for (int id : tclu[p][sl]) raw.clu_id.push_back(id);
            raw.seg_off.push_back((int)raw.seg_beg.size());
            raw.clu_off.push_back((int)raw.clu_id.size());
        }
    }
    split_keys(raw, MAX_KEY_SOURCES, 0, true, vfar_);
    vfar_row_.assign(vfar_.key.size(), 0);
    vfar_rows_len_ = 0;
    for (size_t k = 0; k < vfar_.key.size(); k++) {
        vfar_row_[k] = vfar_rows_len_;
        vfar_rows_len_ += vfar_.tcnt[k];
    }
    for (int p = 1; p <= MAXDEG; p++) dw_[p].rows_off.assign(slot_node[p].size() + 1, 0);
    {
        // keys come degree by degree, slot by slot
        for (size_t k = 0; k < vfar_.key.size(); k++) {
            const int p = vfar_.key[k] >> btk::DEG_SHIFT, sl = vfar_.key[k] & btk::SLOT_MASK;
            dw_[p].rows_off[sl + 1]++;
        }
        int base = 0;
        for (int p = 1; p <= MAXDEG; p++) {
            std::vector<int> &o = dw_[p].rows_off;
            o[0] = base;
            for (size_t sl = 1; sl < o.size(); sl++) o[sl] += o[sl - 1];
            base = o.back();
        }
    }
    make_queue_tiles(vfar_, GRID_TILE, true, 0, vfar_tiles_, vfar_qlen_, true);
    make_pencil_tiles(vfar_, 0, true, vpencil_tiles_);

    // degrees up to the plan's own go through its degree-D grids (exact in
    // exact arithmetic; see bt_degree.hip)
    const int D = prm_.degree;
    {
        std::vector<std::vector<int>> items(ntnodes_);
        for (int p = 1; p <= std::min(D, MAXDEG); p++) {
            if (!used[p]) continue;
            if (!dw_[p].s_ibeg.empty()) fold_up_ = true;
            for (size_t sl = 0; sl < slot_node[p].size(); sl++) {
                items[sys_[slot_sys[p][sl]].tnode_off + slot_node[p][sl]].push_back(multi_id(p, (int)sl));
                fold_down_ = true;
            }
        }
        fold_off_.assign(1, 0);
        for (int n = 0; n < ntnodes_; n++) {
            if (items[n].empty()) continue;
            fold_node_.push_back(n);
            fold_item_.insert(fold_item_.end(), items[n].begin(), items[n].end());
            fold_off_.push_back((int)fold_item_.size());
        }
        // E_p[k][m] = l^p_k(t^D_m), t^n_i = cos(i pi / n) (the node tables'
        // reference points), by the product formula in long double
        emat_off_.assign(D + 1, 0);
        for (int p = 1; p <= D; p++) {
            emat_off_[p] = emat_.size();
            for (int k = 0; k <= p; k++)
                for (int m = 0; m <= D; m++) {
                    const long double t = cosl((long double)m * M_PI / D);
                    const long double tk = cosl((long double)k * M_PI / p);
                    long double v = 1.0L;
                    for (int j = 0; j <= p; j++) {
                        if (j == k) continue;
                        const long double tj = cosl((long double)j * M_PI / p);
                        v *= (t - tj) / (tk - tj);
                    }
                    emat_.push_back((double)v);
                }
        }
    }
    for (int p = 1; p <= MAXDEG; p++) {
        if (!used[p]) continue;
        DegreeWork &w = dw_[p];
        // downpass: each target slot's interpolant at its node's particles,
        // one launch per tree level (nodes of a level have disjoint ranges);
        // degrees up to D are folded instead
        const int T = p <= D ? 0 : (int)slot_node[p].size();
        w.down.assign(tdepth_ + 1, {});
        for (int sl = 0; sl < T; sl++) {
            const System &y = sys_[slot_sys[p][sl]];
            const int t = slot_node[p][sl];
            const int lv = y.ttree.level[t];
            for (int o = 0; o < y.ttree.numpar[t]; o += DOWN_TILE) {
                btk::Tile tl;
                tl.x = sl;
                tl.y = y.toff + y.ttree.ibeg[t] + o;
                tl.z = std::min(DOWN_TILE, y.ttree.numpar[t] - o);
                tl.w = 0;
                w.down[lv].push_back(tl);
            }
        }
        // parts of the source slots (degrees above D)
        w.p_slot.clear(); w.p_ibeg.clear(); w.p_iend.clear(); w.p_off.assign(1, 0);
        for (size_t sl = 0; sl < w.s_ibeg.size(); sl++) {
            for (int b = w.s_ibeg[sl]; p > D && b < w.s_iend[sl]; b += PART_MAX) {
                w.p_slot.push_back((int)sl);
                w.p_ibeg.push_back(b);
                w.p_iend.push_back(std::min(b + PART_MAX, w.s_iend[sl]));
            }
            w.p_off.push_back((int)w.p_slot.size());
        }
    }
    std::string e = validate_var();
    if (!e.empty()) { free_var(); return "internal variable-degree work check failed: " + e; }
    var_planned_ = true;
    if (!st_) return "";  // host-only plan: work built and checked, nothing to upload

    // upload
    alloc_var_ = true;
    std::string r;
    auto fail = [&](const std::string &m) { alloc_var_ = false; free_var(); return m; };
    auto up_csr = [&](const WorkCSR &wk, int **so, btk::Seg **sg, int **co, int **ci) -> std::string {
        std::vector<btk::Seg> sv(wk.seg_beg.size());
        for (size_t i = 0; i < sv.size(); i++) { sv[i].x = wk.seg_beg[i]; sv[i].y = wk.seg_cnt[i]; }
        std::string rr;
        if (!(rr = up(so, wk.seg_off)).empty()) return rr;
        if (!(rr = up(sg, sv)).empty()) return rr;
        if (!(rr = up(co, wk.clu_off)).empty()) return rr;
        return up(ci, wk.clu_id);
    };
    if (!(r = up(&d_vn_clu_off_, vnear_.clu_off)).empty()) return fail(r);
    if (!(r = up(&d_vn_clu_id_, vnear_.clu_id)).empty()) return fail(r);
    if (!(r = up(&d_vn_tiles_, vnear_tiles_)).empty()) return fail(r);
    if (!(r = up_csr(vfar_, &d_vf_seg_off_, &d_vf_segs_, &d_vf_clu_off_, &d_vf_clu_id_)).empty()) return fail(r);
    if (!(r = up(&d_vf_tiles_, vfar_tiles_)).empty()) return fail(r);
    if (!vpencil_tiles_.empty() && !(r = up(&d_vp_tiles_, vpencil_tiles_)).empty()) return fail(r);
    if (!(r = up(&d_vf_row_, vfar_row_)).empty()) return fail(r);
    d_vf_R_ = dalloc<double>(vfar_rows_len_);
    if (ss) d_vf_RW_ = dalloc<double>(vfar_rows_len_);
    if (!d_vf_R_ || (ss && !d_vf_RW_)) return fail("hipMalloc failed (far rows)");
    for (int p = 1; p <= MAXDEG; p++) {
        if (!used[p]) continue;
        DegreeWork &w = dw_[p];
        const size_t S = w.s_ibeg.size(), T = w.num_tslots();
        if (S > 0) {
            if (!(r = up(&w.d_p_slot, w.p_slot)).empty()) return fail(r);
            if (!(r = up(&w.d_p_ibeg, w.p_ibeg)).empty()) return fail(r);
            if (!(r = up(&w.d_p_iend, w.p_iend)).empty()) return fail(r);
            if (!(r = up(&w.d_p_off, w.p_off)).empty()) return fail(r);
            if (!(r = up(&w.d_s_tab, w.s_tab)).empty()) return fail(r);
            const size_t n = S * w.P, npt = w.p_slot.size() * (size_t)w.P;
            w.d_cx = dalloc<double>(n); w.d_cy = dalloc<double>(n); w.d_cz = dalloc<double>(n);
            w.d_cq = dalloc<double>(n);
            w.d_Qp = dalloc<double>(npt);
            if (ss) { w.d_cw = dalloc<double>(n); w.d_Wp = dalloc<double>(npt); }
            if (!w.d_cx || !w.d_cy || !w.d_cz || !w.d_cq || !w.d_Qp || (ss && (!w.d_cw || !w.d_Wp)))
                return fail("hipMalloc failed (slots)");
        }
        if (T > 0) {
            if (!(r = up(&w.d_t_tab, w.t_tab)).empty()) return fail(r);
            if (!(r = up(&w.d_rows_off, w.rows_off)).empty()) return fail(r);
            w.d_Qt = dalloc<double>(T * w.P);
            if (ss) w.d_Wt = dalloc<double>(T * w.P);
            if (!w.d_Qt || (ss && !w.d_Wt)) return fail("hipMalloc failed (slots)");
        }
        w.d_down.assign(w.down.size(), nullptr);
        for (size_t lv = 0; lv < w.down.size(); lv++)
            if (!w.down[lv].empty() && !(r = up(&w.d_down[lv], w.down[lv])).empty()) return fail(r);
        if (S > 0 && p <= D && !(r = up(&w.d_s_node, w.s_node)).empty()) return fail(r);
    }
    if (fold_up_ || fold_down_) {
        if (!(r = up(&d_emat_, emat_)).empty()) return fail(r);
    }
    if (fold_down_) {
        if (!(r = up(&d_fold_node_, fold_node_)).empty()) return fail(r);
        if (!(r = up(&d_fold_off_, fold_off_)).empty()) return fail(r);
        if (!(r = up(&d_fold_item_, fold_item_)).empty()) return fail(r);
    }
    alloc_var_ = false;
    HIPCHK(hipStreamSynchronize(st_));
    var_ = true;
    return "";
}

std::string Plan::validate_var() const
{
    auto slot_ok = [&](int id, bool target) {
        const int p = id >> btk::DEG_SHIFT, sl = id & btk::SLOT_MASK;
        if (p < 1 || p > MAXDEG || (int)dw_.size() <= p) return false;
        const DegreeWork &w = dw_[p];
        return sl >= 0 && sl < (target ? w.num_tslots() : (int)w.s_ibeg.size());
    };
    auto chk = [&](const WorkCSR &w, const std::vector<btk::Tile> &tiles, const int *qlen, bool grid,
                   int tile_t) -> std::string {
        const int nk = (int)w.key.size();
        if ((int)w.seg_off.size() != nk + 1 || (int)w.clu_off.size() != nk + 1) return "csr offsets";
        if (w.seg_off[nk] != (int)w.seg_beg.size() || w.clu_off[nk] != (int)w.clu_id.size()) return "csr ends";
        for (int k = 0; k < nk; k++) {
            if (w.seg_off[k] > w.seg_off[k + 1] || w.clu_off[k] > w.clu_off[k + 1]) return "csr order";
            if (grid && !slot_ok(w.key[k], true)) return "grid key";
            if (grid && w.tcnt[k] != dw_[w.key[k] >> btk::DEG_SHIFT].P) return "grid key size";
        }
        for (size_t i = 0; i < w.seg_beg.size(); i++)
            if (w.seg_beg[i] < 0 || w.seg_cnt[i] < 0 || (long long)w.seg_beg[i] + w.seg_cnt[i] > N_)
                return "source range";
        for (int c : w.clu_id)
            if (!slot_ok(c, false)) return "source slot";
        for (const btk::Tile &t : tiles) {
            if (t.z == 0) continue;
            if (t.x < 0 || t.x >= nk || t.z < 0 || t.z > tile_t || t.y < 0) return "tile";
            if (grid) {
                if (t.w != w.key[t.x] || t.y + t.z > w.tcnt[t.x]) return "grid tile";
            } else if ((long long)t.y + t.z > M_) {
                return "target tile";
            }
        }
        for (int x = 0; x < NQ; x++)
            for (int j = 0; j < qlen[x]; j++) {
                const size_t t = (size_t)j * NQ + x;
                if (t >= tiles.size() || tiles[t].z <= 0) return "tile queues";
            }
        return "";
    };
    std::string e = chk(vnear_, vnear_tiles_, vnear_qlen_, false, tile_targets_);
    if (!e.empty()) return "near " + e;
    if ((e = chk(vfar_, vfar_tiles_, vfar_qlen_, true, GRID_TILE)).empty() == false) return "far " + e;
    if (vpencil_tiles_.size() != vfar_.key.size()) return "pencil tiles";
    for (const btk::Tile &t : vpencil_tiles_) {
        const int d = t.w >> btk::DEG_SHIFT;
        if (t.x < 0 || t.x >= (int)vfar_.key.size() || t.y != 0 || t.w != vfar_.key[t.x] || d < 1 ||
            d > btk::MAX_DEGREE || t.z != (d + 1) * (d + 1) * (d + 1))
            return "pencil tile";
    }
    if (vfar_row_.size() != vfar_.key.size()) return "far rows";
    for (size_t k = 0; k < vfar_.key.size(); k++)
        if (vfar_row_[k] < 0 || vfar_row_[k] + vfar_.tcnt[k] > vfar_rows_len_) return "far row";
    for (const DegreeWork &w : dw_) {
        if (!w.used) continue;
        const int S = (int)w.s_ibeg.size(), T = w.num_tslots();
        if ((int)w.s_iend.size() != S || (int)w.s_tab.size() != S * 3 * w.nl) return "source slots";
        for (int i = 0; i < S; i++)
            if (w.s_ibeg[i] < 0 || w.s_iend[i] > N_ || w.s_ibeg[i] >= w.s_iend[i]) return "source slot range";
        const int NP = (int)w.p_slot.size();
        if ((int)w.p_ibeg.size() != NP || (int)w.p_iend.size() != NP || (int)w.p_off.size() != S + 1 ||
            w.p_off[0] != 0 || w.p_off[S] != NP)
            return "source parts";
        for (int k = 0; k < NP; k++)
            if (w.p_slot[k] < 0 || w.p_slot[k] >= S || w.p_ibeg[k] < w.s_ibeg[w.p_slot[k]] ||
                w.p_iend[k] > w.s_iend[w.p_slot[k]] || w.p_ibeg[k] >= w.p_iend[k])
                return "source part range";
        for (int i = 0; i < S; i++)
            if (w.p_off[i] > w.p_off[i + 1]) return "source part offsets";
        if ((int)w.rows_off.size() != T + 1) return "slot rows";
        for (int sl = 0; sl < T; sl++) {
            if (w.rows_off[sl] > w.rows_off[sl + 1] || w.rows_off[sl + 1] > (int)vfar_.key.size())
                return "slot rows order";
            for (int k = w.rows_off[sl]; k < w.rows_off[sl + 1]; k++)
                if (vfar_.key[k] != multi_id(w.degree, sl)) return "slot row key";
        }
        for (const auto &lv : w.down)
            for (const btk::Tile &t : lv)
                if (t.x < 0 || t.x >= T || t.y < 0 || t.z <= 0 || t.z > DOWN_TILE || (long long)t.y + t.z > M_)
                    return "down tile";
        if ((int)w.s_node.size() != S) return "source slot nodes";
        for (int n : w.s_node)
            if (n < 0 || n >= nsnodes_) return "source slot node";
    }
    const int D = prm_.degree;
    if (fold_off_.empty() || fold_off_.size() != fold_node_.size() + 1 || fold_off_[0] != 0 ||
        fold_off_.back() != (int)fold_item_.size())
        return "fold lists";
    for (size_t b = 0; b < fold_node_.size(); b++) {
        if (fold_node_[b] < 0 || fold_node_[b] >= ntnodes_ || fold_off_[b] > fold_off_[b + 1]) return "fold node";
        for (int k = fold_off_[b]; k < fold_off_[b + 1]; k++)
            if ((fold_item_[k] >> btk::DEG_SHIFT) > D || !slot_ok(fold_item_[k], true)) return "fold item";
    }
    if ((fold_up_ || fold_down_) && (D < 1 || D > btk::MAX_DEGREE || (int)emat_off_.size() != D + 1)) return "fold degree";
    for (int p = 1; p <= D && p < (int)emat_off_.size(); p++)
        if (emat_off_[p] + (size_t)(p + 1) * (D + 1) > emat_.size()) return "fold matrices";
    if (fold_up_ && prm_.compute_type == C_CP) return "fold of source slots in a CP plan";
    if (fold_down_ && prm_.compute_type == C_PC) return "fold of target slots in a PC plan";
    return "";
}

std::string Plan::label_pairs(int first, int count, double tol, int *labels, double *err)
{
    if (!st_) return "plan has no device data";
    if (!coulomb_family(prm_.kernel))
        return "pair labels are set up for the Coulomb, Yukawa and regularized Coulomb kernels";
    if (first < 0 || count < 0 || (long long)first + count > (long long)pairs_.size())
        return "pair range out of bounds";
    if (count == 0) return "";
    std::vector<int> kind(count), tb(count), tc(count), sb(count), sc(count);
    std::vector<double> tbox(6 * (size_t)count), sbox(6 * (size_t)count);
    for (int i = 0; i < count; i++) {
        const FarPair &q = pairs_[first + i];
        const System &y = sys_[q.sys];
        const Tree &tt = y.ttree, &st = y.stree;
        kind[i] = q.kind;
        tb[i] = y.toff + tt.ibeg[q.t]; tc[i] = tt.numpar[q.t];
        sb[i] = y.soff + st.ibeg[q.s]; sc[i] = st.numpar[q.s];
        double *a = &tbox[6 * (size_t)i], *b = &sbox[6 * (size_t)i];
        a[0] = tt.xmin[q.t]; a[1] = tt.xmax[q.t]; a[2] = tt.ymin[q.t];
        a[3] = tt.ymax[q.t]; a[4] = tt.zmin[q.t]; a[5] = tt.zmax[q.t];
        b[0] = st.xmin[q.s]; b[1] = st.xmax[q.s]; b[2] = st.ymin[q.s];
        b[3] = st.ymax[q.s]; b[4] = st.zmin[q.s]; b[5] = st.zmax[q.s];
    }
    // temporary device buffers, freed before returning
    std::vector<void *> tmp;
    auto cleanup = [&]() { for (void *p : tmp) (void)hipFree(p); };
    auto put = [&](const void *h, size_t bytes) -> void * {
        void *d = nullptr;
        if (hipMalloc(&d, std::max<size_t>(bytes, 8)) != hipSuccess) return nullptr;
        tmp.push_back(d);
        if (hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, st_) != hipSuccess) return nullptr;
        return d;
    };
    btl::LabelArgs a{};
    a.npairs = count;
    a.kind = (const int *)put(kind.data(), count * sizeof(int));
    a.tbeg = (const int *)put(tb.data(), count * sizeof(int));
    a.tcnt = (const int *)put(tc.data(), count * sizeof(int));
    a.sbeg = (const int *)put(sb.data(), count * sizeof(int));
    a.scnt = (const int *)put(sc.data(), count * sizeof(int));
    a.tbox = (const double *)put(tbox.data(), tbox.size() * sizeof(double));
    a.sbox = (const double *)put(sbox.data(), sbox.size() * sizeof(double));
    void *dl = nullptr, *de = nullptr;
    if (hipMalloc(&dl, count * sizeof(int)) == hipSuccess) tmp.push_back(dl); else dl = nullptr;
    if (hipMalloc(&de, (size_t)count * btl::LABEL_ND * sizeof(double)) == hipSuccess) tmp.push_back(de); else de = nullptr;
    if (!a.kind || !a.tbeg || !a.tcnt || !a.sbeg || !a.scnt || !a.tbox || !a.sbox || !dl || !de) {
        cleanup();
        return "device buffers for labels failed";
    }
    a.tx = d_tx_; a.ty = d_ty_; a.tz = d_tz_;
    a.sx = d_sx_; a.sy = d_sy_; a.sz = d_sz_; a.sq = d_sq_;
    a.kernel = prm_.kernel;
    a.p0 = prm_.kernel == K_YUKAWA ? prm_.kp[0]
         : (prm_.kernel == K_REG_COULOMB ? (prm_.singularity == S_SUBTRACTION ? prm_.kp[1] : prm_.kp[0]) : 0.0);
    // subtraction plans: labels of the subtraction kernel (both terms)
    if (prm_.singularity == S_SUBTRACTION) {
        a.ss = 1;
        a.tq = d_tq_; a.sw = d_sw_;
        a.p1 = prm_.kernel == K_YUKAWA ? 0.0 : 1.0 / (prm_.kp[0] * prm_.kp[0]);
    }
    a.tol = tol;
    a.labels = (int *)dl;
    a.errors = (double *)de;
    hipError_t e = btl::launch_label(a, st_);
    if (e == hipSuccess) e = hipMemcpyAsync(labels, dl, count * sizeof(int), hipMemcpyDeviceToHost, st_);
    std::vector<double> eh;
    if (e == hipSuccess && err)
        e = hipMemcpyAsync(err, de, (size_t)count * btl::LABEL_ND * sizeof(double), hipMemcpyDeviceToHost, st_);
    if (e == hipSuccess) e = hipStreamSynchronize(st_);
    cleanup();
    if (e != hipSuccess) return std::string("label kernel: ") + hipGetErrorString(e);
    return "";
}

std::string Plan::execute_var(bool timed)
{
    const Params &p = prm_;
    const bool ss = p.singularity == S_SUBTRACTION;
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

    // 1. source slots: modified charges of degree p -- up to the plan's
    //    degree from the node's own degree-D charges, above it from the
    //    node's particles
    const int D = p.degree;
    if (fold_up_) {
        std::string r = upward_fixed();
        if (!r.empty()) return r;
    }
    for (const DegreeWork &w : dw_) {
        const int S = (int)w.s_ibeg.size();
        if (!w.used || S == 0) continue;
        HIPCHK(btk::launch_cluster_points(S, w.nl, w.d_s_tab, w.d_cx, w.d_cy, w.d_cz, st_));
        if (w.degree <= D) {
            btk::DegreeUpArgs u{};
            u.nslots = S; u.p = w.degree; u.D = D; u.subtraction = ss ? 1 : 0;
            u.node = w.d_s_node; u.E = d_emat_ + emat_off_[w.degree];
            u.QD = d_cq_; u.WD = d_cw_; u.Qp = w.d_cq; u.Wp = w.d_cw;
            HIPCHK(btk::launch_degree_up(u, st_));
            continue;
        }
        btk::AnterpArgs a{};
        a.degree = w.degree; a.subtraction = ss;
        a.ibeg = w.d_p_ibeg; a.iend = w.d_p_iend; a.tab = w.d_s_tab;
        a.px = d_sx_; a.py = d_sy_; a.pz = d_sz_; a.pq = d_sq_; a.pw = d_sw_;
        a.Q = w.d_Qp; a.W = w.d_Wp;
        a.mode = 2; a.nodes = w.d_p_slot; a.nblocks = (int)w.p_slot.size();
        HIPCHK(btk::launch_anterp(a, st_));
        HIPCHK(btk::launch_sum_parts(S, w.P, w.d_p_off, w.d_Qp, w.d_Wp, w.d_cq, ss ? w.d_cw : nullptr, st_));
    }
    HIPCHK(hipEventRecord(ev[1], st_));

    btk::InteractArgs ia{};
    ia.kernel = p.kernel; ia.singularity = p.singularity; ia.precision = p.precision;
    ia.tpt = tpt_;
    ia.persistent = persistent_ ? 1 : 0;
    ia.kp = kp_;
    ia.tx = d_tx_; ia.ty = d_ty_; ia.tz = d_tz_; ia.tq = d_tq_;
    ia.sx = d_sx_; ia.sy = d_sy_; ia.sz = d_sz_; ia.sq = d_sq_; ia.sw = d_sw_;
    ia.multi = 1;
    for (const DegreeWork &w : dw_) {
        if (!w.used) continue;
        ia.dcx[w.degree] = w.d_cx; ia.dcy[w.degree] = w.d_cy; ia.dcz[w.degree] = w.d_cz;
        ia.dcq[w.degree] = w.d_cq; ia.dcw[w.degree] = w.d_cw; ia.dtab[w.degree] = w.d_t_tab;
    }

    // 2. near field and the source slots of every degree into the potential
    //    (NaN first: every target is written by exactly one key)
    HIPCHK(hipMemsetAsync(d_pot_, 0xFF, (size_t)M_ * sizeof(double), st_));
    {
        btk::InteractArgs a = ia;
        a.grid_targets = 0; a.ntiles = (int)vnear_tiles_.size(); a.tiles = d_vn_tiles_;
        a.seg_off = d_nseg_off_; a.segs = d_nsegs_; a.clu_off = d_vn_clu_off_; a.clu_id = d_vn_clu_id_;
        a.out1 = d_pot_;
        if (p.precision == PREC_MIXED) a.key_ctr = reinterpret_cast<const double4 *>(d_near_ctr_);
        a.skip_sym = sym_ready() ? 1 : 0;
        std::copy(vnear_qlen_, vnear_qlen_ + NQ, a.qlen);
        a.cursor = d_cursor_;
        HIPCHK(btk::launch_interact(a, st_));
        if (a.skip_sym) {
            std::string r = near_sym(!ov);
            if (!r.empty()) return r;
        }
    }
    HIPCHK(hipEventRecord(ev[2], st_));

    // 3. every target slot's grid values (rows), then the sum per slot
    if (!vfar_.key.empty()) {
        btk::InteractArgs a = ia;
        // acceptance-separated sources: the centred expansion of r^2 (mixed)
        a.grid_scaled = p.precision == PREC_MIXED && p.theta < 0.99;
        a.grid_targets = 1; a.ntiles = (int)vfar_tiles_.size(); a.tiles = d_vf_tiles_;
        a.seg_off = d_vf_seg_off_; a.segs = d_vf_segs_; a.clu_off = d_vf_clu_off_; a.clu_id = d_vf_clu_id_;
        a.out1 = d_vf_R_; a.out2 = d_vf_RW_; a.out_row = d_vf_row_;
        std::copy(vfar_qlen_, vfar_qlen_ + NQ, a.qlen);
        a.cursor = d_cursor_ + NQ;
        if (pencil_grids()) {
            a.ntiles = (int)vpencil_tiles_.size(); a.tiles = d_vp_tiles_;
            HIPCHK(btk::launch_grid_pencil(a, st_));
        } else {
            HIPCHK(btk::launch_interact(a, st_));
        }
        for (const DegreeWork &w : dw_) {
            const int T = w.num_tslots();
            if (!w.used || T == 0) continue;
            HIPCHK(btk::launch_sum_rows(T, w.P, w.d_rows_off, d_vf_row_, d_vf_R_, d_vf_RW_, w.d_Qt,
                                        ss ? w.d_Wt : nullptr, st_));
        }
    }
    HIPCHK(hipEventRecord(ev[3], st_));

    // 4. target slots up to the plan's degree onto their nodes' degree-D
    //    grids and through the plan's downpass; above it, each slot's
    //    interpolant at its node's particles
    if (fold_down_) {
        const size_t nq = (size_t)ntnodes_ * p.P;
        HIPCHK(hipMemsetAsync(d_Qt_, 0, nq * sizeof(double), st_));
        if (ss) HIPCHK(hipMemsetAsync(d_Wt_, 0, nq * sizeof(double), st_));
        btk::DegreeDownArgs f{};
        f.nnodes = (int)fold_node_.size(); f.D = D; f.subtraction = ss ? 1 : 0;
        f.node = d_fold_node_; f.off = d_fold_off_; f.item = d_fold_item_;
        for (const DegreeWork &w : dw_) {
            if (!w.used || w.degree > D) continue;
            f.E[w.degree] = d_emat_ + emat_off_[w.degree];
            f.Qs[w.degree] = w.d_Qt; f.Ws[w.degree] = w.d_Wt;
        }
        f.QD = d_Qt_; f.WD = d_Wt_;
        HIPCHK(btk::launch_degree_down(f, st_));
        std::string r = down_fixed();
        if (!r.empty()) return r;
    }
    for (const DegreeWork &w : dw_) {
        if (!w.used) continue;
        for (size_t lv = 0; lv < w.down.size(); lv++) {
            if (w.down[lv].empty()) continue;
            btk::DownArgs d{};
            d.degree = w.degree; d.subtraction = ss; d.tab = w.d_t_tab; d.parent = nullptr;
            d.tx = d_tx_; d.ty = d_ty_; d.tz = d_tz_; d.tq = d_tq_; d.Q = w.d_Qt; d.W = w.d_Wt;
            d.pot = d_pot_;
            d.mode = 1; d.tiles = w.d_down[lv]; d.nblocks = (int)w.down[lv].size();
            HIPCHK(btk::launch_down(d, st_));
        }
    }
    HIPCHK(hipEventRecord(ev[4], st_));
    if (!(er = join_phases(ov)).empty()) return er;

    // 5. subtraction correction and the caller's order
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

}  // namespace bt
