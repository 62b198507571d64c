// bt_hermite_plan.cpp -- the Hermite approximation in a plan: work lists,
// upload and the GPU phases that differ from the Lagrange path.
//
// Reference order (src/drivers/treedriver.c with approximation HERMITE):
//   PC: every source node's 8-plane charges straight from its particles
//       (clusters.c:182-190), then per target batch the PP and the Hermite
//       PC kernels (interaction_compute_pc.c:121-140 and the *_pc.c Hermite
//       kernels);
//   CP: per source batch the Hermite CP kernel into the target nodes' 8
//       planes (interaction_compute_cp.c:121-140), then every target node's
//       planes to its own particles (interaction_compute_downpass.c:117-120);
//       here each target sums the interpolants of its leaf and ancestors.
// Here the near field keeps its own launch; the Hermite PC clusters of each
// target batch run in a second launch that adds to the potential.
#include <algorithm>

#include "bt_solver.h"

namespace bt {

namespace {
#define HIPCHK(expr)                                                                 \
    do {                                                                             \
        hipError_t e_ = (expr);                                                      \
        if (e_ != hipSuccess) return std::string(#expr ": ") + hipGetErrorString(e_); \
    } while (0)

constexpr int HERM_TILE = 256;  // targets (or grid points) per Hermite tile
constexpr int HERM_PLANES = 8;
// sources per Hermite CP key: heavier keys are split (into rows of their
// own), so that the launch has many blocks of similar work
constexpr long long HERM_KEY_SOURCES = 1 << 14;
}  // namespace

void Plan::build_hermite_work()
{
    const Params &p = prm_;
    const int P = p.P;
    hp_node_.clear(); hp_beg_.clear(); hp_end_.clear(); hp_off_.assign(1, 0);
    hpc_ = WorkCSR();
    hpc_tiles_.clear(); hcp_tiles_.clear(); hdown_tiles_.clear();
    swtab_.clear(); twtab_.clear();
    std::vector<double> tab;

    if (p.compute_type == C_PC) {
        // upward pass: every source node from its own particles, in parts
        for (const System &sy : sys_) {
            for (int i = 0; i < sy.stree.n; i++) {
                const int b = sy.soff + sy.stree.ibeg[i], e = sy.soff + sy.stree.iend[i];
                for (int o = b; o < e; o += PART_MAX) {
                    hp_node_.push_back(sy.snode_off + i);
                    hp_beg_.push_back(o);
                    hp_end_.push_back(std::min(e, o + PART_MAX));
                }
                hp_off_.push_back((int)hp_node_.size());
            }
            hermite_weight_table(sy.stree, p.degree, tab);
            swtab_.insert(swtab_.end(), tab.begin(), tab.end());
        }
        // the near keys' clusters move to the Hermite PC launch
        WorkCSR nn;
        nn.seg_off.push_back(0); nn.clu_off.push_back(0);
        hpc_.seg_off.push_back(0); hpc_.clu_off.push_back(0);
        for (size_t k = 0; k < near_.key.size(); k++) {
            nn.key.push_back(near_.key[k]); nn.tbeg.push_back(near_.tbeg[k]); nn.tcnt.push_back(near_.tcnt[k]);
            for (int s = near_.seg_off[k]; s < near_.seg_off[k + 1]; s++) {
                nn.seg_beg.push_back(near_.seg_beg[s]);
                nn.seg_cnt.push_back(near_.seg_cnt[s]);
                nn.seg_chk.push_back(near_.seg_chk[s]);
                nn.pairs += (long long)near_.tcnt[k] * near_.seg_cnt[s];
            }
            nn.seg_off.push_back((int)nn.seg_beg.size());
            nn.clu_off.push_back(0);
            const int c0 = near_.clu_off[k], c1 = near_.clu_off[k + 1];
            if (c1 == c0) continue;
            const int hk = (int)hpc_.key.size();
            hpc_.key.push_back(near_.key[k]); hpc_.tbeg.push_back(near_.tbeg[k]); hpc_.tcnt.push_back(near_.tcnt[k]);
            hpc_.clu_id.insert(hpc_.clu_id.end(), near_.clu_id.begin() + c0, near_.clu_id.begin() + c1);
            hpc_.clu_off.push_back((int)hpc_.clu_id.size());
            hpc_.seg_off.push_back(0);
            hpc_.pairs += (long long)near_.tcnt[k] * P * (c1 - c0);
            for (int o = 0; o < near_.tcnt[k]; o += HERM_TILE)
                hpc_tiles_.push_back(btk::Tile{hk, near_.tbeg[k] + o, std::min(HERM_TILE, near_.tcnt[k] - o), 0});
        }
        near_ = std::move(nn);
        make_queue_tiles(near_, tile_targets_, false, P, near_tiles_, near_qlen_);
    } else if (p.compute_type == C_CP) {
        // grid keys split finer than the Lagrange ones; chunks of a key keep
        // its node and follow it, so each node's keys stay consecutive
        {
            WorkCSR raw = std::move(grid_);
            split_keys(raw, HERM_KEY_SOURCES, P, false, grid_);
        }
        node_rows_off_.assign(ntnodes_ + 1, 0);
        for (int node : grid_.key) node_rows_off_[node + 1]++;
        for (int n = 0; n < ntnodes_; n++) node_rows_off_[n + 1] += node_rows_off_[n];
        make_queue_tiles(grid_, GRID_TILE, true, P, grid_tiles_, grid_qlen_);
        make_pencil_tiles(grid_, P, false, pencil_tiles_);
        // grid keys write rows of 8 planes
        grid_row_.assign(grid_.key.size(), 0);
        grid_rows_len_ = 0;
        for (size_t k = 0; k < grid_.key.size(); k++) {
            grid_row_[k] = grid_rows_len_;
            grid_rows_len_ += (long long)HERM_PLANES * grid_.tcnt[k];
        }
        // tiles of the heaviest keys first (the order of the rows' sums is
        // fixed by the rows, not by the tiles)
        std::vector<long long> load(grid_.key.size(), 0);
        for (size_t k = 0; k < grid_.key.size(); k++)
            for (int s = grid_.seg_off[k]; s < grid_.seg_off[k + 1]; s++) load[k] += grid_.seg_cnt[s];
        std::vector<int> order(grid_.key.size());
        for (size_t k = 0; k < order.size(); k++) order[k] = (int)k;
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return load[a] > load[b]; });
        for (int k : order)
            for (int o = 0; o < grid_.tcnt[k]; o += HERM_TILE)
                hcp_tiles_.push_back(btk::Tile{k, o, std::min(HERM_TILE, grid_.tcnt[k] - o), grid_.key[k]});
        // downpass: tiles of each target leaf; a target sums its leaf's and
        // its ancestors' interpolants
        for (const System &sy : sys_) {
            for (int leaf : sy.ttree.leaves) {
                const int b = sy.toff + sy.ttree.ibeg[leaf], m = sy.ttree.numpar[leaf];
                for (int o = 0; o < m; o += HERM_TILE)
                    hdown_tiles_.push_back(btk::Tile{sy.tnode_off + leaf, b + o, std::min(HERM_TILE, m - o), 0});
            }
            hermite_weight_table(sy.ttree, p.degree, tab);
            twtab_.insert(twtab_.end(), tab.begin(), tab.end());
        }
    }
}

std::string Plan::validate_hermite() const
{
    const long long P = prm_.P;
    if (prm_.compute_type == C_PC) {
        if ((int)hp_off_.size() != nsnodes_ + 1) return "hermite parts per node";
        for (size_t b = 0; b < hp_node_.size(); b++)
            if (hp_node_[b] < 0 || hp_node_[b] >= nsnodes_ || hp_beg_[b] < 0 || hp_end_[b] > N_ ||
                hp_beg_[b] >= hp_end_[b])
                return "hermite part";
        if (swtab_.size() != (size_t)nsnodes_ * 3 * (prm_.degree + 1)) return "hermite source weights";
        for (const btk::Tile &t : hpc_tiles_)
            if (t.x < 0 || t.x >= (int)hpc_.key.size() || t.y < 0 || t.z <= 0 || t.y + t.z > M_)
                return "hermite PC tile";
        for (int c : hpc_.clu_id)
            if (c < 0 || c >= nsnodes_) return "hermite PC cluster";
        for (int c : near_.clu_id)
            if (c >= 0) return "hermite near keys keep no clusters";
    } else if (prm_.compute_type == C_CP) {
        for (const btk::Tile &t : hcp_tiles_)
            if (t.x < 0 || t.x >= (int)grid_.key.size() || t.w < 0 || t.w >= ntnodes_ || t.y < 0 || t.z <= 0 ||
                t.y + t.z > P)
                return "hermite CP tile";
        for (size_t k = 0; k < grid_row_.size(); k++)
            if (grid_row_[k] < 0 || grid_row_[k] + HERM_PLANES * P > grid_rows_len_) return "hermite grid row";
        for (const btk::Tile &t : hdown_tiles_)
            if (t.x < 0 || t.x >= ntnodes_ || t.y < 0 || t.z <= 0 || t.y + t.z > M_) return "hermite down tile";
        // the leaf-to-root walk ends: parents precede their children
        if ((int)g_tparent_.size() != ntnodes_ || (int)node_rows_off_.size() != ntnodes_ + 1)
            return "hermite target parents";
        for (int n = 0; n < ntnodes_; n++)
            if (g_tparent_[n] < -1 || g_tparent_[n] >= n) return "hermite target parent order";
        if (twtab_.size() != (size_t)ntnodes_ * 3 * (prm_.degree + 1)) return "hermite target weights";
    }
    return "";
}

std::string Plan::upload_hermite()
{
    std::string r;
    const size_t P = prm_.P;
    const bool ss = prm_.singularity == S_SUBTRACTION;
    if (prm_.compute_type == C_PC) {
        if (!(r = up(&d_hp_node_, hp_node_)).empty()) return r;
        if (!(r = up(&d_hp_beg_, hp_beg_)).empty()) return r;
        if (!(r = up(&d_hp_end_, hp_end_)).empty()) return r;
        if (!(r = up(&d_hp_off_, hp_off_)).empty()) return r;
        if (!(r = up(&d_swtab_, swtab_)).empty()) return r;
        const size_t rows = std::max<size_t>(hp_node_.size(), 1) * HERM_PLANES * P;
        d_Qp_ = dalloc<double>(rows);
        if (ss) d_Wp_ = dalloc<double>(rows);
        if (!d_Qp_ || (ss && !d_Wp_)) return "hipMalloc failed (hermite parts)";
        if (!(r = up(&d_hpc_clu_off_, hpc_.clu_off)).empty()) return r;
        if (!(r = up(&d_hpc_clu_id_, hpc_.clu_id)).empty()) return r;
        if (!(r = up(&d_hpc_tiles_, hpc_tiles_)).empty()) return r;
    } else if (prm_.compute_type == C_CP) {
        if (!(r = up(&d_hcp_tiles_, hcp_tiles_)).empty()) return r;
        if (!(r = up(&d_hdown_tiles_, hdown_tiles_)).empty()) return r;
        if (!(r = up(&d_twtab_, twtab_)).empty()) return r;
    }
    return "";
}

// every source node's planes: parts, then the sum of each node's parts
std::string Plan::hermite_upward()
{
    const Params &p = prm_;
    const bool ss = p.singularity == S_SUBTRACTION;
    HIPCHK(btk::launch_cluster_points(nsnodes_, p.degree + 1, d_stab_, d_cx_, d_cy_, d_cz_, st_));
    btk::HermAnterpArgs a{};
    a.degree = p.degree; a.subtraction = ss; a.nparts = (int)hp_node_.size();
    a.part_node = d_hp_node_; a.part_beg = d_hp_beg_; a.part_end = d_hp_end_;
    a.tab = d_stab_; a.wtab = d_swtab_;
    a.px = d_sx_; a.py = d_sy_; a.pz = d_sz_; a.pq = d_sq_; a.pw = d_sw_;
    a.Qp = d_Qp_; a.Wp = d_Wp_;
    HIPCHK(btk::launch_herm_anterp(a, st_));
    HIPCHK(btk::launch_sum_parts(nsnodes_, HERM_PLANES * p.P, d_hp_off_, d_Qp_, d_Wp_, d_cq_, ss ? d_cw_ : nullptr,
                                 st_));
    return "";
}

std::string Plan::hermite_pc()
{
    btk::HermPCArgs a{};
    a.kernel = prm_.kernel; a.singularity = prm_.singularity; a.ntiles = (int)hpc_tiles_.size(); a.P = prm_.P;
    a.tiles = d_hpc_tiles_; a.clu_off = d_hpc_clu_off_; a.clu_id = d_hpc_clu_id_;
    a.cx = d_cx_; a.cy = d_cy_; a.cz = d_cz_; a.cq = d_cq_; a.cw = d_cw_;
    a.tx = d_tx_; a.ty = d_ty_; a.tz = d_tz_; a.tq = d_tq_;
    a.kp = kp_; a.pot = d_pot_;
    HIPCHK(btk::launch_herm_pc(a, st_));
    return "";
}

std::string Plan::hermite_cp()
{
    btk::HermCPArgs a{};
    a.kernel = prm_.kernel; a.ntiles = (int)hcp_tiles_.size(); a.nl = prm_.degree + 1;
    a.tiles = d_hcp_tiles_; a.seg_off = d_gseg_off_; a.segs = d_gsegs_; a.tab = d_ttab_;
    a.sx = d_sx_; a.sy = d_sy_; a.sz = d_sz_; a.sq = d_sq_;
    a.kp = kp_; a.out_row = d_grid_row_; a.rows = d_Qrows_;
    HIPCHK(btk::launch_herm_cp(a, st_));
    HIPCHK(btk::launch_sum_rows(ntnodes_, HERM_PLANES * prm_.P, d_node_rows_off_, d_grid_row_, d_Qrows_, nullptr,
                                d_Qt_, nullptr, st_));
    return "";
}

std::string Plan::hermite_down()
{
    btk::HermDownArgs a{};
    a.degree = prm_.degree; a.ntiles = (int)hdown_tiles_.size(); a.tiles = d_hdown_tiles_;
    a.parent = d_tparent_; a.rows_off = d_node_rows_off_;
    a.tab = d_ttab_; a.wtab = d_twtab_;
    a.tx = d_tx_; a.ty = d_ty_; a.tz = d_tz_; a.Q = d_Qt_; a.pot = d_pot_;
    HIPCHK(btk::launch_herm_down(a, st_));
    return "";
}

}  // namespace bt
