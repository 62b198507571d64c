// bt_solver.h -- plan/execute driver of the MI355X BaryTree path.
//
// Plan  : host work of treedriver (src/drivers/treedriver.c:114-967) -- private
//         particle copies, trees or batches, interaction lists, flattened
//         device work lists -- and the upload of everything to HBM.
// Execute: the interaction hot path on the GPU -- upward pass, near and far
//         field, downpass, subtraction correction, scatter to caller order.
//
// A plan holds one or more independent systems (batching).  Their particles
// are concatenated, their tree nodes numbered globally, and their work lists
// merged, so every phase is one launch for all systems.
#pragma once

#include <string>
#include <vector>

#include "bt_host.h"
#include "bt_kernels.h"

namespace bt {

constexpr int DOWN_TILE = 1024;  // targets per block of the leaf downpass
constexpr int GRID_TILE = 512;   // grid points per block of grid-target tiles (2 per thread)
constexpr int PART_MAX = 16384;  // particles per block of a source slot's charges
constexpr long long MAX_KEY_SOURCES = 1 << 14;  // source points per grid key (heavier keys are split:
                                                 // blocks enough to fill the GPU)
constexpr int SMALL_TARGETS = 1 << 18;  // single-system plans below this many targets: 1 target per thread

struct Timings {
    double host_copy = 0, host_tree = 0, host_lists = 0, host_work = 0, h2d = 0;  // seconds
    double gpu_upward = 0, gpu_near = 0, gpu_far = 0, gpu_down = 0, gpu_finish = 0, gpu_total = 0;
    double gpu_near_sym = 0;  // the symmetric near-field launch alone (part of gpu_near)
};

// One independent system of a plan.
struct System {
    int toff = 0, soff = 0, M = 0, N = 0;  // ranges in the concatenated arrays
    int msl = 1, mtl = 1;                  // leaf sizes used for this system
    Particles S, T;                        // private copies, tree order
    Tree stree, ttree;                     // source tree / batches, target tree / batches
    Lists L;
    int snode_off = 0, tnode_off = 0;      // global node id offsets
};

// An approximated (far-field) pair of the interaction lists: target node t
// and source node s of system sys (node ids local to the system's trees).
//   FAR_PC: source cluster -> target particles (PC lists, CC's PC entries)
//   FAR_CP: source particles -> target grid     (CP lists, CC's CP entries)
//   FAR_CC: source cluster -> target grid       (CC lists)
enum FarKind { FAR_PC = 1, FAR_CP = 2, FAR_CC = 3 };
struct FarPair { int sys, kind, t, s; };

// Features of a far-field pair, for the degree model.
constexpr int NUM_PAIR_FEATURES = 8;

// Data of one interpolation degree when every far-field pair has its own
// degree.  Source slots hold modified charges of (source node, degree),
// computed from the node's particles; target slots hold grid values of
// (target node, degree), interpolated straight to the node's particles.
struct DegreeWork {
    int degree = 0, nl = 0, P = 0;
    bool used = false;                     // some pair has this degree
    std::vector<int> s_ibeg, s_iend;       // global particle range per source slot
    std::vector<int> s_node;               // global source node per source slot
    // a source slot's charges are summed from parts of at most PART_MAX
    // particles (one block each): part k covers [p_ibeg[k], p_iend[k]) of
    // slot p_slot[k]; slot s owns parts [p_off[s], p_off[s+1])
    std::vector<int> p_slot, p_ibeg, p_iend, p_off;
    std::vector<double> s_tab, t_tab;      // 3*nl Chebyshev nodes per slot
    std::vector<int> rows_off;             // target slot s sums far keys [rows_off[s], rows_off[s+1])
    std::vector<std::vector<btk::Tile>> down;  // per tree level: {target slot, first target, count}
    long long evals = 0;                   // point-wise evaluations of this degree's pairs
    int num_tslots() const { return nl ? (int)(t_tab.size() / (3 * nl)) : 0; }
    // device copies
    int *d_p_slot = nullptr, *d_p_ibeg = nullptr, *d_p_iend = nullptr, *d_p_off = nullptr;
    int *d_s_node = nullptr;
    double *d_Qp = nullptr, *d_Wp = nullptr;
    double *d_s_tab = nullptr, *d_cx = nullptr, *d_cy = nullptr, *d_cz = nullptr;
    double *d_cq = nullptr, *d_cw = nullptr;
    double *d_t_tab = nullptr, *d_Qt = nullptr, *d_Wt = nullptr;
    int *d_rows_off = nullptr;
    std::vector<btk::Tile *> d_down;
};

// Deals a CSR's keys into tiles of at most tile_targets targets, as 8
// interleaved per-XCD queues (tile 8*j + x = j-th tile of queue x; padding
// slots have count 0); qlen gets each queue's length.
// With multi, cluster ids are (degree << DEG_SHIFT) | slot and weigh
// (degree+1)^3 points each; else P.
void make_queue_tiles(const WorkCSR &w, int tile_targets, bool grid, int P,
                      std::vector<btk::Tile> &tiles, int *qlen, bool multi = false);

// One tile {key, 0, points, node} per grid key for the tensor-grid launch
// (launch_grid_pencil: one block per grid), heaviest first.  With multi,
// cluster ids carry their degree as in make_queue_tiles.
void make_pencil_tiles(const WorkCSR &w, int P, bool multi, std::vector<btk::Tile> &tiles);

// Splits every key whose sources (particles plus cluster points) exceed
// max_src into consecutive chunks of its particle ranges and clusters; the
// chunks are keys of their own with the same node and targets, in order.
void split_keys(const WorkCSR &in, long long max_src, int P, bool multi, WorkCSR &out);

struct PlanInput {
    int nsys = 1;
    const int *toffs = nullptr, *soffs = nullptr;  // nsys+1 entries each
    const int *msl = nullptr, *mtl = nullptr;      // per-system leaf sizes (optional)
    int tpt = 2;                                   // targets per thread (tile = 256*tpt)
    int persistent = 0;                            // persistent interaction launches
    const double *tx = nullptr, *ty = nullptr, *tz = nullptr, *tq = nullptr;
    const double *sx = nullptr, *sy = nullptr, *sz = nullptr, *sq = nullptr, *sw = nullptr;
    bool device = false;  // the particle arrays are device pointers (read on the device only)
};

class Plan {
public:
    ~Plan();
    // Builds trees and lists and uploads; returns "" or an error message.
    std::string create(const Params &prm, const PlanInput &in, bool host_only = false);
    // Single-system convenience form.
    std::string create(const Params &prm, int M, int N, const double *tx, const double *ty,
                       const double *tz, const double *tq, const double *sx, const double *sy,
                       const double *sz, const double *sq, const double *sw,
                       bool host_only = false, bool device_inputs = false);
    // Runs the GPU path; result in caller order, on the device (out_dev()).
    std::string execute(bool timed);
    // Copies the result to `out` (host memory, or device memory if `device`).
    std::string download(double *out, bool device = false);
    // Frees the plan's device memory, events and stream (the destructor does
    // the same; a released plan cannot execute again).
    void release() { free_all(); }
    // Enqueues the next execution's near-field item launch now (when the
    // items run at the plan's current precision), so that host work before
    // execute() -- a degree model's -- overlaps it; execute() then goes on
    // from it.  Precision and symmetry must not change in between.
    std::string start_items();
    // Host work between start_items() and execute() that uploads to the
    // device (a degree model's per-degree work) goes to the stream of the
    // phases beside the item launch, so its copies do not wait for the
    // items: side_begin() before it, side_end() after (no-ops when the
    // items were not started or the second stream is not available).
    void side_begin();
    void side_end();
    // FP64 direct sum on the GPU for caller-order target indices `idx`
    // (each target against the sources of its own system).
    std::string direct(const std::vector<int> &idx, std::vector<double> &out);
    // The same sums without device allocation: every device buffer lies in
    // the scratch [scratch, scratch + bytes) (m <= NOALLOC_MAX_SAMPLE);
    // no host heap either.  Results equal direct()'s bit for bit.
    std::string direct_noalloc(int m, const int *idx, double *out, void *scratch, size_t bytes);

    void set_precision(int p) { prm_.precision = p; }

    // Far-field pairs in list order (per system; PC: per target batch its
    // approx list; CP: per source batch its approx list; CC: per target node
    // its CC, then PC, then CP entries).
    const std::vector<FarPair> &far_pairs_list() const { return pairs_; }
    // NUM_PAIR_FEATURES floats per pair (see bt_vardeg.cpp).
    void pair_features(std::vector<float> &f) const;
    // Gives every far-field pair its own degree (1..12) and builds and
    // uploads the per-degree work; nullptr returns to the fixed degree.
    std::string set_pair_degrees(const int *deg);
    // far-field pairs a degree model left at the fixed degree (features
    // outside its training range); set by the caller of set_pair_degrees
    void set_model_fallback_pairs(long long n) { model_fallback_pairs_ = n; }
    long long model_fallback_pairs() const { return model_fallback_pairs_; }
    bool variable_degree() const { return var_planned_; }
    double mean_pair_degree() const { return var_mean_deg_; }
    long long variable_far_evals() const { return var_evals_; }
    // Minimal degree in [2, 12] whose pair error (relative L2 over sampled
    // targets, sampled sources) is <= tol, for pairs [first, first+count);
    // 13 where degree 12 does not reach tol.  err (optional) gets the
    // errors at degrees 2..12 (11 per pair; NaN where not computed).
    std::string label_pairs(int first, int count, double tol, int *labels, double *err);
    const Counts &counts() const { return counts_; }
    const Timings &timings() const { return tm_; }
    const Params &params() const { return prm_; }
    double *out_dev() { return d_out_; }
    long long near_pairs() const { return near_.pairs; }
    // pairs (ordered) the item launch delivers at the plan's precision, the
    // pair evaluations it executes and its slots; 0 without it
    long long sym_pairs() const { return items_run() ? sym_pairs_ : 0; }
    long long sym_evals() const { return items_run() ? sym_evals_ : 0; }
    long long sym_slots() const { return items_run() ? sym_slots_len_ : 0; }
    long long far_pairs() const { return grid_.pairs; }
    int source_nodes() const { return nsnodes_; }
    int target_nodes() const { return ntnodes_; }
    int num_systems() const { return (int)sys_.size(); }
    int tpt() const { return tpt_; }
    bool persistent() const { return persistent_; }
    void set_persistent(bool p) { persistent_ = p; }
    // the symmetric near field on (default) or off for later executions
    void set_symmetric(bool on) { sym_enabled_ = on; }
    bool symmetric() const { return sym_enabled_; }
    int num_near_tiles() const { return (int)near_tiles_.size(); }
    int num_far_tiles() const { return (int)grid_tiles_.size(); }
    hipStream_t stream() const { return st_; }

private:
    template <class T> T *dalloc(size_t n);
    void dfree(void *p);  // frees a dalloc buffer before the plan ends
    // trees: on the host from private copies (host-only plans) or on the
    // device (GPU plans); then lists, work and uploads
    std::string host_trees(const PlanInput &in);
    std::string device_trees(const PlanInput &in);
    std::string finish_create(const PlanInput &in, bool host_only, double t2);
    template <class T, class A> std::string up(T **d, const std::vector<T, A> &h);
    template <class T, class A> std::string up_at(T *d, const std::vector<T, A> &h);
    void free_all();
    void build_work();
    // the item launch's items, slots and reduction lists (one system):
    // two-sided and one-sided items when its targets are its sources
    // (sym_ok_), one-sided items only otherwise (items_ok_); set by create
    void build_sym(const System &sy, const std::vector<int> &keys, const std::vector<std::vector<int>> &segs_of,
                   std::vector<std::vector<char>> &sym_of, bool symmetric);
    // the item launch runs in this execution's precision: always for a
    // symmetric plan, in mixed precision for one-sided items only
    bool items_run() const { return sym_ && sym_enabled_ && (sym_ok_ || prm_.precision == PREC_MIXED); }
    std::string upward_fixed();  // upward pass at the plan's degree
    std::string down_fixed();    // downpass at the plan's degree
    bool sym_ready();  // items_run(), and the slots allocated (at the first execution that uses them)
    std::string phase_events();
    // the tensor-grid launch serves this execution's grid targets: Coulomb
    // or Yukawa, or regularized Coulomb with skipping; FP64, or mixed
    // precision with acceptance-separated sources (theta < 0.99)
    bool pencil_grids() const;
    // reduce: add the slots to the potential now (else join_phases does)
    std::string near_sym(bool reduce = true);
    std::string launch_items();  // the fork event, then the item launch on st_
    std::string sym_reduce();
    // The phases beside the item launch: when the items run (and a second
    // stream is available), they go on st_ and every other phase of the
    // execution on aux_, which starts at the fork event recorded before the
    // items; fork_phases swaps st_ and aux_ so the phases' code launches on
    // aux_, join_phases swaps back, makes st_ wait for aux_ and adds the
    // slots to the potential (after the downpass: the order of those two
    // sums is fixed).
    bool overlap_ok();
    std::string fork_phases(bool ov);
    std::string join_phases(bool ov);
    struct AuxGuard {  // swaps back on an early return
        Plan &p;
        ~AuxGuard()
        {
            if (p.swapped_) { std::swap(p.st_, p.aux_); p.swapped_ = false; }
        }
    };
    std::string validate() const;
    void enumerate_pairs();
    void free_var();
    std::string validate_var() const;
    std::string execute_var(bool timed);
    // Hermite approximation (bt_hermite_plan.cpp)
    void build_hermite_work();
    std::string validate_hermite() const;
    std::string upload_hermite();
    std::string hermite_upward();
    std::string hermite_pc();
    std::string hermite_cp();
    std::string hermite_down();

    Params prm_;
    int M_ = 0, N_ = 0, tpt_ = 2, tile_targets_ = 512;
    bool persistent_ = false;
    int near_qlen_[btk::TILE_QUEUES] = {}, grid_qlen_[btk::TILE_QUEUES] = {};
    int nsnodes_ = 0, ntnodes_ = 0, sdepth_ = 0, tdepth_ = 0;
    std::vector<System> sys_;
    WorkCSR near_, grid_;
    std::vector<btk::Tile> near_tiles_, grid_tiles_, leaf_tiles_;
    // grid keys write rows of partial grid values (row of key k at
    // grid_row_[k]); node n sums rows [node_rows_off_[n], node_rows_off_[n+1])
    // per near key: centre of its target node and the r^2 below which the
    // mixed near field recomputes a pair from coordinate differences
    // (negated: the key's clusters are checked too)
    std::vector<double> near_ctr_;
    // Item launch of the near field's particle pairs (Coulomb or Yukawa, or
    // regularized Coulomb with skipping).  Symmetric plans (sym_ok_: one
    // system whose targets are its sources, the same particles in the same
    // tree order): pairs of near-field keys (A, B), each key's particle
    // range a source range of the other's near list, are evaluated once for
    // both directions (two-sided items), every other range of a key's list
    // once for the key (one-sided items), in both precisions.  One system
    // with other targets, skipping, mixed precision (items_ok_): every range
    // as a one-sided item.  The ranges carry SEG_SYM and the near launch
    // passes over them when the items run.  Item k writes A's sums to slots
    // [rslot, +|A|) and B's to [cslot, +|B|); the reduction adds, for each
    // particle of leaf list l, slots[base + i] for the bases
    // [sym_red_off_[l], sym_red_off_[l+1]).
    bool sym_ok_ = false, items_ok_ = false, sym_ = false, sym_enabled_ = true;
    long long model_fallback_pairs_ = 0;
    // the trees were built on the device; one tree and one set of particle
    // arrays serve as sources and targets (every system's targets are its
    // sources and its leaf sizes are equal)
    bool dev_tree_ = false, shared_ = false;
    std::vector<btk::SymItem> sym_items_;
    std::vector<double> sym_ctr_;
    std::vector<btk::Tile> sym_red_tiles_;
    std::vector<int> sym_red_off_;
    std::vector<long long> sym_red_base_;
    long long sym_slots_len_ = 0, sym_pairs_ = 0, sym_evals_ = 0;
    std::vector<long long> grid_row_;
    std::vector<int> node_rows_off_;
    long long grid_rows_len_ = 0;
    std::vector<int> up_levels_off_;  // per level: range into up_nodes_ / child CSR
    std::vector<int> up_nodes_, up_child_off_, up_child_id_;
    std::vector<int> down_levels_off_, down_nodes_;
    // global (all systems) host arrays for the kernels
    std::vector<int> g_sibeg_, g_siend_, g_sleaves_, g_tparent_;
    // upward pass of the leaves in parts (few leaves): part -> leaf node,
    // particle range, CSR of parts per node
    std::vector<int> up_part_node_, up_part_beg_, up_part_end_, up_part_off_;
    HostVec<int> g_torder_;
    Counts counts_;
    Timings tm_;
    btk::KParams kp_{};

    hipStream_t st_ = nullptr;
    hipStream_t aux_ = nullptr;                  // the phases beside the item launch
    hipEvent_t fork_ev_ = nullptr, join_ev_ = nullptr, red_ev_ = nullptr;
    bool swapped_ = false;                       // st_ and aux_ swapped (fork_phases, side_begin)
    bool forked_ = false;                        // the fork event precedes the started items
    std::vector<void *> allocs_;
    // variable degree
    std::vector<FarPair> pairs_;
    bool var_ = false;          // variable-degree work is on the device
    bool var_planned_ = false;  // variable-degree work is built (host)
    double var_mean_deg_ = 0;
    long long var_evals_ = 0;
    std::vector<DegreeWork> dw_;
    // Slots of degree p <= the plan's degree D go through the plan's own
    // degree-D grids: source slots take (E_p x E_p x E_p) of their node's
    // degree-D charges (fold_up_), target slots are carried onto their node's
    // degree-D grid and the plan's downpass (fold_down_); target node
    // fold_node_[b] sums the slots fold_item_[fold_off_[b] .. fold_off_[b+1]).
    // emat_ holds E_p (p = 1..D, (p+1) x (D+1) each) at emat_off_[p].
    bool fold_up_ = false, fold_down_ = false;
    std::vector<int> fold_node_, fold_off_, fold_item_;
    std::vector<double> emat_;
    std::vector<size_t> emat_off_;
    int *d_fold_node_ = nullptr, *d_fold_off_ = nullptr, *d_fold_item_ = nullptr;
    double *d_emat_ = nullptr;
    // one near launch (particle ranges + source slots of every degree) and
    // one far launch (every target slot) for all degrees
    WorkCSR vnear_, vfar_;
    std::vector<btk::Tile> vnear_tiles_, vfar_tiles_;
    std::vector<btk::Tile> vpencil_tiles_;  // vfar_ keys, one tile each
    btk::Tile *d_vp_tiles_ = nullptr;
    int vnear_qlen_[btk::TILE_QUEUES] = {}, vfar_qlen_[btk::TILE_QUEUES] = {};
    std::vector<long long> vfar_row_;
    long long vfar_rows_len_ = 0;
    int *d_vn_clu_off_ = nullptr, *d_vn_clu_id_ = nullptr;
    btk::Tile *d_vn_tiles_ = nullptr;
    int *d_vf_seg_off_ = nullptr, *d_vf_clu_off_ = nullptr, *d_vf_clu_id_ = nullptr;
    btk::Seg *d_vf_segs_ = nullptr;
    btk::Tile *d_vf_tiles_ = nullptr;
    long long *d_vf_row_ = nullptr;
    double *d_vf_R_ = nullptr, *d_vf_RW_ = nullptr;
    std::vector<void *> var_allocs_;
    bool alloc_var_ = false;  // route dalloc to var_allocs_
    // device arrays
    double *d_tx_ = nullptr, *d_ty_ = nullptr, *d_tz_ = nullptr, *d_tq_ = nullptr;
    double *d_sx_ = nullptr, *d_sy_ = nullptr, *d_sz_ = nullptr, *d_sq_ = nullptr, *d_sw_ = nullptr;
    int *d_torder_ = nullptr;
    double *d_stab_ = nullptr, *d_ttab_ = nullptr;
    double *d_cx_ = nullptr, *d_cy_ = nullptr, *d_cz_ = nullptr, *d_cq_ = nullptr, *d_cw_ = nullptr;
    double *d_Qt_ = nullptr, *d_Wt_ = nullptr;
    int *d_sibeg_ = nullptr, *d_siend_ = nullptr, *d_tparent_ = nullptr;
    int *d_sleaves_ = nullptr, *d_up_nodes_ = nullptr, *d_up_child_off_ = nullptr, *d_up_child_id_ = nullptr;
    int *d_up_part_node_ = nullptr, *d_up_part_beg_ = nullptr, *d_up_part_end_ = nullptr, *d_up_part_off_ = nullptr;
    double *d_up_Qp_ = nullptr, *d_up_Wp_ = nullptr;
    // child-to-parent entries: parent of each, identity offsets, rows
    int *d_up_entry_parent_ = nullptr, *d_up_entry_off_ = nullptr;
    double *d_up_Qe_ = nullptr, *d_up_We_ = nullptr;
    int *d_down_nodes_ = nullptr;
    btk::Tile *d_near_tiles_ = nullptr, *d_grid_tiles_ = nullptr, *d_leaf_tiles_ = nullptr;
    // grid_ keys, one tile each (the tensor-grid launch of mixed executions)
    std::vector<btk::Tile> pencil_tiles_;
    btk::Tile *d_pencil_tiles_ = nullptr;
    int *d_nseg_off_ = nullptr, *d_nclu_off_ = nullptr, *d_nclu_id_ = nullptr;
    btk::Seg *d_nsegs_ = nullptr;
    int *d_gseg_off_ = nullptr, *d_gclu_off_ = nullptr, *d_gclu_id_ = nullptr;
    btk::Seg *d_gsegs_ = nullptr;
    double *d_pot_ = nullptr, *d_out_ = nullptr;
    int *d_cursor_ = nullptr;  // tile-queue cursors: near, then far
    double *d_near_ctr_ = nullptr;
    btk::SymItem *d_sym_items_ = nullptr;
    double *d_sym_ctr_ = nullptr, *d_sym_slots_ = nullptr;  // slots: at the first mixed execution
    btk::Tile *d_sym_red_tiles_ = nullptr;
    int *d_sym_red_off_ = nullptr;
    long long *d_sym_red_base_ = nullptr;
    hipEvent_t sym_ev_[2] = {nullptr, nullptr};  // around the symmetric launch
    hipEvent_t ev_[6] = {};                      // phase boundaries of execute()
    bool sym_ran_ = false;                       // it ran in this execution
    bool items_started_ = false;                 // enqueued ahead of execute() (start_items)
    long long *d_grid_row_ = nullptr;
    int *d_node_rows_off_ = nullptr;
    double *d_Qrows_ = nullptr, *d_Wrows_ = nullptr;
    // Hermite: 8 planes per grid point.  Upward pass from every source
    // node's particles in parts (node n: parts [hp_off_[n], hp_off_[n+1]));
    // PC: the near keys' clusters in their own launch (hpc_, hpc_tiles_);
    // CP: grid tiles of 256 points (hcp_tiles_), rows of 8P values; the
    // downpass of every target node with rows (hdown_tiles_)
    bool herm_ = false;
    std::vector<int> hp_node_, hp_beg_, hp_end_, hp_off_;
    WorkCSR hpc_;
    std::vector<btk::Tile> hpc_tiles_, hcp_tiles_, hdown_tiles_;
    std::vector<double> swtab_, twtab_;
    int *d_hp_node_ = nullptr, *d_hp_beg_ = nullptr, *d_hp_end_ = nullptr, *d_hp_off_ = nullptr;
    double *d_Qp_ = nullptr, *d_Wp_ = nullptr, *d_swtab_ = nullptr, *d_twtab_ = nullptr;
    int *d_hpc_clu_off_ = nullptr, *d_hpc_clu_id_ = nullptr;
    btk::Tile *d_hpc_tiles_ = nullptr, *d_hcp_tiles_ = nullptr, *d_hdown_tiles_ = nullptr;
};

}  // namespace bt
