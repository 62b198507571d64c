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

struct Timings {
    double host_copy = 0, host_tree = 0, host_lists = 0, host_work = 0, h2d = 0;  // seconds
    double gpu_upward = 0, gpu_near = 0, gpu_far = 0, gpu_down = 0, gpu_finish = 0, gpu_total = 0;
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

// Input of a plan: nsys systems, system s owning targets [toffs[s],
// toffs[s+1]) and sources [soffs[s], soffs[s+1]) of the arrays.
struct PlanInput {
    int nsys = 1;
    const int *toffs = nullptr, *soffs = nullptr;  // nsys+1 entries each
    const int *msl = nullptr, *mtl = nullptr;      // per-system leaf sizes (optional)
    int tpt = 2;                                   // targets per thread (tile = 256*tpt)
    int persistent = 0;                            // persistent interaction launches
    const double *tx = nullptr, *ty = nullptr, *tz = nullptr, *tq = nullptr;
    const double *sx = nullptr, *sy = nullptr, *sz = nullptr, *sq = nullptr, *sw = nullptr;
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
                       bool host_only = false);
    // Runs the GPU path; result in caller order, on the device (out_dev()).
    std::string execute(bool timed);
    // Copies the device result to host memory.
    std::string download(double *out);
    // FP64 direct sum on the GPU for caller-order target indices `idx`
    // (each target against the sources of its own system).
    std::string direct(const std::vector<int> &idx, std::vector<double> &out);

    void set_precision(int p) { prm_.precision = p; }
    const Counts &counts() const { return counts_; }
    const Timings &timings() const { return tm_; }
    const Params &params() const { return prm_; }
    double *out_dev() { return d_out_; }
    long long near_pairs() const { return near_.pairs; }
    long long far_pairs() const { return grid_.pairs; }
    int source_nodes() const { return nsnodes_; }
    int target_nodes() const { return ntnodes_; }
    int num_systems() const { return (int)sys_.size(); }
    int tpt() const { return tpt_; }
    bool persistent() const { return persistent_; }
    void set_persistent(bool p) { persistent_ = p; }
    int num_near_tiles() const { return (int)near_tiles_.size(); }
    int num_far_tiles() const { return (int)grid_tiles_.size(); }
    hipStream_t stream() const { return st_; }

private:
    template <class T> T *dalloc(size_t n);
    template <class T, class A> std::string up(T **d, const std::vector<T, A> &h);
    template <class T, class A> std::string up_at(T *d, const std::vector<T, A> &h);
    void free_all();
    void build_work();
    std::string validate() const;

    Params prm_;
    int M_ = 0, N_ = 0, tpt_ = 2, tile_targets_ = 512;
    bool persistent_ = false;
    int near_qlen_[btk::TILE_QUEUES] = {}, grid_qlen_[btk::TILE_QUEUES] = {};
    int nsnodes_ = 0, ntnodes_ = 0, sdepth_ = 0, tdepth_ = 0;
    std::vector<System> sys_;
    WorkCSR near_, grid_;
    std::vector<btk::Tile> near_tiles_, grid_tiles_, leaf_tiles_;
    std::vector<int> up_levels_off_;  // per level: range into up_nodes_ / child CSR
    std::vector<int> up_nodes_, up_child_off_, up_child_id_;
    std::vector<int> down_levels_off_, down_nodes_;
    // global (all systems) host arrays for the kernels
    std::vector<int> g_sibeg_, g_siend_, g_sleaves_, g_tparent_;
    HostVec<int> g_torder_;
    Counts counts_;
    Timings tm_;
    btk::KParams kp_{};

    hipStream_t st_ = nullptr;
    std::vector<void *> allocs_;
    // device arrays
    double *d_tx_ = nullptr, *d_ty_ = nullptr, *d_tz_ = nullptr, *d_tq_ = nullptr;
    double *d_sx_ = nullptr, *d_sy_ = nullptr, *d_sz_ = nullptr, *d_sq_ = nullptr, *d_sw_ = nullptr;
    int *d_torder_ = nullptr;
    double *d_stab_ = nullptr, *d_ttab_ = nullptr;
    double *d_cx_ = nullptr, *d_cy_ = nullptr, *d_cz_ = nullptr, *d_cq_ = nullptr, *d_cw_ = nullptr;
    double *d_Qt_ = nullptr, *d_Wt_ = nullptr;
    int *d_sibeg_ = nullptr, *d_siend_ = nullptr, *d_tparent_ = nullptr;
    int *d_sleaves_ = nullptr, *d_up_nodes_ = nullptr, *d_up_child_off_ = nullptr, *d_up_child_id_ = nullptr;
    int *d_down_nodes_ = nullptr;
    btk::Tile *d_near_tiles_ = nullptr, *d_grid_tiles_ = nullptr, *d_leaf_tiles_ = nullptr;
    int *d_nseg_off_ = nullptr, *d_nclu_off_ = nullptr, *d_nclu_id_ = nullptr;
    btk::Seg *d_nsegs_ = nullptr;
    int *d_gseg_off_ = nullptr, *d_gclu_off_ = nullptr, *d_gclu_id_ = nullptr;
    btk::Seg *d_gsegs_ = nullptr;
    double *d_pot_ = nullptr, *d_out_ = nullptr;
    int *d_cursor_ = nullptr;  // tile-queue cursors: near, then far
};

}  // namespace bt
