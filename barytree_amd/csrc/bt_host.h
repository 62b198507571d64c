// bt_host.h -- host-side data structures of the MI355X BaryTree path.
//
// The host owns what the reference also does on the host: run parameters,
// private (tree-ordered) particle copies, the source/target trees or
// batches, and the interaction lists.  Everything numerical after that runs
// in HIP kernels (bt_kernels.hip).
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <utility>
#include <vector>

namespace bt {

// Enum values are the reference ABI (src/interface/BaryTreeInterface.h:7-47).
enum KernelId {
    K_NONE = 0, K_COULOMB = 1, K_YUKAWA = 2, K_REG_COULOMB = 3, K_REG_YUKAWA = 4, K_ATAN = 5,
    K_SIN_OVER_R = 8, K_MQ = 9, K_RBS_U = 10, K_RBS_V = 11, K_USER = 12
};
// the Coulomb family (mixed precision, Hermite, learned degrees); the other
// kernels run the Lagrange approximation in FP64
inline bool coulomb_family(int k) { return k == K_COULOMB || k == K_YUKAWA || k == K_REG_COULOMB; }
enum SingularityId { S_NONE = 0, S_SKIPPING = 1, S_SUBTRACTION = 2 };
enum ApproxId { A_NONE = 0, A_LAGRANGE = 1, A_HERMITE = 2 };
enum ComputeId { C_NONE = 0, C_PC = 1, C_CP = 2, C_CC = 3 };

// Arithmetic of the pair kernels.
enum PrecisionId {
    PREC_FP64 = 0,   // plain mode: everything in FP64
    PREC_MIXED = 1,  // FP32 pair evaluation, compensated FP64 accumulation
};

// RunParams semantics (src/run_params/run_params.c:13-149).
struct Params {
    int kernel = K_NONE, singularity = S_NONE, approximation = A_NONE, compute_type = C_NONE;
    double kp[4] = {0, 0, 0, 0};
    int num_kp = 0;
    double theta = 0.0, size_check_factor = 1.0, beta = -1.0;
    int degree = 0, P = 0;  // P = (degree+1)^3
    int max_source_leaf = 1, max_target_leaf = 1;
    int verbosity = 0;
    int precision = PREC_FP64;
};

// Leaf sizes of the beta auto-tune on this thread (run_params.c:100-117):
// 0 the reference CPU build's (PC 50/5, CP 5/50, CC 50/50; the default, so
// beta runs reproduce the reference CPU results), 1 its GPU build's
// (3000/3000).
void set_beta_leaf_mode(int mode);
int beta_leaf_mode();
// Resolve (theta, degree, leaf sizes) from beta exactly as RunParams_Setup.
void params_setup(Params &p, int kernel, int num_kp, const double *kp, int approximation,
                  int singularity, int compute_type, double theta, int degree, int msl, int mtl,
                  double size_check, double beta, int verbosity);

// Returns an empty string when the combination is supported, otherwise the
// error text (the reference prints and exits for these combinations).
std::string params_check(const Params &p);

// Allocator whose resize() leaves new elements uninitialised, so that large
// arrays are first touched by the (parallel) loop that fills them.
template <class T> struct NoInitAlloc : std::allocator<T> {
    template <class U> struct rebind { using other = NoInitAlloc<U>; };
    NoInitAlloc() = default;
    template <class U> NoInitAlloc(const NoInitAlloc<U> &) {}
    template <class U> void construct(U *p) { ::new (static_cast<void *>(p)) U; }
    template <class U, class... A> void construct(U *p, A &&...a)
    {
        ::new (static_cast<void *>(p)) U(std::forward<A>(a)...);
    }
};
template <class T> using HostVec = std::vector<T, NoInitAlloc<T>>;

// Private SoA particle copy, permuted into tree order; order[i] is the
// caller index (0-based) of sorted particle i.
struct Particles {
    int n = 0;
    HostVec<double> x, y, z, q, w;
    HostVec<int> order;
};

// Array tree (src/tree/struct_tree.h), nodes in preorder.  ibeg/iend are
// 0-based [ibeg, iend) here.
struct Tree {
    int n = 0;
    std::vector<int> ibeg, iend, numpar, nchild, child /*8 per node*/, parent, level;
    std::vector<double> xmin, xmax, ymin, ymax, zmin, zmax, xmid, ymid, zmid, radius;
    int max_depth = 0;
    std::vector<int> leaves;                 // leaf node ids, DFS order
    std::vector<std::vector<int>> levels;    // non-leaf node ids per level, DFS order
    int add_node();
};

// One node of a tree under construction; child ids are local to the tree's
// own node vector (preorder, root = 0).  b and e are 0-based [b, e).
struct NodeRec {
    int b, e, level, nchild;
    int child[8];
    double lo[3], hi[3], mid[3], radius;
};

// From a node's tight box (r.lo, r.hi): its centre and radius
// (tree_linked_list.c:53-69); if numpar > maxpar the dimensions to split, in
// the order x, y, z (dv[d] = 1), by the child count and tie rules of
// tree_linked_list.c:81-91 and partition.c:178-286.  Returns false for a
// leaf (dv all 0).
bool node_split(int numpar, int maxpar, NodeRec &r, int dv[3]);

// Fills a Tree from node records in preorder (leaves_only: the records are
// the leaves only, in DFS order).
void finish_tree(Tree &t, const std::vector<NodeRec> &nodes, int max_depth, bool leaves_only);

// A node of more coincident particles than the leaf size can never be split
// (the reference recurses without end there); the builders report this.
constexpr const char *COINCIDENT_ERROR =
    "more coincident particles than the maximum leaf size: the tree cannot be split further";

// Build the tree over p (permuting p).  leaves_only keeps only the leaves
// (batches, src/tree/batches.c:25-90).  Returns "" or COINCIDENT_ERROR.
std::string build_tree(Tree &t, Particles &p, int max_per_leaf, bool leaves_only);

// Several trees at once (for example the source and target trees of a
// cluster-cluster run), sharing one pool of OpenMP threads.
struct TreeJob {
    Tree *t;
    Particles *p;
    int max_per_leaf;
    bool leaves_only;
};
std::string build_trees(const std::vector<TreeJob> &jobs);

// Jagged interaction lists (src/interaction_lists/interaction_lists.c),
// keyed by the "target" argument of InteractionLists_Make.
struct Lists {
    std::vector<std::vector<int>> pp, approx, cc_pc, cc_cp;
};

void make_pc_lists(Lists &L, const Tree &src, const Tree &key, const Params &p);
void make_cc_lists(Lists &L, const Tree &src, const Tree &tgt, const Params &p);

// Chebyshev node table: per node, 3*(degree+1) coordinates (x nodes, then
// y, then z), computed exactly as src/clusters/clusters.c:436-440.
void node_table(const Tree &t, int degree, std::vector<double> &tab);

// Hermite barycentric weights: per node, 3*(degree+1) values w_d[j] =
// -4 ww[j] / (max_d - min_d) with ww[j] = -cos(t_j) / (2 sin^2 t_j),
// t_j = j*pi/degree, and ww[0] = -ww[degree] = (degree^2/3 + 1/6)/4
// (src/clusters/clusters.c:1296-1318).
void hermite_weight_table(const Tree &t, int degree, std::vector<double> &tab);

// Source-range flags: a source may coincide with a target (the singularity
// skip may trigger); a pair may come closer than its key's threshold.
constexpr int SEG_COINCIDE = 1, SEG_CLOSE = 2, SEG_SYM = 4;

// Device-side work description (CSR), built on the host.
struct WorkCSR {
    // keys: target leaves (particle targets) or target nodes (grid targets)
    std::vector<int> key;          // tree node id of the key (or batch id)
    std::vector<int> tbeg, tcnt;   // target range (particles) or grid size
    std::vector<int> seg_off;      // CSR into seg_beg/seg_cnt (source particle ranges)
    std::vector<int> seg_beg, seg_cnt;
    std::vector<char> seg_chk;     // SEG_COINCIDE | SEG_CLOSE | SEG_SYM flags
    std::vector<int> clu_off;      // CSR into clu_id (source clusters)
    std::vector<int> clu_id;
    long long pairs = 0;           // point-wise evaluations described
};

struct Counts {
    long long num_pp = 0, num_pc = 0, num_cp = 0, num_cc = 0;
    long long pp_ptwise = 0, pc_ptwise = 0, cp_ptwise = 0, cc_ptwise = 0;
    long long pc_replaced = 0, cp_replaced = 0, cc_replaced = 0;
};

}  // namespace bt
