// bt_kernels.h -- launch interface of the gfx950 kernels (bt_kernels.hip).
// Plain C++ declarations so that host code compiled without hipcc can call
// them; every pointer here is a device pointer.
#pragma once

#include <hip/hip_runtime_api.h>
#include <hip/hip_vector_types.h>

namespace btk {

// Work tile {key, first target, target count, node} and source range
// {first, count}; 16- and 8-byte aligned so kernels load them whole.
struct alignas(16) Tile { int x, y, z, w; };
struct alignas(8) Seg { int x, y; };

// Seg.y = count | flags: the singularity skip may trigger in the range;
// a pair of the range may come closer than its key's threshold (key_ctr);
// the range's pairs with the key's particles run in the item launch
// (near launches with skip_sym pass over it).
constexpr unsigned SEG_COINCIDE_BIT = 0x80000000u;
constexpr unsigned SEG_CLOSE_BIT = 0x40000000u;
constexpr unsigned SEG_SYM_BIT = 0x20000000u;
constexpr unsigned SEG_COUNT_MASK = 0x1fffffffu;

// Tile lists are 8 interleaved queues (tile 8*j + x = j-th tile of queue x),
// one per XCD of an MI355X.
constexpr int TILE_QUEUES = 8;

constexpr int MAX_DEGREE = 12;
constexpr int DEG_SHIFT = 24;               // (degree << DEG_SHIFT) | slot
constexpr int SLOT_MASK = (1 << DEG_SHIFT) - 1;

// Kernel parameters, pre-derived on the host.
struct KParams {
    double p0, p1;      // kernel_params[0], kernel_params[1]
    double inv_a2;      // 1/kappa^2 (Coulomb SS), 1/alpha^2 (reg. Coulomb SS)
    double eps2;        // epsilon^2 (regularized Coulomb and Yukawa)
    double ss_param;    // subtraction correction factor
    // atan: delta factor sqrt(1 + 1/delta^2) and wadj; mq: (delta/L)^2 and
    // sqrt(1 + 4 (delta/L)^2); rbs-u, rbs-v: delta^2 (in c1)
    double c1, c2;
};

// Kernel ids beside Coulomb (1), Yukawa (2) and regularized Coulomb (3)
// (src/utilities/enums.h): these run in FP64 whatever the precision asked.
constexpr int KERNEL_REG_YUKAWA = 4, KERNEL_ATAN = 5, KERNEL_SIN_OVER_R = 8, KERNEL_MQ = 9,
              KERNEL_RBS_U = 10, KERNEL_RBS_V = 11, KERNEL_USER = 12;

// Interaction work: a list of tiles {key, first target, target count, node}
// against CSR lists of source particle ranges and source clusters.
struct InteractArgs {
    int kernel, singularity;
    int grid_targets;        // 0: particle targets, 1: grid points of target nodes
    int precision;           // 0: FP64, 1: FP32 pairs with compensated accumulation
    int tpt;                 // targets per thread of particle-target tiles (1 or 2;
                             // grid-target tiles always use 2)
    int ntiles;              // tile slots (8 * longest queue)
    const Tile *tiles;
    int persistent;          // 0: one block per slot; 1: resident blocks pop from queues
    int qlen[TILE_QUEUES];   // tiles per queue (persistent)
    int *cursor;             // TILE_QUEUES counters (persistent; zeroed by the launch)
    const int *seg_off;
    const Seg *segs;        // {first source, count}
    const int *clu_off, *clu_id;
    int P, nl;
    const double *tnode_tab; // target node tables (grid targets)
    const double *tx, *ty, *tz, *tq;
    const double *sx, *sy, *sz, *sq, *sw;
    const double *cx, *cy, *cz, *cq, *cw;  // source cluster points and charges
    KParams kp;
    double *out1, *out2;     // potential (particles) or target-grid rows of Q and W
    int accumulate;          // particle targets: 1 adds to out1, 0 stores
    // grid targets: the output row of key k starts at out_row[k] (grid point
    // j of a tile of key k goes to out1[out_row[k] + j]); null: row = node*P
    const long long *out_row;
    // several degrees in one launch: a cluster id or a tile's node field is
    // (degree << DEG_SHIFT) | slot, and the arrays of degree d are d*[d]
    int multi;
    const double *dcx[MAX_DEGREE + 1], *dcy[MAX_DEGREE + 1], *dcz[MAX_DEGREE + 1];
    const double *dcq[MAX_DEGREE + 1], *dcw[MAX_DEGREE + 1];
    const double *dtab[MAX_DEGREE + 1];  // target slot node tables
    // particle targets, mixed precision (required there): r^2 is formed as
    // |t-c|^2 + |s-c|^2 - 2 (t-c).(s-c) from coordinates relative to the key's
    // centre c = key_ctr[k].xyz; pairs with r^2 below |key_ctr[k].w| (in
    // ranges flagged SEG_CLOSE, and in the key's clusters if .w < 0) are
    // recomputed from the coordinate differences
    const double4 *key_ctr;
    int skip_sym;            // pass over source ranges flagged SEG_SYM_BIT
    // mixed grid targets: r^2 from the expansion around the target node's
    // centre (valid for acceptance-separated sources, theta < 0.99)
    int grid_scaled;
};

hipError_t launch_interact(const InteractArgs &a, hipStream_t st);

// Grid targets of Coulomb, Yukawa (either singularity) and regularized
// Coulomb (skipping), the grids as tensor products: one block per tile
// {key, 0, nl^3, node} of a.tiles, nl = a.nl (a.multi: the node field's
// degree + 1, 2..13); the other fields as for launch_interact's grid
// targets.  a.precision 0: FP64 displacements; 1: the centred expansion
// (grid_scaled), for acceptance-separated sources only.
hipError_t launch_grid_pencil(const InteractArgs &a, hipStream_t st);

// Item launch of the near field's particle pairs (Coulomb and Yukawa with
// either singularity, regularized Coulomb with skipping; FP64 or mixed
// precision): one system whose targets are its sources, or (one-sided items
// only) one system with other targets.  A two-sided item (cslot >= 0) is a pair
// of distinct leaves A, B, each in the other's near list: every pair (a, b)
// is evaluated once and gives both a's term from b and b's term from a.
// A one-sided item (cslot < 0) is a leaf A with a source range B of its near
// list that does not list A back (or A itself): A's terms only.  a's sums go
// to slots[rslot + i] (i-th particle of A), b's to slots[cslot + j];
// ctr[k] = centre of the pair's boxes and its r^2 threshold (negative:
// pairs below it are recomputed from coordinate differences, with the skip
// of coincident pairs).
struct SymItem {
    int a_beg, a_cnt, b_beg, b_cnt;
    long long rslot, cslot;
};
struct SymArgs {
    int kernel, singularity, nitems;
    int fp64;  // 1: FP64 arithmetic (plain executions), 0: mixed precision
    const SymItem *items;
    const double4 *ctr;
    const double *x, *y, *z, *q, *w;  // the sources, tree order (w: subtraction only)
    const double *tx, *ty, *tz, *tq;   // the targets, tree order (symmetric plans: the sources)
    KParams kp;
    double *slots;
};
hipError_t launch_near_sym(const SymArgs &a, hipStream_t st);

// Adds the slots of every particle to its potential, in list order: tile
// {list, first particle, count}; particle i of list l adds
// slots[base[k] + i] for k in [off[l], off[l+1]).
constexpr int TPB_SYM_REDUCE = 256;  // particles per reduction tile (one per thread)
struct SymReduceArgs {
    int ntiles;
    const Tile *tiles;
    const int *off;
    const long long *base;
    const double *slots;
    double *pot;
};
hipError_t launch_sym_reduce(const SymReduceArgs &a, hipStream_t st);

// Hermite approximation (bt_hermite.hip), FP64.  Clusters carry 8 planes
// per grid point; the planes of node c sit at [8*P*c + m*P + j].
// Upward pass: block (b, slice) = particles [part_beg[b], part_end[b]) of
// node part_node[b] against grid points slice*256 + thread; the part's row
// of 8P values goes to Qp[8P*b] (and its weights to Wp).
struct HermAnterpArgs {
    int degree, subtraction, nparts;
    const int *part_node, *part_beg, *part_end;
    const double *tab, *wtab;  // per node 3(n+1): Chebyshev nodes; Hermite weights w
    const double *px, *py, *pz, *pq, *pw;
    double *Qp, *Wp;
};
hipError_t launch_herm_anterp(const HermAnterpArgs &a, hipStream_t st);
// Particle-cluster far field: tile {key, first target, count} against the
// clusters clu_id[clu_off[key] ..) of its key, added to pot.
struct HermPCArgs {
    int kernel, singularity, ntiles, P;
    const Tile *tiles;
    const int *clu_off, *clu_id;
    const double *cx, *cy, *cz, *cq, *cw;
    const double *tx, *ty, *tz, *tq;
    KParams kp;
    double *pot;
};
hipError_t launch_herm_pc(const HermPCArgs &a, hipStream_t st);
// Cluster-particle far field: tile {key, first grid point, count, node}
// against the source ranges of its key; grid point j's 8 plane sums go to
// rows[out_row[key] + m*P + j].
struct HermCPArgs {
    int kernel, ntiles, nl;
    const Tile *tiles;
    const int *seg_off;
    const Seg *segs;
    const double *tab;
    const double *sx, *sy, *sz, *sq;
    KParams kp;
    const long long *out_row;
    double *rows;
};
hipError_t launch_herm_cp(const HermCPArgs &a, hipStream_t st);
// Downpass: tile {leaf, first target, count}; each target adds the 8-plane
// interpolants of its leaf and every ancestor n with rows (rows_off[n+1] >
// rows_off[n]) to pot.
struct HermDownArgs {
    int degree, ntiles;
    const Tile *tiles;
    const int *parent, *rows_off;
    const double *tab, *wtab;
    const double *tx, *ty, *tz;
    const double *Q;
    double *pot;
};
hipError_t launch_herm_down(const HermDownArgs &a, hipStream_t st);

// Source grid point coordinates from node tables: c{x,y,z}[node*P + j].
hipError_t launch_cluster_points(int numnodes, int nl, const double *tab, double *cx, double *cy,
                                 double *cz, hipStream_t st);

// Upward pass.  mode 0: one block per leaf in `nodes`, points = particles
// [ibeg, iend) of the node, output row = the node; mode 1: one block per
// parent in `nodes`, points = the grids of its children (child CSR
// child_off/child_id); mode 2: block b takes the node tables of nodes[b],
// points = particles [ibeg[b], iend[b]), output row b (partial charges).
// mode 0: block b anterpolates the particles of node nodes[b] into its row;
// mode 2: the particles [ibeg[b], iend[b]) of node nodes[b] into row b;
// mode 1: the grids of node nodes[b]'s children child_id[child_off[b] ..
// child_off[b+1]) into its row; mode 3: as mode 1 into row child_off[b] of
// Qo (Wo), the children's charges read from Q (W).
struct AnterpArgs {
    int mode, degree, subtraction, nblocks;
    const int *nodes;
    const int *ibeg, *iend;
    const int *child_off, *child_id;
    const double *tab;                                  // node tables
    const double *px, *py, *pz, *pq, *pw;              // particles (mode 0)
    double *Q, *W;                                      // cluster charges (in/out)
    double *Qo, *Wo;  // mode 3: the entries' rows (the children's charges are read from Q, W)
};
hipError_t launch_anterp(const AnterpArgs &a, hipStream_t st);
// Q[d][j] = sum_{k in [off[s], off[s+1])} Qp[k][j] (W likewise if W is set), in k order,
// d = dst[s] (s without dst).
hipError_t launch_sum_parts(int nslots, int P, const int *off, const double *Qp, const double *Wp,
                            double *Q, double *W, hipStream_t st, const int *dst = nullptr);
// Q[s][j] = sum_{k in [off[s], off[s+1])} R[row[k] + j] (W from RW likewise), in k order;
// zero where a slot has no rows.
hipError_t launch_sum_rows(int nslots, int P, const int *off, const long long *row, const double *R,
                           const double *RW, double *Q, double *W, hipStream_t st);

// Downward pass.  mode 0: one block per child in `nodes`; the child grid
// gets the interpolant of its parent's values.  mode 1: one block per tile
// {leaf, first target, count} of `tiles`; targets get the leaf interpolant.
struct DownArgs {
    int mode, degree, subtraction, nblocks;
    const int *nodes, *parent;
    const Tile *tiles;
    const double *tab;
    const double *tx, *ty, *tz, *tq;
    double *Q, *W;
    double *pot;
};
hipError_t launch_down(const DownArgs &a, hipStream_t st);

// Change of degree of tensor Chebyshev grids (bt_degree.hip), for plans
// whose far-field pairs carry their own degrees.  E[p] is the (p+1) x (D+1)
// matrix of the 1D degree-p Lagrange basis at the degree-D Chebyshev
// points of the reference interval, row major.
// up: slot s of degree p gets (E x E x E) of the degree-D grid of node[s]
// (QD, WD: degree-D grids of all nodes; Qp, Wp: the slots' grids).
struct DegreeUpArgs {
    int nslots, p, D, subtraction;
    const int *node;
    const double *E;
    const double *QD, *WD;
    double *Qp, *Wp;
};
hipError_t launch_degree_up(const DegreeUpArgs &a, hipStream_t st);
// down: block b stores into the degree-D grid of node[b] the sum over its
// items [off[b], off[b+1]) ((p << DEG_SHIFT) | slot, in that order) of
// (E^T x E^T x E^T) of the slot's degree-p grid (Qs[p], Ws[p]).
struct DegreeDownArgs {
    int nnodes, D, subtraction;
    const int *node, *off, *item;
    const double *E[MAX_DEGREE + 1];
    const double *Qs[MAX_DEGREE + 1], *Ws[MAX_DEGREE + 1];
    double *QD, *WD;
};
hipError_t launch_degree_down(const DegreeDownArgs &a, hipStream_t st);

// Subtraction correction and scatter to the caller's order:
// out[order[i]] = pot[i] + ss_param*tq[i] (ss_param = 0 without subtraction).
hipError_t launch_finish(int n, const double *pot, const double *tq, double ss_param,
                         const int *order, double *out, hipStream_t st);

// FP64 direct sum (directdriver semantics) for sampled targets, no tree:
// out[i] = sum_j G(t_i, s_j) (+ correction).  The sources are split into
// nchunks chunks over blocks; part holds nchunks*m partial sums.
hipError_t launch_direct(int kernel, int singularity, KParams kp, int m, const double *tx,
                         const double *ty, const double *tz, const double *tq, int n,
                         const double *sx, const double *sy, const double *sz, const double *sq,
                         const double *sw, double *part, int nchunks, double *out, hipStream_t st);

}  // namespace btk
