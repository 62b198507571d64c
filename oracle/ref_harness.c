/*
 * ref_harness.c -- drives the reference's own compiled sources (test
 * infrastructure only; built into oracle/_ref by `make -C oracle ref`).
 *
 * The reference library needs MPI for tree.c, batches.c, clusters.c,
 * particles.c and the drivers, and this image has no MPI.  The rest of the
 * hot path is MPI-free and is compiled from where it lies:
 *   run_params.c, tools.c, xmalloc.c, partition.c, tree_linked_list.c,
 *   interaction_lists.c, interaction_compute_{pc,cp,cc,downpass,direct,
 *   correction}.c and the kernel sources.
 * ref_treecode() below runs treedriver's single-rank phase order
 * (src/drivers/treedriver.c:114-967) with those reference functions, and
 * fills the gaps (Tree_Fill and the level lists of tree.c, the batch fill of
 * batches.c, the cluster upward pass of clusters.c -- Lagrange and Hermite --,
 * the reorder of particles.c) with the oracle's restatement, which is
 * included below.
 *
 * Agreement of ref_treecode() with orc_treecode() therefore pins the
 * oracle's partition, tree, interaction-list, kernel, dispatch and downpass
 * code against the reference itself; the upward pass is pinned separately
 * (invariants and the known-answer values).
 */
#include "bt_oracle.c"

#include "utilities/enums.h"
#include "utilities/tools.h"
#include "run_params/struct_run_params.h"
#include "run_params/run_params.h"
#include "particles/struct_particles.h"
#include "tree/struct_tree.h"
#include "tree/struct_tree_linked_list_node.h"
#include "tree/tree_linked_list.h"
#include "tree/partition.h"
#include "clusters/struct_clusters.h"
#include "interaction_lists/struct_interaction_lists.h"
#include "interaction_lists/interaction_lists.h"
#include "interaction_compute/interaction_compute.h"

static int *ivec(int n) { return calloc((size_t)(n > 0 ? n : 1), sizeof(int)); }
static double *dvec(int n) { return calloc((size_t)(n > 0 ? n : 1), sizeof(double)); }

static struct Tree *h_tree_alloc(int n)
{
    struct Tree *t = calloc(1, sizeof(struct Tree));
    t->numnodes = n;
    t->ibeg = ivec(n); t->iend = ivec(n); t->numpar = ivec(n); t->cluster_ind = ivec(n);
    t->x_mid = dvec(n); t->y_mid = dvec(n); t->z_mid = dvec(n);
    t->x_min = dvec(n); t->y_min = dvec(n); t->z_min = dvec(n);
    t->x_max = dvec(n); t->y_max = dvec(n); t->z_max = dvec(n);
    t->radius = dvec(n); t->num_children = ivec(n); t->children = ivec(8 * n); t->parent = ivec(n);
    return t;
}

static void h_tree_fill(struct Tree *t, struct TreeLinkedListNode *p)
{
    int i = p->node_index;
    t->x_mid[i] = p->x_mid; t->y_mid[i] = p->y_mid; t->z_mid[i] = p->z_mid;
    t->x_min[i] = p->x_min; t->y_min[i] = p->y_min; t->z_min[i] = p->z_min;
    t->x_max[i] = p->x_max; t->y_max[i] = p->y_max; t->z_max[i] = p->z_max;
    t->ibeg[i] = p->ibeg; t->iend[i] = p->iend; t->numpar[i] = p->numpar;
    t->radius[i] = p->radius; t->cluster_ind[i] = i; t->num_children[i] = p->num_children;
    t->parent[i] = p->parent ? p->parent->node_index : -1;
    for (int k = 0; k < p->num_children; k++) {
        t->children[8 * i + k] = p->child[k]->node_index;
        h_tree_fill(t, p->child[k]);
    }
}

static void h_levels_rec(struct Tree *t, int idx, int level, int *cap_lev, int *cap_leaves)
{
    if (t->num_children[idx] == 0) {
        if (t->leaves_list_num >= *cap_leaves) {
            *cap_leaves *= 2;
            t->leaves_list = realloc(t->leaves_list, (size_t)*cap_leaves * sizeof(int));
        }
        t->leaves_list[t->leaves_list_num++] = idx;
        return;
    }
    if (t->levels_list_num[level] >= cap_lev[level]) {
        cap_lev[level] *= 2;
        t->levels_list[level] = realloc(t->levels_list[level], (size_t)cap_lev[level] * sizeof(int));
    }
    t->levels_list[level][t->levels_list_num[level]++] = idx;
    for (int k = 0; k < t->num_children[idx]; k++)
        h_levels_rec(t, t->children[8 * idx + k], level + 1, cap_lev, cap_leaves);
}

static void h_free_ll(struct TreeLinkedListNode *p)
{
    for (int k = 0; k < p->num_children; k++) h_free_ll(p->child[k]);
    free(p);
}

static struct Tree *h_tree_construct(struct Particles *p, int maxpar, int is_source)
{
    struct TreeLinkedListNode *ll = NULL;
    double mm[6];
    int numnodes = 0, numleaves = 0, max_depth = 1, min_leaf = INT_MAX, max_leaf = 0;
    mm[0] = minval(p->x, p->num); mm[1] = maxval(p->x, p->num);
    mm[2] = minval(p->y, p->num); mm[3] = maxval(p->y, p->num);
    mm[4] = minval(p->z, p->num); mm[5] = maxval(p->z, p->num);
    if (is_source)
        TreeLinkedList_Sources_Construct(&ll, NULL, p, 1, p->num, maxpar, mm, &numnodes, &numleaves,
                                         &min_leaf, &max_leaf, &max_depth, 0);
    else
        TreeLinkedList_Targets_Construct(&ll, NULL, p, 1, p->num, maxpar, mm, &numnodes, &numleaves,
                                         &min_leaf, &max_leaf, &max_depth, 0);
    TreeLinkedList_SetIndex(ll, 0);
    struct Tree *t = h_tree_alloc(numnodes);
    h_tree_fill(t, ll);
    t->numleaves = numleaves; t->max_depth = max_depth;
    t->min_leaf_size = min_leaf; t->max_leaf_size = max_leaf;
    t->levels_list = calloc((size_t)max_depth + 1, sizeof(int *));
    t->levels_list_num = ivec(max_depth);
    int *cap_lev = ivec(max_depth);
    for (int i = 0; i < max_depth; i++) { cap_lev[i] = 8; t->levels_list[i] = ivec(8); }
    int cap_leaves = 8;
    t->leaves_list = ivec(8);
    t->leaves_list_num = 0;
    h_levels_rec(t, 0, 0, cap_lev, &cap_leaves);
    free(cap_lev);
    h_free_ll(ll);
    return t;
}

/* batches.c:197-412, with the reference's own partition functions */
static void h_batches_rec(struct Tree *b, int *cap, struct Particles *p, int ibeg, int iend,
                          int maxpar, int is_source)
{
    int ind[8][2];
    double xyzmms[6][8];
    memset(ind, 0, sizeof(ind));
    memset(xyzmms, 0, sizeof(xyzmms));
    int numpar = iend - ibeg + 1;
    double x_min = minval(p->x + ibeg - 1, numpar), x_max = maxval(p->x + ibeg - 1, numpar);
    double y_min = minval(p->y + ibeg - 1, numpar), y_max = maxval(p->y + ibeg - 1, numpar);
    double z_min = minval(p->z + ibeg - 1, numpar), z_max = maxval(p->z + ibeg - 1, numpar);
    double xl = x_max - x_min, yl = y_max - y_min, zl = z_max - z_min;
    double x_mid = (x_max + x_min) / 2.0, y_mid = (y_max + y_min) / 2.0, z_mid = (z_max + z_min) / 2.0;
    double radius = sqrt(xl * xl + yl * yl + zl * zl) / 2.0;
    if (numpar > maxpar) {
        int maxch = numpar < 2 * maxpar ? 2 : (numpar < 4 * maxpar ? 4 : 8);
        int nposs;
        ind[0][0] = ibeg; ind[0][1] = iend;
        xyzmms[0][0] = x_min; xyzmms[1][0] = x_max; xyzmms[2][0] = y_min;
        xyzmms[3][0] = y_max; xyzmms[4][0] = z_min; xyzmms[5][0] = z_max;
        if (is_source)
            pc_partition_8(p->x, p->y, p->z, p->q, p->w, p->order, xyzmms, xl, yl, zl, &nposs, maxch,
                           x_mid, y_mid, z_mid, ind);
        else
            cp_partition_8(p->x, p->y, p->z, p->q, p->order, xyzmms, xl, yl, zl, &nposs, maxch,
                           x_mid, y_mid, z_mid, ind);
        for (int i = 0; i < nposs; i++)
            if (ind[i][0] <= ind[i][1]) h_batches_rec(b, cap, p, ind[i][0], ind[i][1], maxpar, is_source);
        return;
    }
    if (b->numnodes >= *cap) {
        *cap *= 2;
        b->ibeg = realloc(b->ibeg, (size_t)*cap * sizeof(int));
        b->iend = realloc(b->iend, (size_t)*cap * sizeof(int));
        b->numpar = realloc(b->numpar, (size_t)*cap * sizeof(int));
        b->x_mid = realloc(b->x_mid, (size_t)*cap * sizeof(double));
        b->y_mid = realloc(b->y_mid, (size_t)*cap * sizeof(double));
        b->z_mid = realloc(b->z_mid, (size_t)*cap * sizeof(double));
        b->radius = realloc(b->radius, (size_t)*cap * sizeof(double));
    }
    int k = b->numnodes++;
    b->ibeg[k] = ibeg; b->iend[k] = iend; b->numpar[k] = numpar;
    b->x_mid[k] = x_mid; b->y_mid[k] = y_mid; b->z_mid[k] = z_mid; b->radius[k] = radius;
}

static struct Tree *h_batches(struct Particles *p, int maxpar, int is_source)
{
    struct Tree *b = calloc(1, sizeof(struct Tree));
    int cap = 16;
    b->ibeg = ivec(cap); b->iend = ivec(cap); b->numpar = ivec(cap);
    b->x_mid = dvec(cap); b->y_mid = dvec(cap); b->z_mid = dvec(cap); b->radius = dvec(cap);
    h_batches_rec(b, &cap, p, 1, p->num, maxpar, is_source);
    return b;
}

/* struct Tree -> oracle Tree view (for the oracle's upward pass) */
static void h_view(Tree *o, const struct Tree *t)
{
    memset(o, 0, sizeof(*o));
    o->n = t->numnodes;
    o->ibeg = t->ibeg; o->iend = t->iend; o->numpar = t->numpar; o->nchild = t->num_children;
    o->child = t->children;
    o->xmin = t->x_min; o->xmax = t->x_max; o->ymin = t->y_min; o->ymax = t->y_max;
    o->zmin = t->z_min; o->zmax = t->z_max; o->xmid = t->x_mid; o->ymid = t->y_mid; o->zmid = t->z_mid;
    o->radius = t->radius; o->max_depth = t->max_depth; o->numleaves = t->numleaves;
    o->leaves = t->leaves_list; o->nleaves = t->leaves_list_num;
    o->lev = t->levels_list; o->lev_n = t->levels_list_num;
}

static struct Clusters *h_clusters(const struct Tree *t, struct Particles *p, const Prm *prm, int sources)
{
    Tree o;
    h_view(&o, t);
    Clus c;
    memset(&c, 0, sizeof(c));
    int herm = prm->approximation == APX_HERMITE;
    if (sources) {
        Parts s;
        s.n = p->num; s.x = p->x; s.y = p->y; s.z = p->z; s.q = p->q; s.w = p->w; s.order = p->order;
        if (herm) clusters_sources_hermite(&c, &o, &s, prm);
        else clusters_sources(&c, &o, &s, prm);
    } else {
        clusters_targets(&c, &o, prm);
    }
    struct Clusters *r = calloc(1, sizeof(struct Clusters));
    r->num = t->numnodes * prm->P;
    r->num_charges = r->num * (herm ? 8 : 1);
    r->num_weights = r->num * (herm && prm->singularity == SNG_SUB ? 8 : 1);
    r->x = c.x; r->y = c.y; r->z = c.z; r->q = c.q; r->w = c.w;
    return r;
}

static struct Particles *h_parts(int n, const double *x, const double *y, const double *z,
                                 const double *q, const double *w)
{
    struct Particles *p = calloc(1, sizeof(struct Particles));
    p->num = n;
    p->x = dvec(n); p->y = dvec(n); p->z = dvec(n); p->q = dvec(n); p->w = dvec(n);
    p->order = ivec(n);
    for (int i = 0; i < n; i++) {
        p->x[i] = x[i]; p->y[i] = y[i]; p->z[i] = z[i];
        p->q[i] = q ? q[i] : 0.0; p->w[i] = w ? w[i] : 1.0; p->order[i] = i + 1;
    }
    return p;
}

/* Phase times of the last ref_treecode call (seconds) and its point-wise
 * evaluation counts: the computed ones (a sample when a stride is set) and
 * those of the full lists.  With stride k > 1 only every k-th key (target
 * batch for PC, source batch for CP, target node for CC) keeps its
 * interaction lists, so the compute phase runs a sample of the treecode's
 * evaluations; the trees, clusters, lists and downpass run in full. */
typedef struct {
    double t_setup, t_upward, t_lists, t_compute, t_downpass, t_finish;
    long long near_done, far_done, near_total, far_total;
} ref_timing;
static __thread ref_timing g_rt;
static __thread int g_ref_stride = 1;

void ref_set_sample_stride(int stride) { g_ref_stride = stride > 0 ? stride : 1; }
void ref_last_timing(ref_timing *t) { *t = g_rt; }

/* point-wise evaluations of key i's lists (treedriver.c:191-200, 472-483,
 * 763-792), and the sampling cut of the lists */
static void h_count_and_sample(struct InteractionLists *L, int nkeys, const struct Tree *keys,
                               const struct Tree *ent, int P, int ct)
{
    for (int i = 0; i < nkeys; i++) {
        long long nn = 0, nf = 0;
        const long long nk = keys->numpar[i];
        for (int j = 0; j < L->num_pp[i]; j++) nn += nk * ent->numpar[L->pp_interactions[i][j]];
        /* PC lists only num_pc, CP only num_cp (keys: source batches), CC all */
        if (ct == CT_CC) {
            nf += (long long)L->num_cc[i] * P * P;
            nf += (long long)L->num_pc[i] * nk * P;
            for (int j = 0; j < L->num_cp[i]; j++) nf += (long long)P * ent->numpar[L->cp_interactions[i][j]];
        } else if (ct == CT_PC) {
            nf += (long long)L->num_pc[i] * nk * P;
        } else {
            nf += (long long)L->num_cp[i] * nk * P;
        }
        g_rt.near_total += nn; g_rt.far_total += nf;
        if (i % g_ref_stride) {
            L->num_pp[i] = 0;
            if (L->num_cc) L->num_cc[i] = 0;
            if (L->num_pc) L->num_pc[i] = 0;
            if (L->num_cp) L->num_cp[i] = 0;
        } else {
            g_rt.near_done += nn; g_rt.far_done += nf;
        }
    }
}

int ref_treecode(int M, int N,
                 const double *tx, const double *ty, const double *tz, const double *tq,
                 const double *sx, const double *sy, const double *sz, const double *sq,
                 const double *sw, double *out,
                 int kernel, int nkp, const double *kp,
                 int singularity, int approximation, int compute_type,
                 double theta, int degree, int msl, int mtl,
                 double size_check, double beta, int verbosity)
{
    struct RunParams *rp = NULL;
    RunParams_Setup(&rp, (KERNEL)kernel, nkp, (double *)kp, (APPROXIMATION)approximation,
                    (SINGULARITY)singularity, (COMPUTE_TYPE)compute_type, theta, degree, msl, mtl,
                    size_check, beta, 0);
    RunParams_Validate(rp);
    Prm prm;
    memset(&prm, 0, sizeof(prm));
    prm.kernel = kernel; prm.singularity = singularity; prm.approximation = approximation;
    prm.compute_type = compute_type;
    for (int i = 0; i < nkp && i < 4; i++) prm.kp[i] = kp[i];
    prm.theta = rp->theta; prm.degree = rp->interp_degree; prm.P = rp->interp_pts_per_cluster;
    prm.size_check_factor = rp->size_check_factor; prm.msl = rp->max_per_source_leaf;
    prm.mtl = rp->max_per_target_leaf;
    if (!supported(&prm)) return ORC_ERR_UNSUPPORTED;

    struct Particles *S = h_parts(N, sx, sy, sz, sq, sw);
    struct Particles *T = h_parts(M, tx, ty, tz, tq, NULL);
    double *pot = dvec(M);
    struct InteractionLists *L = NULL;

    memset(&g_rt, 0, sizeof(g_rt));
    double t0 = wall();
    /* the phases of treedriver.c:114-967 in its order */
    if (compute_type == CT_PC) {
        struct Tree *tree = h_tree_construct(S, rp->max_per_source_leaf, 1);
        struct Tree *batches = h_batches(T, rp->max_per_target_leaf, 0);
        g_rt.t_setup = wall() - t0; t0 = wall();
        struct Clusters *cl = h_clusters(tree, S, &prm, 1);
        g_rt.t_upward = wall() - t0; t0 = wall();
        InteractionLists_Make(&L, tree, batches, rp);
        h_count_and_sample(L, batches->numnodes, batches, tree, prm.P, compute_type);
        g_rt.t_lists = wall() - t0; t0 = wall();
        InteractionCompute_PC(pot, tree, batches, L, S, T, cl, rp);
        g_rt.t_compute = wall() - t0; t0 = wall();
    } else if (compute_type == CT_CP) {
        struct Tree *tree = h_tree_construct(T, rp->max_per_target_leaf, 0);
        struct Tree *batches = h_batches(S, rp->max_per_source_leaf, 1);
        struct Clusters *cl = h_clusters(tree, T, &prm, 0);
        g_rt.t_setup = wall() - t0; t0 = wall();
        InteractionLists_Make(&L, tree, batches, rp);
        h_count_and_sample(L, batches->numnodes, batches, tree, prm.P, compute_type);
        g_rt.t_lists = wall() - t0; t0 = wall();
        InteractionCompute_CP(pot, tree, batches, L, S, T, cl, rp);
        g_rt.t_compute = wall() - t0; t0 = wall();
        InteractionCompute_Downpass(pot, tree, T, cl, rp);
        g_rt.t_downpass = wall() - t0; t0 = wall();
    } else {
        struct Tree *stree = h_tree_construct(S, rp->max_per_source_leaf, 1);
        struct Tree *ttree = h_tree_construct(T, rp->max_per_target_leaf, 0);
        g_rt.t_setup = wall() - t0; t0 = wall();
        struct Clusters *scl = h_clusters(stree, S, &prm, 1);
        struct Clusters *tcl = h_clusters(ttree, T, &prm, 0);
        g_rt.t_upward = wall() - t0; t0 = wall();
        InteractionLists_Make(&L, stree, ttree, rp);
        h_count_and_sample(L, ttree->numnodes, ttree, stree, prm.P, compute_type);
        g_rt.t_lists = wall() - t0; t0 = wall();
        InteractionCompute_CC(pot, stree, ttree, L, S, T, scl, tcl, rp);
        g_rt.t_compute = wall() - t0; t0 = wall();
        InteractionCompute_Downpass(pot, ttree, T, tcl, rp);
        g_rt.t_downpass = wall() - t0; t0 = wall();
    }
    InteractionCompute_SubtractionPotentialCorrection(pot, T, rp);
    for (int i = 0; i < M; i++) out[T->order[i] - 1] = pot[i];
    g_rt.t_finish = wall() - t0;
    /* Test harness: the small allocations above are left to process exit. */
    return 0;
}
