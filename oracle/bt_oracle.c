/*
 * bt_oracle.c -- CPU restatement of BaryTree's FP64 interaction path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product): only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load this file's
 * library.  Compile with -ffp-contract=off so the arithmetic is the
 * reference's arithmetic without fused multiply-adds.
 *
 * Scope: Lagrange and Hermite approximations; PC, CP and CC compute types; the Coulomb,
 * Yukawa and regularized-Coulomb kernels with skipping and subtraction, in
 * exactly the combinations the reference dispatchers accept
 * (src/interaction_compute/interaction_compute_{pc,cp,cc}.c), and the
 * Hermite approximation where those dispatchers accept it (PC with either
 * singularity, CP with skipping).
 *
 * Every function cites the reference lines it follows.  Deviations, each
 * deliberate:
 *   - caller arrays are copied, never permuted in place;
 *   - target-cluster charges start at zero (the reference's CPU path leaves
 *     them uninitialised, src/clusters/clusters.c:245-253 allocates without
 *     clearing; its OpenACC path clears them, :258-271);
 *   - a NULL source weight is treated as 1 for skipping (the reference uses
 *     an uninitialised dummy that the skipping path never reads);
 *   - errors return a negative code instead of exit(1).
 */
#define _GNU_SOURCE
#include <float.h>
#include <limits.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "bt_oracle.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

enum { KER_NONE = 0, KER_COULOMB = 1, KER_YUKAWA = 2, KER_REGCOULOMB = 3, KER_REGYUKAWA = 4, KER_ATAN = 5,
       KER_SINOVERR = 8, KER_MQ = 9, KER_RBSU = 10, KER_RBSV = 11, KER_USER = 12 };
enum { SNG_NONE = 0, SNG_SKIP = 1, SNG_SUB = 2 };
enum { APX_NONE = 0, APX_LAGRANGE = 1, APX_HERMITE = 2 };
enum { CT_NONE = 0, CT_PC = 1, CT_CP = 2, CT_CC = 3 };

#define ORC_ERR_ARGS (-1)
#define ORC_ERR_UNSUPPORTED (-2)
#define ORC_ERR_NOMEM (-3)

static __thread orc_stats g_stats;
static __thread int g_stride = 1;

static double wall(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* ------------------------------------------------------------------ */
/* Parameters: src/run_params/run_params.c:13-131                       */
/* ------------------------------------------------------------------ */

typedef struct {
    int kernel, singularity, approximation, compute_type;
    double kp[4];
    double theta, size_check_factor;
    int degree, P;
    int msl, mtl;
} Prm;

void orc_params_resolve(int compute_type, int approximation, double theta, int degree,
                        int msl, int mtl, double size_check, double beta,
                        double *theta_out, int *degree_out, double *size_check_out,
                        int *msl_out, int *mtl_out)
{
    if (beta < 0 || beta > 1) {
        *theta_out = theta;
        *degree_out = degree;
        *size_check_out = size_check;
        *msl_out = msl;
        *mtl_out = mtl;
        return;
    }
    double tmin = 0.55, tmax = 0.95, nmin = 1, nmax = 12, es = 2, et = 3, scf = 1.0;
    if (compute_type == CT_PC || compute_type == CT_CP) {
        if (approximation != APX_LAGRANGE) { scf = 4.0; nmax = 9; }
    } else {
        es = 1.75;
    }
    *size_check_out = scf;
    *theta_out = tmax - (tmax - tmin) * pow(beta, es);
    *degree_out = (int)(nmax - (nmax - nmin) * pow(1. - beta, et));
    /* CPU-build leaf sizes (run_params.c:107-119) */
    if (compute_type == CT_PC) { *msl_out = 50; *mtl_out = 5; }
    else if (compute_type == CT_CP) { *msl_out = 5; *mtl_out = 50; }
    else { *msl_out = 50; *mtl_out = 50; }
}

/* ------------------------------------------------------------------ */
/* Particles: SoA copies with the 1-based order permutation            */
/* (src/particles/particles.c:58-144)                                  */
/* ------------------------------------------------------------------ */

typedef struct {
    int n;
    double *x, *y, *z, *q, *w; /* w may be NULL (targets) */
    int *order;
} Parts;

static int parts_copy(Parts *p, int n, const double *x, const double *y, const double *z,
                      const double *q, const double *w, int want_w)
{
    size_t b = (size_t)(n > 0 ? n : 1) * sizeof(double);
    p->n = n;
    p->x = malloc(b); p->y = malloc(b); p->z = malloc(b); p->q = malloc(b);
    p->w = want_w ? malloc(b) : NULL;
    p->order = malloc((size_t)(n > 0 ? n : 1) * sizeof(int));
    if (!p->x || !p->y || !p->z || !p->q || (want_w && !p->w) || !p->order) return ORC_ERR_NOMEM;
    for (int i = 0; i < n; i++) {
        p->x[i] = x[i]; p->y[i] = y[i]; p->z[i] = z[i];
        p->q[i] = q ? q[i] : 0.0;
        if (want_w) p->w[i] = w ? w[i] : 1.0;
        p->order[i] = i + 1;
    }
    return 0;
}

static void parts_free(Parts *p)
{
    free(p->x); free(p->y); free(p->z); free(p->q); free(p->w); free(p->order);
    memset(p, 0, sizeof(*p));
}

/* ------------------------------------------------------------------ */
/* Partition: src/tree/partition.c:14-175 (in-place, pivot hole).       */
/* Key dimension d; all arrays follow the same permutation.  Returns   */
/* the 1-based index of the last element of the lower part.            */
/* ------------------------------------------------------------------ */

static double *key_of(Parts *p, int d) { return d == 0 ? p->x : (d == 1 ? p->y : p->z); }

static void move_el(Parts *p, int dst, int src)
{
    p->x[dst] = p->x[src]; p->y[dst] = p->y[src]; p->z[dst] = p->z[src];
    p->q[dst] = p->q[src];
    if (p->w) p->w[dst] = p->w[src];
    p->order[dst] = p->order[src];
}

static int split_range(Parts *p, int d, int ibeg, int iend, double val)
{
    double *a = key_of(p, d);
    if (ibeg < iend) {
        double hx = p->x[ibeg - 1], hy = p->y[ibeg - 1], hz = p->z[ibeg - 1];
        double hq = p->q[ibeg - 1], hw = p->w ? p->w[ibeg - 1] : 0.0;
        int ho = p->order[ibeg - 1];
        double hkey = a[ibeg - 1];
        a[ibeg - 1] = val;
        int up = ibeg, lo = iend;
        while (up != lo) {
            while (up < lo && val < a[lo - 1]) lo--;
            if (up != lo) move_el(p, up - 1, lo - 1);
            while (up < lo && val >= a[up - 1]) up++;
            if (up != lo) move_el(p, lo - 1, up - 1);
        }
        int mid = up;
        if (hkey > val) mid = up - 1;
        p->x[up - 1] = hx; p->y[up - 1] = hy; p->z[up - 1] = hz; p->q[up - 1] = hq;
        if (p->w) p->w[up - 1] = hw;
        p->order[up - 1] = ho;
        return mid;
    } else if (ibeg == iend) {
        return a[ibeg - 1] < val ? ibeg : ibeg - 1;
    }
    return ibeg - 1;
}

static double max3d(double a, double b, double c)
{
    double m = a;
    if (m < b) m = b;
    if (m < c) m = c;
    return m;
}

/* src/tree/partition.c:178-286 (pc_partition_8 / cp_partition_8) */
static int split_8(Parts *p, double xl, double yl, double zl, int maxch,
                   double xm, double ym, double zm, int ind[8][2])
{
    int n = 1;
    double crit = max3d(xl, yl, zl) / sqrt(2.0);
    int dx = xl >= crit, dy = yl >= crit, dz = zl >= crit;
    if (maxch == 4) {
        if (xl < yl && xl < zl) dx = 0;
        if (yl < xl && yl < zl) dy = 0;
        if (zl < xl && zl < yl) dz = 0;
        if (dx + dy + dz == 3) dx = 0;
    }
    if (maxch == 2) {
        if (xl < yl || xl < zl) dx = 0;
        if (yl < xl || yl < zl) dy = 0;
        if (zl < xl || zl < yl) dz = 0;
        if (dx + dy + dz == 3) { dx = 0; dy = 0; }
        if (dx + dy + dz == 2) {
            if (dx == 0) dy = 0;
            else if (dy == 0) dz = 0;
            else if (dz == 0) dx = 0;
        }
    }
    const int div[3] = {dx, dy, dz};
    const double mid[3] = {xm, ym, zm};
    for (int d = 0; d < 3; d++) {
        if (!div[d]) continue;
        for (int i = 0; i < n; i++) {
            int t = split_range(p, d, ind[i][0], ind[i][1], mid[d]);
            ind[n + i][0] = t + 1;
            ind[n + i][1] = ind[i][1];
            ind[i][1] = t;
        }
        n *= 2;
    }
    return n;
}

/* ------------------------------------------------------------------ */
/* Tree: src/tree/tree_linked_list.c:16-280, tree.c:19-158, 235-268    */
/* Nodes are stored in creation order, which is the preorder that      */
/* TreeLinkedList_SetIndex assigns.                                    */
/* ------------------------------------------------------------------ */

typedef struct {
    int n, cap;
    int *ibeg, *iend, *numpar, *nchild, *child, *level;
    double *xmin, *xmax, *ymin, *ymax, *zmin, *zmax, *xmid, *ymid, *zmid, *radius;
    int max_depth, numleaves;
    int *leaves, nleaves;
    int **lev, *lev_n, *lev_cap;
    int leaves_only; /* batches: keep only leaves */
} Tree;

static int tree_grow(Tree *t)
{
    if (t->n < t->cap) return 0;
    int nc = t->cap ? t->cap * 2 : 64;
#define GROW(f, T) do { void *q_ = realloc(t->f, (size_t)nc * sizeof(T)); if (!q_) return ORC_ERR_NOMEM; t->f = q_; } while (0)
    GROW(ibeg, int); GROW(iend, int); GROW(numpar, int); GROW(nchild, int); GROW(level, int);
    { void *q_ = realloc(t->child, (size_t)nc * 8 * sizeof(int)); if (!q_) return ORC_ERR_NOMEM; t->child = q_; }
    GROW(xmin, double); GROW(xmax, double); GROW(ymin, double); GROW(ymax, double);
    GROW(zmin, double); GROW(zmax, double); GROW(xmid, double); GROW(ymid, double);
    GROW(zmid, double); GROW(radius, double);
#undef GROW
    t->cap = nc;
    return 0;
}

static void tree_free(Tree *t)
{
    free(t->ibeg); free(t->iend); free(t->numpar); free(t->nchild); free(t->child); free(t->level);
    free(t->xmin); free(t->xmax); free(t->ymin); free(t->ymax); free(t->zmin); free(t->zmax);
    free(t->xmid); free(t->ymid); free(t->zmid); free(t->radius); free(t->leaves);
    if (t->lev) { for (int i = 0; i < t->max_depth; i++) free(t->lev[i]); }
    free(t->lev); free(t->lev_n); free(t->lev_cap);
    memset(t, 0, sizeof(*t));
}

static double minv(const double *x, int n) { double m = x[0]; for (int i = 1; i < n; ++i) if (m > x[i]) m = x[i]; return m; }
static double maxv(const double *x, int n) { double m = x[0]; for (int i = 1; i < n; ++i) if (m < x[i]) m = x[i]; return m; }

/* One node (or, for batches, possibly a leaf record): returns node index
 * (>= 0), -1 when a batch interior node is not recorded, or an error. */
static int tree_build_rec(Tree *t, Parts *p, int ibeg, int iend, int maxpar, int level)
{
    int numpar = iend - ibeg + 1;
    double x_min = minv(p->x + ibeg - 1, numpar), x_max = maxv(p->x + ibeg - 1, numpar);
    double y_min = minv(p->y + ibeg - 1, numpar), y_max = maxv(p->y + ibeg - 1, numpar);
    double z_min = minv(p->z + ibeg - 1, numpar), z_max = maxv(p->z + ibeg - 1, numpar);
    double xl = x_max - x_min, yl = y_max - y_min, zl = z_max - z_min;
    double x_mid = (x_max + x_min) / 2.0, y_mid = (y_max + y_min) / 2.0, z_mid = (z_max + z_min) / 2.0;
    double radius = sqrt(xl * xl + yl * yl + zl * zl) / 2.0;
    int is_leaf = !(numpar > maxpar);
    if (level + 1 > t->max_depth) t->max_depth = level + 1;

    int idx = -1;
    if (!t->leaves_only || is_leaf) {
        int e = tree_grow(t);
        if (e) return e;
        idx = t->n++;
        t->ibeg[idx] = ibeg; t->iend[idx] = iend; t->numpar[idx] = numpar;
        t->xmin[idx] = x_min; t->xmax[idx] = x_max; t->ymin[idx] = y_min;
        t->ymax[idx] = y_max; t->zmin[idx] = z_min; t->zmax[idx] = z_max;
        t->xmid[idx] = x_mid; t->ymid[idx] = y_mid; t->zmid[idx] = z_mid;
        t->radius[idx] = radius; t->nchild[idx] = 0; t->level[idx] = level;
    }
    if (!is_leaf) {
        int maxch = numpar < 2 * maxpar ? 2 : (numpar < 4 * maxpar ? 4 : 8);
        int ind[8][2];
        memset(ind, 0, sizeof(ind));
        ind[0][0] = ibeg; ind[0][1] = iend;
        int nposs = split_8(p, xl, yl, zl, maxch, x_mid, y_mid, z_mid, ind);
        for (int i = 0; i < nposs; i++) {
            if (ind[i][0] <= ind[i][1]) {
                int c = tree_build_rec(t, p, ind[i][0], ind[i][1], maxpar, level + 1);
                if (c < -1) return c;
                if (idx >= 0) t->child[8 * idx + t->nchild[idx]++] = c;
            }
        }
    } else {
        t->numleaves++;
    }
    return idx;
}

static int push_int(int **arr, int *n, int *cap, int v)
{
    if (*n >= *cap) {
        int nc = *cap ? *cap * 2 : 16;
        int *q = realloc(*arr, (size_t)nc * sizeof(int));
        if (!q) return ORC_ERR_NOMEM;
        *arr = q; *cap = nc;
    }
    (*arr)[(*n)++] = v;
    return 0;
}

/* src/tree/tree.c:106-158 */
static int tree_levels_rec(Tree *t, int idx, int level, int *leaves_cap)
{
    if (t->nchild[idx] == 0) return push_int(&t->leaves, &t->nleaves, leaves_cap, idx);
    int e = push_int(&t->lev[level], &t->lev_n[level], &t->lev_cap[level], idx);
    if (e) return e;
    for (int i = 0; i < t->nchild[idx]; i++) {
        e = tree_levels_rec(t, t->child[8 * idx + i], level + 1, leaves_cap);
        if (e) return e;
    }
    return 0;
}

static int tree_construct(Tree *t, Parts *p, int maxpar, int leaves_only)
{
    memset(t, 0, sizeof(*t));
    t->leaves_only = leaves_only;
    if (p->n <= 0) return ORC_ERR_ARGS;
    int r = tree_build_rec(t, p, 1, p->n, maxpar, 0);
    if (r < -1) return r;
    if (leaves_only) return 0;
    t->lev = calloc((size_t)t->max_depth, sizeof(int *));
    t->lev_n = calloc((size_t)t->max_depth, sizeof(int));
    t->lev_cap = calloc((size_t)t->max_depth, sizeof(int));
    if (!t->lev || !t->lev_n || !t->lev_cap) return ORC_ERR_NOMEM;
    int leaves_cap = 0;
    return tree_levels_rec(t, 0, 0, &leaves_cap);
}

/* ------------------------------------------------------------------ */
/* Interaction lists: src/interaction_lists/interaction_lists.c:48-234, */
/* 464-748.  Jagged lists keyed by the "target" side.                  */
/* ------------------------------------------------------------------ */

typedef struct { int *v; int n, cap; } IList;

static int il_push(IList *l, int v) { return push_int(&l->v, &l->n, &l->cap, v); }

typedef struct {
    int n;             /* number of keys (target nodes / batches) */
    IList *pp, *ap;    /* direct, approx (pc / cp / cc)             */
    IList *spc, *tcp;  /* CC only: source-approx (PC), target-approx (CP) */
} Lists;

static void lists_free(Lists *L)
{
    for (int i = 0; i < L->n; i++) {
        free(L->pp[i].v); free(L->ap[i].v);
        if (L->spc) free(L->spc[i].v);
        if (L->tcp) free(L->tcp[i].v);
    }
    free(L->pp); free(L->ap); free(L->spc); free(L->tcp);
    memset(L, 0, sizeof(*L));
}

/* interaction_lists.c:464-533 */
static int pc_list_rec(const Tree *tr, int node, double br, double bx, double by, double bz,
                       IList *ap, IList *pp, const Prm *prm)
{
    double tx = bx - tr->xmid[node];
    double ty = by - tr->ymid[node];
    double tz = bz - tr->zmid[node];
    double dist = sqrt(tx * tx + ty * ty + tz * tz);
    if (((tr->radius[node] + br) < dist * prm->theta) && (tr->radius[node] != 0.00)
        && (prm->size_check_factor * prm->P < tr->numpar[node])) {
        return il_push(ap, node);
    }
    if (tr->nchild[node] == 0) return il_push(pp, node);
    for (int i = 0; i < tr->nchild[node]; i++) {
        int e = pc_list_rec(tr, tr->child[8 * node + i], br, bx, by, bz, ap, pp, prm);
        if (e) return e;
    }
    return 0;
}

/* interaction_lists.c:537-748 */
static int cc_list_rec(const Tree *st, int sn, const Tree *tt, int tn, Lists *L, const Prm *prm)
{
    int size_check = (int)(prm->size_check_factor * prm->P);
    double tx = tt->xmid[tn] - st->xmid[sn];
    double ty = tt->ymid[tn] - st->ymid[sn];
    double tz = tt->zmid[tn] - st->zmid[sn];
    double dist = sqrt(tx * tx + ty * ty + tz * tz);
    if ((st->radius[sn] + tt->radius[tn]) < dist * prm->theta) {
        if (st->numpar[sn] <= size_check && tt->numpar[tn] <= size_check) return il_push(&L->pp[tn], sn);
        if (st->numpar[sn] <= size_check) return il_push(&L->tcp[tn], sn);
        if (tt->numpar[tn] <= size_check) return il_push(&L->spc[tn], sn);
        return il_push(&L->ap[tn], sn);
    }
    int e = 0;
    if (tt->nchild[tn] == 0 && st->nchild[sn] == 0) return il_push(&L->pp[tn], sn);
    if (st->nchild[sn] == 0 || (tt->nchild[tn] != 0 && st->numpar[sn] < tt->numpar[tn])) {
        for (int i = 0; i < tt->nchild[tn] && !e; i++)
            e = cc_list_rec(st, sn, tt, tt->child[8 * tn + i], L, prm);
    } else {
        for (int i = 0; i < st->nchild[sn] && !e; i++)
            e = cc_list_rec(st, st->child[8 * sn + i], tt, tn, L, prm);
    }
    return e;
}

static int lists_make(Lists *L, const Tree *src, const Tree *tgt, const Prm *prm)
{
    memset(L, 0, sizeof(*L));
    L->n = tgt->n;
    L->pp = calloc((size_t)L->n, sizeof(IList));
    L->ap = calloc((size_t)L->n, sizeof(IList));
    if (!L->pp || !L->ap) return ORC_ERR_NOMEM;
    if (prm->compute_type == CT_CC) {
        L->spc = calloc((size_t)L->n, sizeof(IList));
        L->tcp = calloc((size_t)L->n, sizeof(IList));
        if (!L->spc || !L->tcp) return ORC_ERR_NOMEM;
        return cc_list_rec(src, 0, tgt, 0, L, prm);
    }
    for (int i = 0; i < tgt->n; i++) {
        int e = pc_list_rec(src, 0, tgt->radius[i], tgt->xmid[i], tgt->ymid[i], tgt->zmid[i],
                            &L->ap[i], &L->pp[i], prm);
        if (e) return e;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* Chebyshev nodes and barycentric weights: clusters.c:436-457          */
/* ------------------------------------------------------------------ */

static void cheb_nodes(int n, double a0, double a1, double *node)
{
    for (int i = 0; i < n + 1; i++) {
        double tt = cos(i * M_PI / n);
        node[i] = a0 + (tt + 1.0) / 2.0 * (a1 - a0);
    }
}

static void bary_weights(int n, double *wts)
{
    for (int j = 0; j < n + 1; j++) {
        double dj = 1.0;
        if (j == 0) dj = 0.5;
        if (j == n) dj = 0.5;
        wts[j] = ((j % 2 == 0) ? 1 : -1) * dj;
    }
}

typedef struct {
    int numnodes, P;
    double *x, *y, *z, *q, *w; /* w NULL unless subtraction */
} Clus;

static void clus_free(Clus *c) { free(c->x); free(c->y); free(c->z); free(c->q); free(c->w); memset(c, 0, sizeof(*c)); }

/* qplanes (wplanes) values per grid point: 1 for Lagrange, 8 for Hermite
 * (run_params.c:122-128); wplanes 0: no weights */
static int clus_alloc_planes(Clus *c, int numnodes, int P, int qplanes, int wplanes)
{
    size_t n = (size_t)numnodes * (size_t)P;
    c->numnodes = numnodes; c->P = P;
    c->x = calloc(n, sizeof(double)); c->y = calloc(n, sizeof(double)); c->z = calloc(n, sizeof(double));
    c->q = calloc(n * qplanes, sizeof(double));
    c->w = wplanes ? calloc(n * wplanes, sizeof(double)) : NULL;
    if (!c->x || !c->y || !c->z || !c->q || (wplanes && !c->w)) return ORC_ERR_NOMEM;
    return 0;
}

static int clus_alloc(Clus *c, int numnodes, int P, int with_w)
{
    return clus_alloc_planes(c, numnodes, P, 1, with_w ? 1 : 0);
}

/*
 * Anterpolation of m "points" (particles, or a child's grid) onto the grid of
 * the box [x0,x1]x[y0,y1]x[z0,z1]: clusters.c:800-1004 (leaf, skipping),
 * :1008-1219 (leaf, subtraction), :379-582 and :585-795 (child to parent).
 * f1 holds the charges (q, q*w or child Q); f2 (optional) the second plane
 * (w or child W).  Grid coordinates are written to gx/gy/gz.
 */
static int anterp(int n, double x0, double x1, double y0, double y1, double z0, double z1,
                  int m, const double *px, const double *py, const double *pz,
                  const double *f1in, const double *f2in,
                  double *gx, double *gy, double *gz, double *Q, double *W)
{
    int nl = n + 1, P = nl * nl * nl;
    double nodeX[32], nodeY[32], nodeZ[32], wts[32];
    double *f1 = malloc((size_t)(m > 0 ? m : 1) * sizeof(double));
    double *f2 = f2in ? malloc((size_t)(m > 0 ? m : 1) * sizeof(double)) : NULL;
    int *ex = malloc((size_t)(m > 0 ? m : 1) * 3 * sizeof(int));
    if (!f1 || (f2in && !f2) || !ex) { free(f1); free(f2); free(ex); return ORC_ERR_NOMEM; }
    int *ey = ex + m, *ez = ex + 2 * m;
    for (int j = 0; j < m; j++) {
        f1[j] = f1in[j];
        if (f2) f2[j] = f2in[j];
        ex[j] = -1; ey[j] = -1; ez[j] = -1;
    }
    cheb_nodes(n, x0, x1, nodeX);
    cheb_nodes(n, y0, y1, nodeY);
    cheb_nodes(n, z0, z1, nodeZ);
    bary_weights(n, wts);

    for (int i = 0; i < m; i++) {
        double sumX = 0.0, sumY = 0.0, sumZ = 0.0;
        double sx = px[i], sy = py[i], sz = pz[i];
        for (int j = 0; j < nl; j++) {
            double cx = sx - nodeX[j], cy = sy - nodeY[j], cz = sz - nodeZ[j];
            if (fabs(cx) < DBL_MIN) ex[i] = j;
            if (fabs(cy) < DBL_MIN) ey[i] = j;
            if (fabs(cz) < DBL_MIN) ez[i] = j;
            double w = wts[j];
            sumX += w / cx;
            sumY += w / cy;
            sumZ += w / cz;
        }
        double denominator = 1.0;
        if (ex[i] == -1) denominator *= sumX;
        if (ey[i] == -1) denominator *= sumY;
        if (ez[i] == -1) denominator *= sumZ;
        f1[i] /= denominator;
        if (f2) f2[i] /= denominator;
    }

    for (int j = 0; j < P; j++) {
        int k1 = j % nl;
        int kk = (j - k1) / nl;
        int k2 = kk % nl;
        kk = kk - k2;
        int k3 = kk / nl;
        double cz = nodeZ[k3], w3 = wts[k3];
        double cy = nodeY[k2], w2 = wts[k2];
        double cx = nodeX[k1], w1 = wts[k1];
        gx[j] = cx; gy[j] = cy; gz[j] = cz;
        double temp = 0.0, temp2 = 0.0;
        for (int i = 0; i < m; i++) {
            double sx = px[i], sy = py[i], sz = pz[i];
            double numerator = 1.0;
            if (ex[i] == -1) numerator *= w1 / (sx - cx);
            else if (ex[i] != k1) numerator *= 0;
            if (ey[i] == -1) numerator *= w2 / (sy - cy);
            else if (ey[i] != k2) numerator *= 0;
            if (ez[i] == -1) numerator *= w3 / (sz - cz);
            else if (ez[i] != k3) numerator *= 0;
            temp += numerator * f1[i];
            if (f2) temp2 += numerator * f2[i];
        }
        Q[j] += temp;
        if (f2) W[j] += temp2;
    }
    free(f1); free(f2); free(ex);
    return 0;
}

/* clusters.c:50-200 (Lagrange branches) */
static int clusters_sources(Clus *c, const Tree *t, const Parts *s, const Prm *prm)
{
    int sub = prm->singularity == SNG_SUB;
    int P = prm->P, n = prm->degree;
    int e = clus_alloc(c, t->n, P, sub);
    if (e) return e;
    double *qw = NULL;
    if (sub) {
        qw = malloc((size_t)s->n * sizeof(double));
        if (!qw) return ORC_ERR_NOMEM;
        for (int i = 0; i < s->n; i++) qw[i] = s->q[i] * s->w[i];
    }
    /* leaves, then parents level by level: every node's grid is written by
       one thread, its children in order (errors: the least code) */
#pragma omp parallel for schedule(dynamic) reduction(min : e)
    for (int li = 0; li < t->nleaves; li++) {
        int idx = t->leaves[li];
        int b = t->ibeg[idx] - 1, m = t->iend[idx] - t->ibeg[idx] + 1;
        size_t cs = (size_t)idx * P;
        int r = anterp(n, t->xmin[idx], t->xmax[idx], t->ymin[idx], t->ymax[idx], t->zmin[idx], t->zmax[idx],
                       m, s->x + b, s->y + b, s->z + b, sub ? qw + b : s->q + b, sub ? s->w + b : NULL,
                       c->x + cs, c->y + cs, c->z + cs, c->q + cs, sub ? c->w + cs : NULL);
        if (r < e) e = r;
    }
    for (int level = t->max_depth - 1; level >= 0 && !e; --level) {
#pragma omp parallel for schedule(dynamic) reduction(min : e)
        for (int ci = 0; ci < t->lev_n[level]; ++ci) {
            int par = t->lev[level][ci];
            size_t ps = (size_t)par * P;
            for (int k = 0; k < t->nchild[par]; ++k) {
                int ch = t->child[8 * par + k];
                size_t cs = (size_t)ch * P;
                int r = anterp(n, t->xmin[par], t->xmax[par], t->ymin[par], t->ymax[par], t->zmin[par],
                               t->zmax[par], P, c->x + cs, c->y + cs, c->z + cs, c->q + cs, sub ? c->w + cs : NULL,
                               c->x + ps, c->y + ps, c->z + ps, c->q + ps, sub ? c->w + ps : NULL);
                if (r < e) e = r;
            }
        }
    }
    free(qw);
    return e;
}

/* clusters.c:205-288, 1840-1908: grid coordinates of every target node;
 * charges (and weights) start at zero. */
static int clusters_targets(Clus *c, const Tree *t, const Prm *prm)
{
    int P = prm->P, n = prm->degree, nl = n + 1;
    int planes = prm->approximation == APX_HERMITE ? 8 : 1;
    int e = clus_alloc_planes(c, t->n, P, planes, prm->singularity == SNG_SUB ? planes : 0);
    if (e) return e;
    double nodeX[32], nodeY[32], nodeZ[32];
    for (int idx = 0; idx < t->n; idx++) {
        cheb_nodes(n, t->xmin[idx], t->xmax[idx], nodeX);
        cheb_nodes(n, t->ymin[idx], t->ymax[idx], nodeY);
        cheb_nodes(n, t->zmin[idx], t->zmax[idx], nodeZ);
        size_t cs = (size_t)idx * P;
        for (int j = 0; j < P; j++) {
            int k1 = j % nl;
            int kk = (j - k1) / nl;
            int k2 = kk % nl;
            kk = kk - k2;
            int k3 = kk / nl;
            c->x[cs + j] = nodeX[k1]; c->y[cs + j] = nodeY[k2]; c->z[cs + j] = nodeZ[k3];
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* Kernels: the coulomb, yukawa and regularized-coulomb dirs of src/kernels */
/* ------------------------------------------------------------------ */

/* Target particles [ts, ts+nt) against source particles [ss, ss+ns):
 * K_*_PP and K_*_SS_PP (coulomb_pp.c:16-58, coulomb_ss_pp.c:18-57,
 * yukawa_pp.c:19-54, yukawa_ss_pp.c:19-57, regularized-coulomb_pp.c:18-58,
 * regularized-coulomb_ss_pp.c:19-56). */
/* ------------------------------------------------------------------ */
/* The kernels beside the Coulomb family, one pair's term as each of    */
/* their sources writes it (src/kernels/<name>/<name>_{pp,pc,cp}.c).    */
/* ATAN and MQ act on z alone (periodic in domainLength = kp[0], delta  */
/* = kp[1]); RBS_U and RBS_V on (x, y) (delta = kp[0]); SIN_OVER_R and  */
/* USER take k = kp[0]; REGULARIZED_YUKAWA (kappa, epsilon).            */
/* ------------------------------------------------------------------ */

static int other_kernel(int k) { return k == KER_REGYUKAWA || k == KER_ATAN || (k >= KER_SINOVERR && k <= KER_USER); }

/* periodic z offset (atan_pp.c:47-53, mq_pp.c:47-53) */
static double wrap_dz(double tz, double sz, double L)
{
    double dz = (tz - sz) / L;
    if (dz < -0.5) dz += 1.0;
    if (dz > 0.5) dz -= 1.0;
    return dz;
}

/* the factor ATAN applies to each kernel call's sum (atan_pp.c:26-28, 58) */
static double atan_wadj(const Prm *prm)
{
    double delta = prm->kp[1];
    return 1. / (1. - delta / sqrt(1. + delta * delta));
}

/* PP: regularized-yukawa_pp.c:40-43, regularized-yukawa_ss_pp.c:44-48,
 * atan_pp.c:44-55, sin-over-r_pp.c:40-46, mq_pp.c:44-55, rbs-u_pp.c:41-44,
 * rbs-v_pp.c:41-44, user_kernel_pp.c:40-46 */
static double other_pp(const Prm *prm, double tx, double ty, double tz, double tq, double sx, double sy,
                       double sz, double q, double w)
{
    double dx = tx - sx, dy = ty - sy, dz = tz - sz;
    switch (prm->kernel) {
    case KER_REGYUKAWA: {
        double kappa = prm->kp[0];
        double r = sqrt(dx * dx + dy * dy + dz * dz);
        if (prm->singularity == SNG_SUB) {
            double epsilon = prm->kp[1];
            return (q - tq) * w * exp(-kappa * r) / sqrt(r * r + epsilon * epsilon);
        }
        double epsilon2 = prm->kp[1] * prm->kp[1];
        return q * exp(-kappa * r) / sqrt(r * r + epsilon2);
    }
    case KER_ATAN: {
        double delta = prm->kp[1];
        double delta_factor = sqrt(1. + 1.0 / (delta * delta));
        double d = wrap_dz(tz, sz, prm->kp[0]);
        return q * (1.0 / M_PI * atan(delta_factor * tan(M_PI * d)) - d);
    }
    case KER_MQ: {
        double L = prm->kp[0], delta = prm->kp[1];
        double deltaLsq = delta * delta / L / L;
        double norm_delta_L = sqrt(1 + 4 * deltaLsq);
        double d = wrap_dz(tz, sz, L);
        return q * (.5 * d * norm_delta_L / sqrt(d * d + deltaLsq) - d);
    }
    case KER_SINOVERR: {
        double r = sqrt(dx * dx + dy * dy + dz * dz);
        return r > DBL_MIN ? q * sin(prm->kp[0] * r) / r : 0.0;
    }
    case KER_RBSU: {
        double r = dx * dx + dy * dy + prm->kp[0] * prm->kp[0];
        return 1. / (2. * M_PI) * q * dy / r;
    }
    case KER_RBSV: {
        double r = dx * dx + dy * dy + prm->kp[0] * prm->kp[0];
        return -1. / (2. * M_PI) * q * dx / r;
    }
    default: { /* KER_USER */
        double r = sqrt(dx * dx + dy * dy + dz * dz);
        return r > DBL_MIN ? q * exp(-prm->kp[0] * r) / r : 0.0;
    }
    }
}

/* PC: target against a cluster point with charge Q (and weight W):
 * regularized-yukawa_pc.c:39-44, regularized-yukawa_ss_pc.c:42-47,
 * atan_pc.c:44-55, sin-over-r_pc.c:41-46, mq_pc.c:44-55, rbs-u_pc.c:39-44,
 * rbs-v_pc.c:39-44, user_kernel_pc.c:38-43 (no skip in any of them) */
static double other_pc(const Prm *prm, double tx, double ty, double tz, double tq, double cx, double cy,
                       double cz, double Q, double W)
{
    double dx = tx - cx, dy = ty - cy, dz = tz - cz;
    switch (prm->kernel) {
    case KER_REGYUKAWA: {
        double kappa = prm->kp[0];
        if (prm->singularity == SNG_SUB) {
            double epsilon = prm->kp[1];
            double r = sqrt(dx * dx + dy * dy + dz * dz);
            return (Q - tq * W) * exp(-kappa * r) / sqrt(r * r + epsilon * epsilon);
        }
        double epsilon2 = prm->kp[1] * prm->kp[1];
        double r2 = dx * dx + dy * dy + dz * dz;
        double r = sqrt(r2);
        return Q * exp(-kappa * r) / sqrt(r2 + epsilon2);
    }
    case KER_SINOVERR: {
        double r = sqrt(dx * dx + dy * dy + dz * dz);
        return Q * sin(prm->kp[0] * r) / r;
    }
    case KER_USER: {
        double r = sqrt(dx * dx + dy * dy + dz * dz);
        return Q * exp(-prm->kp[0] * r) / r;
    }
    default: /* ATAN, MQ, RBS: the PP expressions (no skip there either) */
        return other_pp(prm, tx, ty, tz, tq, cx, cy, cz, Q, 0.0);
    }
}

/* CP (and CC on cluster charges): grid point c against a source s:
 * regularized-yukawa_cp.c:40-45, sin-over-r_cp.c:40-45, rbs-u_cp.c:41-45,
 * rbs-v_cp.c:41-45, user_kernel_cp.c:38-43 */
static double other_cp(const Prm *prm, double cx, double cy, double cz, double sx, double sy, double sz,
                       double q)
{
    double dx = cx - sx, dy = cy - sy, dz = cz - sz;
    switch (prm->kernel) {
    case KER_REGYUKAWA: {
        double kappa = prm->kp[0], epsilon2 = prm->kp[1] * prm->kp[1];
        double r2 = dx * dx + dy * dy + dz * dz;
        double r = sqrt(r2);
        return q * exp(-kappa * r) / sqrt(r2 + epsilon2);
    }
    case KER_SINOVERR: {
        double r = sqrt(dx * dx + dy * dy + dz * dz);
        return q * sin(prm->kp[0] * r) / r;
    }
    case KER_RBSU: {
        double r = dx * dx + dy * dy + prm->kp[0] * prm->kp[0];
        return 1. / (2. * M_PI) * q * dy / r;
    }
    case KER_RBSV: {
        double r = dx * dx + dy * dy + prm->kp[0] * prm->kp[0];
        return -1. / (2. * M_PI) * q * dx / r;
    }
    default: { /* KER_USER */
        double r = sqrt(dx * dx + dy * dy + dz * dz);
        return q * exp(-prm->kp[0] * r) / r;
    }
    }
}

static void k_pp(const Prm *prm, int nt, int ns, int ts, int ss, const Parts *T, const Parts *S, double *pot)
{
    const double *sx = S->x, *sy = S->y, *sz = S->z, *sq = S->q, *sw = S->w;
    int sub = prm->singularity == SNG_SUB;
    if (other_kernel(prm->kernel)) {
        double wadj = prm->kernel == KER_ATAN ? atan_wadj(prm) : 1.0;
#pragma omp parallel for schedule(static) if (nt >= 256)
        for (int i = 0; i < nt; i++) {
            int ii = ts + i;
            double temporary_potential = 0.0;
            for (int j = 0; j < ns; j++) {
                int jj = ss + j;
                temporary_potential += other_pp(prm, T->x[ii], T->y[ii], T->z[ii], T->q[ii], sx[jj], sy[jj], sz[jj],
                                                sq[jj], sw[jj]);
            }
            pot[ii] += prm->kernel == KER_ATAN ? wadj * temporary_potential : temporary_potential;
        }
        return;
    }
#pragma omp parallel for schedule(static) if (nt >= 256)
    for (int i = 0; i < nt; i++) {
        int ii = ts + i;
        double temporary_potential = 0.0;
        double tx = T->x[ii], ty = T->y[ii], tz = T->z[ii], tq = T->q[ii];
        for (int j = 0; j < ns; j++) {
            int jj = ss + j;
            double dx = tx - sx[jj], dy = ty - sy[jj], dz = tz - sz[jj];
            if (prm->kernel == KER_COULOMB) {
                if (!sub) {
                    double r2 = dx * dx + dy * dy + dz * dz;
                    if (r2 > DBL_MIN) temporary_potential += sq[jj] / sqrt(r2);
                } else {
                    double k2 = prm->kp[0] * prm->kp[0];
                    double r = sqrt(dx * dx + dy * dy + dz * dz);
                    if (r > DBL_MIN)
                        temporary_potential += (sq[jj] - tq * exp(-r * r / k2)) * sw[jj] / r;
                }
            } else if (prm->kernel == KER_YUKAWA) {
                double kappa = prm->kp[0];
                double r = sqrt(dx * dx + dy * dy + dz * dz);
                if (!sub) {
                    if (r > DBL_MIN) temporary_potential += sq[jj] * exp(-kappa * r) / r;
                } else {
                    if (r > DBL_MIN)
                        temporary_potential += (sq[jj] - tq) * sw[jj] * exp(-kappa * r) / r;
                }
            } else {
                if (!sub) {
                    double eps = prm->kp[0];
                    double r2 = dx * dx + dy * dy + dz * dz + eps * eps;
                    temporary_potential += sq[jj] / sqrt(r2);
                } else {
                    double alpha2 = prm->kp[0] * prm->kp[0];
                    double epsilon2 = prm->kp[1] * prm->kp[1];
                    double r2 = dx * dx + dy * dy + dz * dz;
                    temporary_potential += (sq[jj] - tq * exp(-r2 / alpha2)) * sw[jj] / sqrt(r2 + epsilon2);
                }
            }
        }
        pot[ii] += temporary_potential;
    }
}

/* Target particles against a source cluster grid: K_*_PC_Lagrange and
 * K_*_SS_PC_Lagrange (coulomb_pc.c:17-58, coulomb_ss_pc.c:19-56,
 * yukawa_pc.c:19-53, yukawa_ss_pc.c:18-55, regularized-coulomb_pc.c:19-59,
 * regularized-coulomb_ss_pc.c:19-56). */
static void k_pc(const Prm *prm, int nt, int P, int ts, size_t cs, const Parts *T, const Clus *C, double *pot)
{
    int sub = prm->singularity == SNG_SUB;
    if (other_kernel(prm->kernel)) {
        double wadj = prm->kernel == KER_ATAN ? atan_wadj(prm) : 1.0;
#pragma omp parallel for schedule(static) if (nt >= 256)
        for (int i = 0; i < nt; i++) {
            int ii = ts + i;
            double temporary_potential = 0.0;
            for (int j = 0; j < P; j++) {
                size_t jj = cs + j;
                temporary_potential += other_pc(prm, T->x[ii], T->y[ii], T->z[ii], T->q[ii], C->x[jj], C->y[jj],
                                                C->z[jj], C->q[jj], sub ? C->w[jj] : 0.0);
            }
            pot[ii] += prm->kernel == KER_ATAN ? wadj * temporary_potential : temporary_potential;
        }
        return;
    }
#pragma omp parallel for schedule(static) if (nt >= 256)
    for (int i = 0; i < nt; i++) {
        int ii = ts + i;
        double temporary_potential = 0.0;
        double tx = T->x[ii], ty = T->y[ii], tz = T->z[ii], tq = T->q[ii];
        for (int j = 0; j < P; j++) {
            size_t jj = cs + j;
            double dx = tx - C->x[jj], dy = ty - C->y[jj], dz = tz - C->z[jj];
            if (prm->kernel == KER_COULOMB) {
                if (!sub) {
                    double r2 = dx * dx + dy * dy + dz * dz;
                    temporary_potential += C->q[jj] / sqrt(r2);
                } else {
                    double k2 = prm->kp[0] * prm->kp[0];
                    double r = sqrt(dx * dx + dy * dy + dz * dz);
                    temporary_potential += (C->q[jj] - tq * C->w[jj] * exp(-r * r / k2)) / r;
                }
            } else if (prm->kernel == KER_YUKAWA) {
                double kappa = prm->kp[0];
                double r = sqrt(dx * dx + dy * dy + dz * dz);
                if (!sub) temporary_potential += C->q[jj] * exp(-kappa * r) / r;
                else temporary_potential += (C->q[jj] - tq * C->w[jj]) * exp(-kappa * r) / r;
            } else {
                if (!sub) {
                    double epsilon2 = prm->kp[0] * prm->kp[0];
                    double r2 = dx * dx + dy * dy + dz * dz + epsilon2;
                    temporary_potential += C->q[jj] / sqrt(r2);
                } else {
                    double alpha2 = prm->kp[0] * prm->kp[0];
                    double epsilon2 = prm->kp[1] * prm->kp[1];
                    double r2 = dx * dx + dy * dy + dz * dz;
                    temporary_potential += (C->q[jj] - tq * C->w[jj] * exp(-r2 / alpha2)) / sqrt(r2 + epsilon2);
                }
            }
        }
        pot[ii] += temporary_potential;
    }
}

/* Target cluster grid against source particles: K_*_CP_Lagrange and
 * K_*_SS_CP_Lagrange (coulomb_cp.c:17-58, coulomb_ss_cp.c:20-68,
 * yukawa_cp.c:19-60, yukawa_ss_cp.c:19-67, regularized-coulomb_cp.c:19-60). */
static void k_cp(const Prm *prm, int ns, int P, int ss, size_t cs, const Parts *S, Clus *C)
{
    int sub = prm->singularity == SNG_SUB;
    if (other_kernel(prm->kernel)) {
#pragma omp parallel for schedule(static) if (P >= 256)
        for (int i = 0; i < P; i++) {
            double temporary_potential = 0.0;
            for (int j = 0; j < ns; j++) {
                int jj = ss + j;
                temporary_potential += other_cp(prm, C->x[cs + i], C->y[cs + i], C->z[cs + i], S->x[jj], S->y[jj],
                                                S->z[jj], S->q[jj]);
            }
            C->q[cs + i] += temporary_potential;
        }
        return;
    }
#pragma omp parallel for schedule(static) if (P >= 256)
    for (int i = 0; i < P; i++) {
        double temporary_potential = 0.0, temporary_weight = 0.0;
        double cx = C->x[cs + i], cy = C->y[cs + i], cz = C->z[cs + i];
        for (int j = 0; j < ns; j++) {
            int jj = ss + j;
            double dx = cx - S->x[jj], dy = cy - S->y[jj], dz = cz - S->z[jj];
            if (prm->kernel == KER_COULOMB) {
                if (!sub) {
                    double r2 = dx * dx + dy * dy + dz * dz;
                    temporary_potential += S->q[jj] / sqrt(r2);
                } else {
                    double k2 = prm->kp[0] * prm->kp[0];
                    double r = sqrt(dx * dx + dy * dy + dz * dz);
                    temporary_potential += S->q[jj] * S->w[jj] / r;
                    temporary_weight += exp(-r * r / k2) * S->w[jj] / r;
                }
            } else if (prm->kernel == KER_YUKAWA) {
                double kappa = prm->kp[0];
                double r = sqrt(dx * dx + dy * dy + dz * dz);
                if (!sub) {
                    temporary_potential += S->q[jj] * exp(-kappa * r) / r;
                } else {
                    temporary_potential += S->q[jj] * S->w[jj] * exp(-kappa * r) / r;
                    temporary_weight += S->w[jj] * exp(-kappa * r) / r;
                }
            } else {
                double epsilon2 = prm->kp[0] * prm->kp[0];
                double r2 = dx * dx + dy * dy + dz * dz + epsilon2;
                temporary_potential += S->q[jj] / sqrt(r2);
            }
        }
        C->q[cs + i] += temporary_potential;
        if (sub) C->w[cs + i] += temporary_weight;
    }
}

/* Target cluster grid against a source cluster grid: K_*_CP_Lagrange on
 * cluster charges (interaction_compute_cc.c:121-127), K_Coulomb_SS_CC_Lagrange
 * (coulomb_ss_cc.c:20-68), K_Yukawa_SS_CC_Lagrange (yukawa_ss_cc.c:19-67). */
static void k_cc(const Prm *prm, int P, size_t ss, size_t ts, const Clus *S, Clus *C)
{
    int sub = prm->singularity == SNG_SUB;
    if (other_kernel(prm->kernel)) {
#pragma omp parallel for schedule(static) if (P >= 256)
        for (int i = 0; i < P; i++) {
            double temporary_potential = 0.0;
            for (int j = 0; j < P; j++) {
                size_t jj = ss + j;
                temporary_potential += other_cp(prm, C->x[ts + i], C->y[ts + i], C->z[ts + i], S->x[jj], S->y[jj],
                                                S->z[jj], S->q[jj]);
            }
            C->q[ts + i] += temporary_potential;
        }
        return;
    }
#pragma omp parallel for schedule(static) if (P >= 256)
    for (int i = 0; i < P; i++) {
        double temporary_potential = 0.0, temporary_weight = 0.0;
        double cx = C->x[ts + i], cy = C->y[ts + i], cz = C->z[ts + i];
        for (int j = 0; j < P; j++) {
            size_t jj = ss + j;
            double dx = cx - S->x[jj], dy = cy - S->y[jj], dz = cz - S->z[jj];
            if (prm->kernel == KER_COULOMB) {
                if (!sub) {
                    double r2 = dx * dx + dy * dy + dz * dz;
                    temporary_potential += S->q[jj] / sqrt(r2);
                } else {
                    double k2 = prm->kp[0] * prm->kp[0];
                    double r = sqrt(dx * dx + dy * dy + dz * dz);
                    temporary_potential += S->q[jj] / r;
                    temporary_weight += exp(-r * r / k2) * S->w[jj] / r;
                }
            } else if (prm->kernel == KER_YUKAWA) {
                double kappa = prm->kp[0];
                double r = sqrt(dx * dx + dy * dy + dz * dz);
                if (!sub) {
                    temporary_potential += S->q[jj] * exp(-kappa * r) / r;
                } else {
                    temporary_potential += S->q[jj] * exp(-kappa * r) / r;
                    temporary_weight += S->w[jj] * exp(-kappa * r) / r;
                }
            } else {
                double epsilon2 = prm->kp[0] * prm->kp[0];
                double r2 = dx * dx + dy * dy + dz * dz + epsilon2;
                temporary_potential += S->q[jj] / sqrt(r2);
            }
        }
        C->q[ts + i] += temporary_potential;
        if (sub) C->w[ts + i] += temporary_weight;
    }
}

/* ------------------------------------------------------------------ */
/* Downpass: interaction_compute_downpass.c:37-142, 151-497, 502-856   */
/* Evaluates the interpolant of the grid values (Q, W) of the box       */
/* [x0,x1]x... at m points; adds to out1 (and out2).  When tq is given  */
/* (subtraction, leaf to targets), the value interpolated is Q - tq*W. */
/* ------------------------------------------------------------------ */

static void interp_eval(int n, double x0, double x1, double y0, double y1, double z0, double z1,
                        const double *gq, const double *gw, int m,
                        const double *px, const double *py, const double *pz, const double *tq,
                        double *out1, double *out2)
{
    int nl = n + 1, P = nl * nl * nl;
    double nodeX[32], nodeY[32], nodeZ[32], wts[32];
    cheb_nodes(n, x0, x1, nodeX);
    cheb_nodes(n, y0, y1, nodeY);
    cheb_nodes(n, z0, z1, nodeZ);
    bary_weights(n, wts);
    for (int i = 0; i < m; i++) {
        double sumX = 0.0, sumY = 0.0, sumZ = 0.0;
        double tx = px[i], ty = py[i], tz = pz[i];
        int eix = -1, eiy = -1, eiz = -1;
        for (int j = 0; j < nl; j++) {
            double cx = tx - nodeX[j], cy = ty - nodeY[j], cz = tz - nodeZ[j];
            if (fabs(cx) < DBL_MIN) eix = j;
            if (fabs(cy) < DBL_MIN) eiy = j;
            if (fabs(cz) < DBL_MIN) eiz = j;
            double w = wts[j];
            sumX += w / cx;
            sumY += w / cy;
            sumZ += w / cz;
        }
        double denominator = 1.0;
        if (eix == -1) denominator /= sumX;
        if (eiy == -1) denominator /= sumY;
        if (eiz == -1) denominator /= sumZ;
        double temp = 0.0, temp2 = 0.0;
        double tqi = tq ? tq[i] : 0.0;
        for (int j = 0; j < P; j++) {
            int k1 = j % nl;
            int kk = (j - k1) / nl;
            int k2 = kk % nl;
            kk = kk - k2;
            int k3 = kk / nl;
            double w3 = wts[k3], w2 = wts[k2], w1 = wts[k1];
            double cx = nodeX[k1], cy = nodeY[k2], cz = nodeZ[k3];
            double numerator = 1.0;
            if (eix == -1) numerator *= w1 / (tx - cx);
            else if (eix != k1) numerator *= 0;
            if (eiy == -1) numerator *= w2 / (ty - cy);
            else if (eiy != k2) numerator *= 0;
            if (eiz == -1) numerator *= w3 / (tz - cz);
            else if (eiz != k3) numerator *= 0;
            if (tq) {
                temp += numerator * denominator * (gq[j] - tqi * gw[j]);
            } else {
                temp += numerator * denominator * gq[j];
                if (gw && out2) temp2 += numerator * denominator * gw[j];
            }
        }
        out1[i] += temp;
        if (!tq && gw && out2) out2[i] += temp2;
    }
}

static void downpass(const Tree *t, Clus *c, const Parts *T, const Prm *prm, double *pot)
{
    int P = prm->P, n = prm->degree;
    int sub = prm->singularity == SNG_SUB;
    /* parents write their own children's grids, leaves their own targets */
    for (int level = 0; level < t->max_depth; ++level) {
#pragma omp parallel for schedule(dynamic)
        for (int ci = 0; ci < t->lev_n[level]; ++ci) {
            int par = t->lev[level][ci];
            size_t ps = (size_t)par * P;
            for (int k = 0; k < t->nchild[par]; ++k) {
                int ch = t->child[8 * par + k];
                size_t cs = (size_t)ch * P;
                interp_eval(n, t->xmin[par], t->xmax[par], t->ymin[par], t->ymax[par], t->zmin[par], t->zmax[par],
                            c->q + ps, sub ? c->w + ps : NULL, P, c->x + cs, c->y + cs, c->z + cs, NULL,
                            c->q + cs, sub ? c->w + cs : NULL);
            }
        }
    }
#pragma omp parallel for schedule(dynamic)
    for (int li = 0; li < t->nleaves; ++li) {
        int idx = t->leaves[li];
        int b = t->ibeg[idx] - 1, m = t->iend[idx] - t->ibeg[idx] + 1;
        size_t cs = (size_t)idx * P;
        interp_eval(n, t->xmin[idx], t->xmax[idx], t->ymin[idx], t->ymax[idx], t->zmin[idx], t->zmax[idx],
                    c->q + cs, sub ? c->w + cs : NULL, m, T->x + b, T->y + b, T->z + b,
                    sub ? T->q + b : NULL, pot + b, NULL);
    }
}

/* ------------------------------------------------------------------ */
/* Hermite approximation: 8 planes per grid point (value, the three     */
/* first, the three mixed second and the mixed third derivative terms,  */
/* in the order 0, x, y, z, xy, yz, xz, xyz).  Planes of node c sit at  */
/* q[8*P*c + m*P + j]; grid coordinates at x[P*c + j].                  */
/* ------------------------------------------------------------------ */

/* Nodes, Hermite barycentric weights dj and w of one dimension of the box
 * [a0,a1] (clusters.c:1296-1318, interaction_compute_downpass.c:913-933). */
static void herm_weights(int n, double a0, double a1, double *node, double *dj, double *w)
{
    double ww[32];
    for (int i = 0; i < n + 1; i++) {
        double xx = i * M_PI / n;
        double tt = cos(xx);
        ww[i] = -cos(xx) / (2 * sin(xx) * sin(xx));
        node[i] = a0 + (tt + 1.0) / 2.0 * (a1 - a0);
    }
    ww[0] = 0.25 * (n * n / 3.0 + 1.0 / 6.0);
    ww[n] = -ww[0];
    for (int j = 0; j < n + 1; j++) {
        dj[j] = 1.0;
        w[j] = -4.0 * ww[j] / (a1 - a0);
    }
    dj[0] = 0.25;
    dj[n] = 0.25;
}

/* The 8 numerators of grid point (k1,k2,k3) for a point at displacement
 * (dx,dy,dz) from it, with exact-hit indices e* (-1: none):
 * clusters.c:1392-1462, interaction_compute_downpass.c:1010-1082. */
static void herm_numerators(int k1, int k2, int k3, double dx, double dy, double dz, int ex, int ey, int ez,
                            const double *djx, const double *wx, const double *wy, const double *wz,
                            double *num)
{
    for (int m = 0; m < 8; m++) num[m] = 1.0;
    double Ax = djx[k1] / (dx * dx) + wx[k1] / dx;
    double Ay = djx[k2] / (dy * dy) + wy[k2] / dy;
    double Az = djx[k3] / (dz * dz) + wz[k3] / dz;
    double Bx = djx[k1] / dx, By = djx[k2] / dy, Bz = djx[k3] / dz;
    if (ex == -1) {
        num[0] *= Ax; num[1] *= Bx; num[2] *= Ax; num[3] *= Ax;
        num[4] *= Bx; num[5] *= Ax; num[6] *= Bx; num[7] *= Bx;
    } else if (ex != k1) {
        for (int m = 0; m < 8; m++) num[m] *= 0;
    } else {
        num[1] *= 0; num[4] *= 0; num[6] *= 0; num[7] *= 0;
    }
    if (ey == -1) {
        num[0] *= Ay; num[1] *= Ay; num[2] *= By; num[3] *= Ay;
        num[4] *= By; num[5] *= By; num[6] *= Ay; num[7] *= By;
    } else if (ey != k2) {
        for (int m = 0; m < 8; m++) num[m] *= 0;
    } else {
        num[2] *= 0; num[4] *= 0; num[5] *= 0; num[7] *= 0;
    }
    if (ez == -1) {
        num[0] *= Az; num[1] *= Az; num[2] *= Az; num[3] *= Bz;
        num[4] *= Az; num[5] *= Bz; num[6] *= Bz; num[7] *= Bz;
    } else if (ez != k3) {
        for (int m = 0; m < 8; m++) num[m] *= 0;
    } else {
        num[3] *= 0; num[5] *= 0; num[6] *= 0; num[7] *= 0;
    }
}

/* Upward pass of one node straight from its particles: clusters.c:1224-1836
 * (pc_comp_ms_modifiedF_hermite, _hermite_SS).  f1 = q (or q*w), f2 = w
 * (subtraction) or NULL; Q (and W) receive 8 planes of P values. */
static int anterp_hermite(int n, double x0, double x1, double y0, double y1, double z0, double z1, int m,
                          const double *px, const double *py, const double *pz, const double *f1in,
                          const double *f2in, double *gx, double *gy, double *gz, double *Q, double *W)
{
    int nl = n + 1, P = nl * nl * nl;
    double nodeX[32], nodeY[32], nodeZ[32], dj[32], wx[32], wy[32], wz[32];
    double *f1 = malloc((size_t)(m > 0 ? m : 1) * sizeof(double));
    double *f2 = f2in ? malloc((size_t)(m > 0 ? m : 1) * sizeof(double)) : NULL;
    int *ex = malloc((size_t)(m > 0 ? m : 1) * 3 * sizeof(int));
    if (!f1 || (f2in && !f2) || !ex) { free(f1); free(f2); free(ex); return ORC_ERR_NOMEM; }
    int *ey = ex + m, *ez = ex + 2 * m;
    herm_weights(n, x0, x1, nodeX, dj, wx);
    herm_weights(n, y0, y1, nodeY, dj, wy);
    herm_weights(n, z0, z1, nodeZ, dj, wz);
    for (int i = 0; i < m; i++) {
        f1[i] = f1in[i];
        if (f2) f2[i] = f2in[i];
        ex[i] = ey[i] = ez[i] = -1;
        double sumX = 0.0, sumY = 0.0, sumZ = 0.0;
        for (int j = 0; j < nl; j++) {
            double dx = px[i] - nodeX[j], dy = py[i] - nodeY[j], dz = pz[i] - nodeZ[j];
            if (fabs(dx) < DBL_MIN) ex[i] = j;
            if (fabs(dy) < DBL_MIN) ey[i] = j;
            if (fabs(dz) < DBL_MIN) ez[i] = j;
            sumX += dj[j] / (dx * dx) + wx[j] / dx;
            sumY += dj[j] / (dy * dy) + wy[j] / dy;
            sumZ += dj[j] / (dz * dz) + wz[j] / dz;
        }
        double denominator = 1.0;
        if (ex[i] == -1) denominator *= sumX;
        if (ey[i] == -1) denominator *= sumY;
        if (ez[i] == -1) denominator *= sumZ;
        f1[i] /= denominator;
        if (f2) f2[i] /= denominator;
    }
    for (int j = 0; j < P; j++) {
        int k1 = j % nl;
        int kk = (j - k1) / nl;
        int k2 = kk % nl;
        kk = kk - k2;
        int k3 = kk / nl;
        double cx = nodeX[k1], cy = nodeY[k2], cz = nodeZ[k3];
        gx[j] = cx; gy[j] = cy; gz[j] = cz;
        double tq[8] = {0}, tw[8] = {0}, num[8];
        for (int i = 0; i < m; i++) {
            herm_numerators(k1, k2, k3, px[i] - cx, py[i] - cy, pz[i] - cz, ex[i], ey[i], ez[i],
                            dj, wx, wy, wz, num);
            for (int q = 0; q < 8; q++) tq[q] += num[q] * f1[i];
            if (f2)
                for (int q = 0; q < 8; q++) tw[q] += num[q] * f2[i];
        }
        for (int q = 0; q < 8; q++) Q[q * P + j] += tq[q];
        if (f2)
            for (int q = 0; q < 8; q++) W[q * P + j] += tw[q];
    }
    free(f1); free(f2); free(ex);
    return 0;
}

/* clusters.c:182-190: every node from its own particles */
static int clusters_sources_hermite(Clus *c, const Tree *t, const Parts *s, const Prm *prm)
{
    int sub = prm->singularity == SNG_SUB;
    int P = prm->P, n = prm->degree;
    int e = clus_alloc_planes(c, t->n, P, 8, sub ? 8 : 0);
    if (e) return e;
    double *qw = NULL;
    if (sub) {
        qw = malloc((size_t)s->n * sizeof(double));
        if (!qw) return ORC_ERR_NOMEM;
        for (int i = 0; i < s->n; i++) qw[i] = s->q[i] * s->w[i];
    }
    for (int idx = 0; idx < t->n && !e; idx++) {
        int b = t->ibeg[idx] - 1, m = t->iend[idx] - t->ibeg[idx] + 1;
        size_t cs = (size_t)idx * P;
        e = anterp_hermite(n, t->xmin[idx], t->xmax[idx], t->ymin[idx], t->ymax[idx], t->zmin[idx],
                           t->zmax[idx], m, s->x + b, s->y + b, s->z + b, sub ? qw + b : s->q + b,
                           sub ? s->w + b : NULL, c->x + cs, c->y + cs, c->z + cs, c->q + 8 * cs,
                           sub ? c->w + 8 * cs : NULL);
    }
    free(qw);
    return e;
}

/* Target particles against a source cluster's 8 planes: K_Coulomb_PC_Hermite
 * (coulomb_pc.c:66-135), K_Yukawa_PC_Hermite (yukawa_pc.c:61-140),
 * K_RegularizedCoulomb_PC_Hermite (regularized-coulomb_pc.c:67-139) and the
 * subtraction forms (coulomb_ss_pc.c:63-163, yukawa_ss_pc.c:61-167,
 * regularized-coulomb_ss_pc.c:63-160). */
static void k_pc_hermite(const Prm *prm, int nt, int P, int ts, size_t cs, const Parts *T, const Clus *C,
                         double *pot)
{
    int sub = prm->singularity == SNG_SUB;
    const double *q0 = C->q + 8 * cs, *q1 = q0 + P, *q2 = q0 + 2 * P, *q3 = q0 + 3 * P, *q4 = q0 + 4 * P,
                 *q5 = q0 + 5 * P, *q6 = q0 + 6 * P, *q7 = q0 + 7 * P;
    const double *w0 = NULL, *w1 = NULL, *w2 = NULL, *w3 = NULL, *w4 = NULL, *w5 = NULL, *w6 = NULL, *w7 = NULL;
    if (sub) {
        w0 = C->w + 8 * cs; w1 = w0 + P; w2 = w0 + 2 * P; w3 = w0 + 3 * P;
        w4 = w0 + 4 * P; w5 = w0 + 5 * P; w6 = w0 + 6 * P; w7 = w0 + 7 * P;
    }
    for (int i = 0; i < nt; i++) {
        int ii = ts + i;
        double temporary_potential = 0.0;
        double tx = T->x[ii], ty = T->y[ii], tz = T->z[ii], tq = T->q[ii];
        for (int j = 0; j < P; j++) {
            size_t jj = cs + j;
            double dx = tx - C->x[jj], dy = ty - C->y[jj], dz = tz - C->z[jj];
            if (prm->kernel == KER_SINOVERR) {
                /* sin-over-r_pc.c:68-155 */
                double k = prm->kp[0], k2 = k * k, k3 = k * k2;
                double r = sqrt(dx * dx + dy * dy + dz * dz);
                if (r > DBL_MIN) {
                    double rinv = 1 / r;
                    double r2inv = rinv * rinv, r3inv = rinv * r2inv, r4inv = r2inv * r2inv;
                    double r5inv = r3inv * r2inv, r6inv = r3inv * r3inv, r7inv = r4inv * r3inv;
                    double sinr = sin(k * r), cosr = cos(k * r);
                    temporary_potential += sinr * rinv * (q0[j])
                        + (sinr * r3inv - k * cosr * r2inv) * (q1[j] * dx + q2[j] * dy + q3[j] * dz)
                        + (sinr * (3 * r5inv - k2 * r3inv) - 3 * k * cosr * r4inv)
                              * (q4[j] * dx * dy + q5[j] * dy * dz + q6[j] * dx * dz)
                        + (sinr * (15 * r7inv - 6 * k2 * r5inv) + cosr * (k3 * r4inv - 15 * k * r6inv))
                              * q7[j] * dx * dy * dz;
                }
            } else if (prm->kernel == KER_YUKAWA) {
                double k = prm->kp[0], k2 = k * k, k3 = k * k2;
                double r = sqrt(dx * dx + dy * dy + dz * dz);
                double r2 = r * r, r3 = r2 * r;
                double rinv = 1 / r;
                double r3inv = rinv * rinv * rinv, r5inv = r3inv * rinv * rinv, r7inv = r5inv * rinv * rinv;
                if (!sub) {
                    temporary_potential += exp(-k * r)
                        * (rinv * (q0[j])
                        + r3inv * (1 + k * r) * (q1[j] * dx + q2[j] * dy + q3[j] * dz)
                        + r5inv * (3 + 3 * k * r + k2 * r2) * (q4[j] * dx * dy + q5[j] * dy * dz + q6[j] * dx * dz)
                        + r7inv * (15 + 15 * k * r + 6 * k2 * r2 + k3 * r3) * q7[j] * dx * dy * dz);
                } else {
                    double kr = k * r, k2r2 = kr * kr, k3r3 = k2r2 * kr;
                    (void)k2r2; (void)k3r3;
                    double d0 = q0[j] - w0[j] * tq, d1 = q1[j] - w1[j] * tq, d2 = q2[j] - w2[j] * tq,
                           d3 = q3[j] - w3[j] * tq, d4 = q4[j] - w4[j] * tq, d5 = q5[j] - w5[j] * tq,
                           d6 = q6[j] - w6[j] * tq, d7 = q7[j] - w7[j] * tq;
                    temporary_potential += exp(-k * r)
                        * (rinv * (d0)
                        + r3inv * (1 + k * r) * (d1 * dx + d2 * dy + d3 * dz)
                        + r5inv * (3 + 3 * k * r + k2 * r2) * (d4 * dx * dy + d5 * dy * dz + d6 * dx * dz)
                        + r7inv * (15 + 15 * k * r + 6 * k2 * r2 + k3 * r3) * d7 * dx * dy * dz);
                }
            } else {
                /* Coulomb and regularized Coulomb: the charge part is the
                 * derivative expansion of 1/r (r including eps^2 for the
                 * regularized kernel); subtraction adds that of exp(-r^2/k^2)/r */
                double r;
                if (prm->kernel == KER_COULOMB) r = sqrt(dx * dx + dy * dy + dz * dz);
                else r = sqrt(dx * dx + dy * dy + dz * dz + (sub ? prm->kp[1] * prm->kp[1] : prm->kp[0] * prm->kp[0]));
                double rinv = 1 / r;
                double r3inv = rinv * rinv * rinv, r5inv = r3inv * rinv * rinv, r7inv = r5inv * rinv * rinv;
                if (!sub) {
                    temporary_potential += rinv * (q0[j])
                        + r3inv * (q1[j] * dx + q2[j] * dy + q3[j] * dz)
                        + 3 * r5inv * (q4[j] * dx * dy + q5[j] * dy * dz + q6[j] * dx * dz)
                        + 15 * r7inv * q7[j] * dx * dy * dz;
                } else {
                    double k2 = prm->kp[0] * prm->kp[0];
                    double r2 = r * r;
                    double rk2 = r2 / k2, rk4 = rk2 * rk2, rk6 = rk4 * rk2;
                    temporary_potential += rinv * (q0[j])
                        + r3inv * (q1[j] * dx + q2[j] * dy + q3[j] * dz)
                        + 3 * r5inv * (q4[j] * dx * dy + q5[j] * dy * dz + q6[j] * dx * dz)
                        + 15 * r7inv * q7[j] * dx * dy * dz
                        - tq * exp(-rk2)
                            * (rinv * (w0[j])
                            + r3inv * (1 + 2 * rk2) * (w1[j] * dx + w2[j] * dy + w3[j] * dz)
                            + r5inv * (3 + 4 * rk2 + 4 * rk4) * (w4[j] * dx * dy + w5[j] * dy * dz + w6[j] * dx * dz)
                            + r7inv * (15 + 18 * rk2 + 12 * rk4 + 8 * rk6) * w7[j] * dx * dy * dz);
                }
            }
        }
        pot[ii] += temporary_potential;
    }
}

/* Target cluster grid (8 planes) against source particles:
 * K_Coulomb_CP_Hermite (coulomb_cp.c:66-184), K_Yukawa_CP_Hermite
 * (yukawa_cp.c:68-194), K_RegularizedCoulomb_CP_Hermite
 * (regularized-coulomb_cp.c:68-188). */
static void k_cp_hermite(const Prm *prm, int ns, int P, int ss, size_t cs, const Parts *S, Clus *C)
{
    double *q0 = C->q + 8 * cs;
    for (int i = 0; i < P; i++) {
        double t0 = 0.0, t1 = 0.0, t2 = 0.0, t3 = 0.0, t4 = 0.0, t5 = 0.0, t6 = 0.0, t7 = 0.0;
        double cx = C->x[cs + i], cy = C->y[cs + i], cz = C->z[cs + i];
        for (int j = 0; j < ns; j++) {
            int jj = ss + j;
            double dx = S->x[jj] - cx, dy = S->y[jj] - cy, dz = S->z[jj] - cz;
            if (prm->kernel == KER_SINOVERR) {
                /* sin-over-r_cp.c:68-198 */
                double k = prm->kp[0], k2 = k * k, k3 = k * k2;
                double r = sqrt(dx * dx + dy * dy + dz * dz);
                double rinv = 1 / r;
                double r2inv = rinv * rinv, r3inv = rinv * r2inv, r4inv = r2inv * r2inv;
                double r5inv = r3inv * r2inv, r6inv = r3inv * r3inv, r7inv = r4inv * r3inv;
                double sinr = sin(k * r) * S->q[jj], cosr = cos(k * r) * S->q[jj];
                double d0 = sinr * rinv;
                double d1 = sinr * r3inv - k * cosr * r2inv;
                double d2 = sinr * (3 * r5inv - k2 * r3inv) - 3 * k * cosr * r4inv;
                double d3 = sinr * (15 * r7inv - 6 * k2 * r5inv) + cosr * (k3 * r4inv - 15 * k * r6inv);
                t0 += d0;
                t1 += d1 * dx; t2 += d1 * dy; t3 += d1 * dz;
                t4 += d2 * dx * dy; t5 += d2 * dy * dz; t6 += d2 * dx * dz;
                t7 += d3 * dx * dy * dz;
            } else if (prm->kernel == KER_YUKAWA) {
                double k = prm->kp[0], k2 = k * k, k3 = k * k2;
                double r2 = dx * dx + dy * dy + dz * dz;
                double r = sqrt(r2), r3 = r2 * r;
                double r2inv = 1 / r2;
                double rinvq = S->q[jj] / r * exp(-k * r);
                double r3inv = rinvq * r2inv, r5inv = r3inv * r2inv, r7inv = r5inv * r2inv;
                double d1 = r3inv * (1 + k * r);
                double d2 = r5inv * (3 + 3 * k * r + k2 * r2);
                double d3 = r7inv * (15 + 15 * k * r + 6 * k2 * r2 + k3 * r3);
                t0 += rinvq;
                t1 += d1 * dx; t2 += d1 * dy; t3 += d1 * dz;
                t4 += d2 * dx * dy; t5 += d2 * dy * dz; t6 += d2 * dx * dz;
                t7 += d3 * dx * dy * dz;
            } else {
                double r2 = dx * dx + dy * dy + dz * dz;
                if (prm->kernel == KER_REGCOULOMB) r2 += prm->kp[0] * prm->kp[0];
                double r2inv = 1 / r2;
                double rinvq = S->q[jj] / sqrt(r2);
                double r3inv = rinvq * r2inv, r5inv = r3inv * r2inv, r7inv = r5inv * r2inv;
                r5inv *= 3.0;
                t0 += rinvq;
                t1 += r3inv * dx; t2 += r3inv * dy; t3 += r3inv * dz;
                t4 += r5inv * dx * dy; t5 += r5inv * dy * dz; t6 += r5inv * dx * dz;
                t7 += r7inv * dx * dy * dz * 15.0;
            }
        }
        q0[i] += t0; q0[P + i] += t1; q0[2 * P + i] += t2; q0[3 * P + i] += t3;
        q0[4 * P + i] += t4; q0[5 * P + i] += t5; q0[6 * P + i] += t6; q0[7 * P + i] += t7;
    }
}

/* Downpass: every target node's 8 planes to its own particles
 * (interaction_compute_downpass.c:117-120, 860-1109). */
static void downpass_hermite(const Tree *t, const Clus *c, const Parts *T, const Prm *prm, double *pot)
{
    int n = prm->degree, nl = n + 1, P = prm->P;
    for (int idx = 0; idx < t->n; idx++) {
        int b = t->ibeg[idx] - 1, m = t->iend[idx] - t->ibeg[idx] + 1;
        const double *q0 = c->q + (size_t)8 * P * idx;
        double nodeX[32], nodeY[32], nodeZ[32], dj[32], wx[32], wy[32], wz[32];
        herm_weights(n, t->xmin[idx], t->xmax[idx], nodeX, dj, wx);
        herm_weights(n, t->ymin[idx], t->ymax[idx], nodeY, dj, wy);
        herm_weights(n, t->zmin[idx], t->zmax[idx], nodeZ, dj, wz);
        for (int i = 0; i < m; i++) {
            double tx = T->x[b + i], ty = T->y[b + i], tz = T->z[b + i];
            double sumX = 0.0, sumY = 0.0, sumZ = 0.0;
            int eix = -1, eiy = -1, eiz = -1;
            for (int j = 0; j < nl; j++) {
                double cx = tx - nodeX[j], cy = ty - nodeY[j], cz = tz - nodeZ[j];
                if (fabs(cx) < DBL_MIN) eix = j;
                if (fabs(cy) < DBL_MIN) eiy = j;
                if (fabs(cz) < DBL_MIN) eiz = j;
                sumX += dj[j] / (cx * cx) + wx[j] / cx;
                sumY += dj[j] / (cy * cy) + wy[j] / cy;
                sumZ += dj[j] / (cz * cz) + wz[j] / cz;
            }
            double denominator = 1.0;
            if (eix == -1) denominator /= sumX;
            if (eiy == -1) denominator /= sumY;
            if (eiz == -1) denominator /= sumZ;
            double temp = 0.0, num[8];
            for (int j = 0; j < P; j++) {
                int k1 = j % nl;
                int kk = (j - k1) / nl;
                int k2 = kk % nl;
                kk = kk - k2;
                int k3 = kk / nl;
                herm_numerators(k1, k2, k3, tx - nodeX[k1], ty - nodeY[k2], tz - nodeZ[k3], eix, eiy, eiz,
                                dj, wx, wy, wz, num);
                temp += denominator * (num[0] * q0[j] + num[1] * q0[P + j] + num[2] * q0[2 * P + j]
                                     + num[3] * q0[3 * P + j] + num[4] * q0[4 * P + j] + num[5] * q0[5 * P + j]
                                     + num[6] * q0[6 * P + j] + num[7] * q0[7 * P + j]);
            }
            pot[b + i] += temp;
        }
    }
}

/* interaction_compute_correction.c:22-45 and the *_ss_correction.c kernels */
static void ss_correction(const Prm *prm, const Parts *T, double *pot)
{
    if (prm->singularity != SNG_SUB) return;
    double param = 0.0;
    if (prm->kernel == KER_COULOMB) param = 2.0 * M_PI * prm->kp[0] * prm->kp[0];
    else if (prm->kernel == KER_REGCOULOMB) param = 2.0 * M_PI * prm->kp[0] * prm->kp[0];
    else if (prm->kernel == KER_YUKAWA) param = 4.0 * M_PI / prm->kp[0] / prm->kp[0];
    else if (prm->kernel == KER_REGYUKAWA) param = 4.0 * M_PI / prm->kp[0] / prm->kp[0];
    for (int i = 0; i < T->n; i++) pot[i] += param * T->q[i];
}

/* Which (kernel, singularity, approximation, compute type) the reference
 * dispatchers accept (interaction_compute_{pc,cp,cc}.c; Hermite: PC with
 * either singularity, CP with skipping (interaction_compute_cp.c:133, 187,
 * interaction_compute_downpass.c:126-128), no CC (interaction_compute_cc.c:146)). */
static int supported(const Prm *p)
{
    if (p->approximation != APX_LAGRANGE && p->approximation != APX_HERMITE) return 0;
    if (other_kernel(p->kernel)) {
        /* the Lagrange combinations the dispatchers compute as written:
         * ATAN and MQ have PC kernels only; REGULARIZED_YUKAWA also
         * subtraction in PC; every other combination exits there or
         * silently drops the far field */
        if (p->compute_type < CT_PC || p->compute_type > CT_CC) return 0;
        /* Hermite: sin-over-r in PC and CP with skipping only
         * (interaction_compute_pc.c:351-358, interaction_compute_cp.c:317-324) */
        if (p->approximation == APX_HERMITE)
            return p->kernel == KER_SINOVERR && p->compute_type != CT_CC && p->singularity == SNG_SKIP &&
                   p->degree >= 1 && p->degree <= 30;
        if (p->approximation != APX_LAGRANGE) return 0;
        if ((p->kernel == KER_ATAN || p->kernel == KER_MQ) && p->compute_type != CT_PC) return 0;
        if (p->singularity == SNG_SUB && !(p->kernel == KER_REGYUKAWA && p->compute_type == CT_PC)) return 0;
        if (p->singularity != SNG_SKIP && p->singularity != SNG_SUB) return 0;
        return p->degree >= 1 && p->degree <= 30;
    }
    if (p->kernel != KER_COULOMB && p->kernel != KER_YUKAWA && p->kernel != KER_REGCOULOMB) return 0;
    if (p->singularity != SNG_SKIP && p->singularity != SNG_SUB) return 0;
    if (p->compute_type < CT_PC || p->compute_type > CT_CC) return 0;
    if (p->approximation == APX_HERMITE) {
        if (p->compute_type == CT_CC) return 0;
        if (p->compute_type == CT_CP && p->singularity == SNG_SUB) return 0;
    } else if (p->kernel == KER_REGCOULOMB && p->singularity == SNG_SUB && p->compute_type != CT_PC) {
        return 0;
    }
    if (p->degree < 1 || p->degree > 30) return 0;
    return 1;
}

static void stats_counts(const Lists *L, const Tree *tgt, const Tree *src, const Prm *prm)
{
    long long P = prm->P;
    for (int i = 0; i < L->n; i++) {
        long long nt = tgt->numpar[i];
        g_stats.num_pp += L->pp[i].n;
        for (int j = 0; j < L->pp[i].n; j++) g_stats.pp_ptwise += nt * src->numpar[L->pp[i].v[j]];
        if (prm->compute_type == CT_PC) {
            g_stats.num_pc += L->ap[i].n;
            for (int j = 0; j < L->ap[i].n; j++) {
                g_stats.pc_ptwise += nt * P;
                g_stats.pc_replaced += nt * src->numpar[L->ap[i].v[j]];
            }
        } else if (prm->compute_type == CT_CP) {
            g_stats.num_cp += L->ap[i].n;
            for (int j = 0; j < L->ap[i].n; j++) {
                g_stats.cp_ptwise += nt * P;
                g_stats.cp_replaced += nt * src->numpar[L->ap[i].v[j]];
            }
        } else {
            g_stats.num_cc += L->ap[i].n;
            g_stats.num_pc += L->spc[i].n;
            g_stats.num_cp += L->tcp[i].n;
            for (int j = 0; j < L->ap[i].n; j++) {
                g_stats.cc_ptwise += P * P;
                g_stats.cc_replaced += nt * src->numpar[L->ap[i].v[j]];
            }
            for (int j = 0; j < L->spc[i].n; j++) {
                g_stats.pc_ptwise += nt * P;
                g_stats.pc_replaced += nt * src->numpar[L->spc[i].v[j]];
            }
            for (int j = 0; j < L->tcp[i].n; j++) {
                g_stats.cp_ptwise += P * src->numpar[L->tcp[i].v[j]];
                g_stats.cp_replaced += nt * src->numpar[L->tcp[i].v[j]];
            }
        }
    }
}

int orc_treecode(int M, int N,
                 const double *tx, const double *ty, const double *tz, const double *tq,
                 const double *sx, const double *sy, const double *sz, const double *sq,
                 const double *sw, double *out,
                 int kernel, int nkp, const double *kp,
                 int singularity, int approximation, int compute_type,
                 double theta, int degree, int msl, int mtl,
                 double size_check, double beta, int verbosity)
{
    (void)verbosity;
    memset(&g_stats, 0, sizeof(g_stats));
    if (M <= 0 || N <= 0 || !out || !tx || !ty || !tz || !sx || !sy || !sz || !sq) return ORC_ERR_ARGS;
    memset(out, 0, (size_t)M * sizeof(double));
    Prm prm;
    memset(&prm, 0, sizeof(prm));
    prm.kernel = kernel; prm.singularity = singularity; prm.approximation = approximation;
    prm.compute_type = compute_type;
    for (int i = 0; i < nkp && i < 4; i++) prm.kp[i] = kp[i];
    orc_params_resolve(compute_type, approximation, theta, degree, msl, mtl, size_check, beta,
                       &prm.theta, &prm.degree, &prm.size_check_factor, &prm.msl, &prm.mtl);
    prm.P = (prm.degree + 1) * (prm.degree + 1) * (prm.degree + 1);
    if (!supported(&prm)) return ORC_ERR_UNSUPPORTED;
    if (singularity == SNG_SUB && (!sw || !tq)) return ORC_ERR_ARGS;
    if (prm.msl < 1 || prm.mtl < 1) return ORC_ERR_ARGS;
    g_stats.interp_degree = prm.degree; g_stats.pts_per_cluster = prm.P; g_stats.theta = prm.theta;
    g_stats.max_per_source_leaf = prm.msl; g_stats.max_per_target_leaf = prm.mtl;

    int rc = 0;
    double t0 = wall();
    Parts S, T;
    memset(&S, 0, sizeof(S)); memset(&T, 0, sizeof(T));
    if ((rc = parts_copy(&S, N, sx, sy, sz, sq, sw, 1))) goto done0;
    if ((rc = parts_copy(&T, M, tx, ty, tz, tq, NULL, 0))) goto done0;
    double *pot = calloc((size_t)M, sizeof(double));
    if (!pot) { rc = ORC_ERR_NOMEM; goto done0; }

    Tree src, tgt;
    Clus sc, tc;
    Lists L;
    memset(&src, 0, sizeof(src)); memset(&tgt, 0, sizeof(tgt));
    memset(&sc, 0, sizeof(sc)); memset(&tc, 0, sizeof(tc)); memset(&L, 0, sizeof(L));
    int P = prm.P;
    int stride = g_stride > 0 ? g_stride : 1;
    int herm = prm.approximation == APX_HERMITE;

    if (compute_type == CT_PC) {
        /* treedriver.c:350-627 */
        if ((rc = tree_construct(&src, &S, prm.msl, 0))) goto done;
        if ((rc = tree_construct(&tgt, &T, prm.mtl, 1))) goto done;
        g_stats.t_setup = wall() - t0; t0 = wall();
        if (herm) rc = clusters_sources_hermite(&sc, &src, &S, &prm);
        else rc = clusters_sources(&sc, &src, &S, &prm);
        if (rc) goto done;
        g_stats.t_upward = wall() - t0; t0 = wall();
        if ((rc = lists_make(&L, &src, &tgt, &prm))) goto done;
        g_stats.t_lists = wall() - t0; t0 = wall();
        stats_counts(&L, &tgt, &src, &prm);
        for (int i = 0; i < tgt.n; i++) {
            if (i % stride) continue;
            int bs = tgt.ibeg[i] - 1, nt = tgt.iend[i] - tgt.ibeg[i] + 1;
            for (int j = 0; j < L.ap[i].n; j++) {
                if (herm) k_pc_hermite(&prm, nt, P, bs, (size_t)P * L.ap[i].v[j], &T, &sc, pot);
                else k_pc(&prm, nt, P, bs, (size_t)P * L.ap[i].v[j], &T, &sc, pot);
                g_stats.far_done += (long long)nt * P;
            }
            for (int j = 0; j < L.pp[i].n; j++) {
                int nd = L.pp[i].v[j];
                k_pp(&prm, nt, src.iend[nd] - src.ibeg[nd] + 1, bs, src.ibeg[nd] - 1, &T, &S, pot);
                g_stats.near_done += (long long)nt * (src.iend[nd] - src.ibeg[nd] + 1);
            }
        }
        g_stats.t_compute = wall() - t0; t0 = wall();
    } else if (compute_type == CT_CP) {
        /* treedriver.c:114-339 */
        if ((rc = tree_construct(&tgt, &T, prm.mtl, 0))) goto done;
        if ((rc = tree_construct(&src, &S, prm.msl, 1))) goto done;
        if ((rc = clusters_targets(&tc, &tgt, &prm))) goto done;
        g_stats.t_setup = wall() - t0; t0 = wall();
        if ((rc = lists_make(&L, &tgt, &src, &prm))) goto done;
        g_stats.t_lists = wall() - t0; t0 = wall();
        stats_counts(&L, &src, &tgt, &prm);
        for (int i = 0; i < src.n; i++) {
            if (i % stride) continue;
            int bs = src.ibeg[i] - 1, ns = src.iend[i] - src.ibeg[i] + 1;
            for (int j = 0; j < L.ap[i].n; j++) {
                if (herm) k_cp_hermite(&prm, ns, P, bs, (size_t)P * L.ap[i].v[j], &S, &tc);
                else k_cp(&prm, ns, P, bs, (size_t)P * L.ap[i].v[j], &S, &tc);
                g_stats.far_done += (long long)ns * P;
            }
            for (int j = 0; j < L.pp[i].n; j++) {
                int nd = L.pp[i].v[j];
                k_pp(&prm, tgt.iend[nd] - tgt.ibeg[nd] + 1, ns, tgt.ibeg[nd] - 1, bs, &T, &S, pot);
                g_stats.near_done += (long long)ns * (tgt.iend[nd] - tgt.ibeg[nd] + 1);
            }
        }
        g_stats.t_compute = wall() - t0; t0 = wall();
        if (herm) downpass_hermite(&tgt, &tc, &T, &prm, pot);
        else downpass(&tgt, &tc, &T, &prm, pot);
        g_stats.t_downpass = wall() - t0; t0 = wall();
    } else {
        /* treedriver.c:638-967 */
        if ((rc = tree_construct(&src, &S, prm.msl, 0))) goto done;
        if ((rc = tree_construct(&tgt, &T, prm.mtl, 0))) goto done;
        g_stats.t_setup = wall() - t0; t0 = wall();
        if ((rc = clusters_sources(&sc, &src, &S, &prm))) goto done;
        if ((rc = clusters_targets(&tc, &tgt, &prm))) goto done;
        g_stats.t_upward = wall() - t0; t0 = wall();
        if ((rc = lists_make(&L, &src, &tgt, &prm))) goto done;
        g_stats.t_lists = wall() - t0; t0 = wall();
        stats_counts(&L, &tgt, &src, &prm);
        /* interaction_compute_cc.c:87-1205 */
        for (int i = 0; i < tgt.n; i++) {
            if (i % stride) continue;
            int ts = tgt.ibeg[i] - 1, nt = tgt.iend[i] - tgt.ibeg[i] + 1;
            size_t tcs = (size_t)P * i;
            for (int j = 0; j < L.ap[i].n; j++)
                k_cc(&prm, P, (size_t)P * L.ap[i].v[j], tcs, &sc, &tc);
            for (int j = 0; j < L.spc[i].n; j++)
                k_pc(&prm, nt, P, ts, (size_t)P * L.spc[i].v[j], &T, &sc, pot);
            g_stats.far_done += (long long)L.ap[i].n * P * P + (long long)L.spc[i].n * nt * P;
            for (int j = 0; j < L.tcp[i].n; j++) {
                int nd = L.tcp[i].v[j];
                k_cp(&prm, src.iend[nd] - src.ibeg[nd] + 1, P, src.ibeg[nd] - 1, tcs, &S, &tc);
                g_stats.far_done += (long long)P * (src.iend[nd] - src.ibeg[nd] + 1);
            }
            for (int j = 0; j < L.pp[i].n; j++) {
                int nd = L.pp[i].v[j];
                k_pp(&prm, nt, src.iend[nd] - src.ibeg[nd] + 1, ts, src.ibeg[nd] - 1, &T, &S, pot);
                g_stats.near_done += (long long)nt * (src.iend[nd] - src.ibeg[nd] + 1);
            }
        }
        g_stats.t_compute = wall() - t0; t0 = wall();
        downpass(&tgt, &tc, &T, &prm, pot);
        g_stats.t_downpass = wall() - t0; t0 = wall();
    }
    g_stats.source_nodes = src.n; g_stats.target_nodes = tgt.n;
    g_stats.source_leaves = src.numleaves; g_stats.target_leaves = tgt.numleaves;
    g_stats.max_depth_source = src.max_depth; g_stats.max_depth_target = tgt.max_depth;
    g_stats.batches = compute_type == CT_PC ? tgt.n : (compute_type == CT_CP ? src.n : 0);

    /* treedriver.c:589-592: correction, then restore the caller's order */
    ss_correction(&prm, &T, pot);
    for (int i = 0; i < M; i++) out[T.order[i] - 1] = pot[i];
    g_stats.t_finish = wall() - t0;

done:
    lists_free(&L);
    clus_free(&sc); clus_free(&tc);
    tree_free(&src); tree_free(&tgt);
    free(pot);
done0:
    parts_free(&S); parts_free(&T);
    return rc;
}

/* ------------------------------------------------------------------ */
/* Far-field pairs at their own degree (the product's variable-degree   */
/* mode).  A pair's approximation is formed for that pair alone: source */
/* charges of degree p from the source node's particles (anterp, as for */
/* a leaf), a target grid of degree p for the target node (grid values  */
/* by the CP/CC kernels, then the interpolant at the target node's      */
/* particles).  Pair order: PC per target batch, CP per source batch,  */
/* CC per target node its CC, then PC, then CP entries.                 */
/* ------------------------------------------------------------------ */

static int src_cluster_deg(const Prm *q, const Tree *src, int s, const Parts *S, const double *qw, Clus *c)
{
    int sub = q->singularity == SNG_SUB;
    int e = clus_alloc(c, 1, q->P, sub);
    if (e) return e;
    int b = src->ibeg[s] - 1, m = src->iend[s] - src->ibeg[s] + 1;
    return anterp(q->degree, src->xmin[s], src->xmax[s], src->ymin[s], src->ymax[s], src->zmin[s],
                  src->zmax[s], m, S->x + b, S->y + b, S->z + b, sub ? qw + b : S->q + b,
                  sub ? S->w + b : NULL, c->x, c->y, c->z, c->q, sub ? c->w : NULL);
}

static int tgt_grid_deg(const Prm *q, const Tree *tgt, int t, Clus *g)
{
    int e = clus_alloc(g, 1, q->P, q->singularity == SNG_SUB);
    if (e) return e;
    int n = q->degree, nl = n + 1;
    double nodeX[32], nodeY[32], nodeZ[32];
    cheb_nodes(n, tgt->xmin[t], tgt->xmax[t], nodeX);
    cheb_nodes(n, tgt->ymin[t], tgt->ymax[t], nodeY);
    cheb_nodes(n, tgt->zmin[t], tgt->zmax[t], nodeZ);
    for (int j = 0; j < q->P; j++) {
        int k1 = j % nl, k2 = (j / nl) % nl, k3 = j / (nl * nl);
        g->x[j] = nodeX[k1]; g->y[j] = nodeY[k2]; g->z[j] = nodeZ[k3];
    }
    return 0;
}

static int far_pair(const Prm *prm, int kind, int p, const Tree *src, int s, const Parts *S,
                    const double *qw, const Tree *tgt, int t, const Parts *T, double *pot)
{
    Prm q = *prm;
    q.degree = p;
    q.P = (p + 1) * (p + 1) * (p + 1);
    int sub = q.singularity == SNG_SUB, e = 0;
    int ts = tgt->ibeg[t] - 1, nt = tgt->iend[t] - tgt->ibeg[t] + 1;
    Clus c, g;
    memset(&c, 0, sizeof(c)); memset(&g, 0, sizeof(g));
    if (kind != 2 && (e = src_cluster_deg(&q, src, s, S, qw, &c))) goto out;
    if (kind == 1) {
        k_pc(&q, nt, q.P, ts, 0, T, &c, pot);
        goto out;
    }
    if ((e = tgt_grid_deg(&q, tgt, t, &g))) goto out;
    if (kind == 2) k_cp(&q, src->iend[s] - src->ibeg[s] + 1, q.P, src->ibeg[s] - 1, 0, S, &g);
    else k_cc(&q, q.P, 0, 0, &c, &g);
    interp_eval(p, tgt->xmin[t], tgt->xmax[t], tgt->ymin[t], tgt->ymax[t], tgt->zmin[t], tgt->zmax[t],
                g.q, sub ? g.w : NULL, nt, T->x + ts, T->y + ts, T->z + ts, sub ? T->q + ts : NULL,
                pot + ts, NULL);
out:
    clus_free(&c);
    clus_free(&g);
    return e;
}

int orc_treecode_vardeg(int M, int N,
                        const double *tx, const double *ty, const double *tz, const double *tq,
                        const double *sx, const double *sy, const double *sz, const double *sq,
                        const double *sw, double *out,
                        int kernel, int nkp, const double *kp,
                        int singularity, int approximation, int compute_type,
                        double theta, int degree, int msl, int mtl,
                        double size_check, double beta, int npairs, const int *pair_degree)
{
    memset(&g_stats, 0, sizeof(g_stats));
    if (M <= 0 || N <= 0 || !out || !tx || !ty || !tz || !sx || !sy || !sz || !sq) return ORC_ERR_ARGS;
    memset(out, 0, (size_t)M * sizeof(double));
    Prm prm;
    memset(&prm, 0, sizeof(prm));
    prm.kernel = kernel; prm.singularity = singularity; prm.approximation = approximation;
    prm.compute_type = compute_type;
    for (int i = 0; i < nkp && i < 4; i++) prm.kp[i] = kp[i];
    orc_params_resolve(compute_type, approximation, theta, degree, msl, mtl, size_check, beta,
                       &prm.theta, &prm.degree, &prm.size_check_factor, &prm.msl, &prm.mtl);
    prm.P = (prm.degree + 1) * (prm.degree + 1) * (prm.degree + 1);
    /* per-pair degrees are a Lagrange extension */
    if (!supported(&prm) || approximation != APX_LAGRANGE) return ORC_ERR_UNSUPPORTED;
    if (singularity == SNG_SUB && (!sw || !tq)) return ORC_ERR_ARGS;
    int sub = singularity == SNG_SUB;

    int rc = 0;
    Parts S, T;
    memset(&S, 0, sizeof(S)); memset(&T, 0, sizeof(T));
    if ((rc = parts_copy(&S, N, sx, sy, sz, sq, sw, 1))) goto done0;
    if ((rc = parts_copy(&T, M, tx, ty, tz, tq, NULL, 0))) goto done0;
    double *pot = calloc((size_t)M, sizeof(double));
    double *qw = calloc((size_t)N, sizeof(double));
    if (!pot || !qw) { free(pot); free(qw); rc = ORC_ERR_NOMEM; goto done0; }

    Tree src, tgt;
    Lists L;
    memset(&src, 0, sizeof(src)); memset(&tgt, 0, sizeof(tgt)); memset(&L, 0, sizeof(L));
    if (compute_type == CT_PC) {
        if ((rc = tree_construct(&src, &S, prm.msl, 0))) goto done;
        if ((rc = tree_construct(&tgt, &T, prm.mtl, 1))) goto done;
        if ((rc = lists_make(&L, &src, &tgt, &prm))) goto done;
    } else if (compute_type == CT_CP) {
        if ((rc = tree_construct(&tgt, &T, prm.mtl, 0))) goto done;
        if ((rc = tree_construct(&src, &S, prm.msl, 1))) goto done;
        if ((rc = lists_make(&L, &tgt, &src, &prm))) goto done;
    } else {
        if ((rc = tree_construct(&src, &S, prm.msl, 0))) goto done;
        if ((rc = tree_construct(&tgt, &T, prm.mtl, 0))) goto done;
        if ((rc = lists_make(&L, &src, &tgt, &prm))) goto done;
    }
    for (int i = 0; i < N; i++) qw[i] = sub ? S.q[i] * S.w[i] : S.q[i];  /* tree order */
    {
        long long count = 0;
        for (int i = 0; i < L.n; i++)
            count += L.ap[i].n + (compute_type == CT_CC ? L.spc[i].n + L.tcp[i].n : 0);
        if (count != npairs) { rc = ORC_ERR_ARGS; goto done; }
    }
    int k = 0;
    for (int i = 0; i < L.n && !rc; i++) {
        if (compute_type == CT_PC) {
            int bs = tgt.ibeg[i] - 1, nt = tgt.iend[i] - tgt.ibeg[i] + 1;
            for (int j = 0; j < L.ap[i].n && !rc; j++)
                rc = far_pair(&prm, 1, pair_degree[k++], &src, L.ap[i].v[j], &S, qw, &tgt, i, &T, pot);
            for (int j = 0; j < L.pp[i].n; j++) {
                int nd = L.pp[i].v[j];
                k_pp(&prm, nt, src.iend[nd] - src.ibeg[nd] + 1, bs, src.ibeg[nd] - 1, &T, &S, pot);
            }
        } else if (compute_type == CT_CP) {
            int bs = src.ibeg[i] - 1, ns = src.iend[i] - src.ibeg[i] + 1;
            for (int j = 0; j < L.ap[i].n && !rc; j++)
                rc = far_pair(&prm, 2, pair_degree[k++], &src, i, &S, qw, &tgt, L.ap[i].v[j], &T, pot);
            for (int j = 0; j < L.pp[i].n; j++) {
                int nd = L.pp[i].v[j];
                k_pp(&prm, tgt.iend[nd] - tgt.ibeg[nd] + 1, ns, tgt.ibeg[nd] - 1, bs, &T, &S, pot);
            }
        } else {
            int ts = tgt.ibeg[i] - 1, nt = tgt.iend[i] - tgt.ibeg[i] + 1;
            for (int j = 0; j < L.ap[i].n && !rc; j++)
                rc = far_pair(&prm, 3, pair_degree[k++], &src, L.ap[i].v[j], &S, qw, &tgt, i, &T, pot);
            for (int j = 0; j < L.spc[i].n && !rc; j++)
                rc = far_pair(&prm, 1, pair_degree[k++], &src, L.spc[i].v[j], &S, qw, &tgt, i, &T, pot);
            for (int j = 0; j < L.tcp[i].n && !rc; j++)
                rc = far_pair(&prm, 2, pair_degree[k++], &src, L.tcp[i].v[j], &S, qw, &tgt, i, &T, pot);
            for (int j = 0; j < L.pp[i].n; j++) {
                int nd = L.pp[i].v[j];
                k_pp(&prm, nt, src.iend[nd] - src.ibeg[nd] + 1, ts, src.ibeg[nd] - 1, &T, &S, pot);
            }
        }
    }
    if (!rc) {
        ss_correction(&prm, &T, pot);
        for (int i = 0; i < M; i++) out[T.order[i] - 1] = pot[i];
    }
done:
    lists_free(&L);
    tree_free(&src); tree_free(&tgt);
    free(pot); free(qw);
done0:
    parts_free(&S); parts_free(&T);
    return rc;
}

int orc_direct(int M, int N,
               const double *tx, const double *ty, const double *tz, const double *tq,
               const double *sx, const double *sy, const double *sz, const double *sq,
               const double *sw, double *out,
               int kernel, int nkp, const double *kp, int singularity)
{
    if (M <= 0 || N <= 0 || !out) return ORC_ERR_ARGS;
    Prm prm;
    memset(&prm, 0, sizeof(prm));
    prm.kernel = kernel; prm.singularity = singularity;
    for (int i = 0; i < nkp && i < 4; i++) prm.kp[i] = kp[i];
    if ((kernel < KER_COULOMB || kernel > KER_REGCOULOMB) && !other_kernel(kernel)) return ORC_ERR_UNSUPPORTED;
    if (singularity != SNG_SKIP && singularity != SNG_SUB) return ORC_ERR_UNSUPPORTED;
    if (singularity == SNG_SUB && (!sw || !tq)) return ORC_ERR_ARGS;
    Parts S, T;
    memset(&S, 0, sizeof(S)); memset(&T, 0, sizeof(T));
    int rc = parts_copy(&S, N, sx, sy, sz, sq, sw, 1);
    if (!rc) rc = parts_copy(&T, M, tx, ty, tz, tq, NULL, 0);
    if (!rc) {
        memset(out, 0, (size_t)M * sizeof(double));
        /* interaction_compute_direct.c:62-218: one PP call over all pairs */
        k_pp(&prm, M, N, 0, 0, &T, &S, out);
        /* directdriver.c:117 */
        ss_correction(&prm, &T, out);
    }
    parts_free(&S); parts_free(&T);
    return rc;
}

void orc_last_stats(orc_stats *s) { *s = g_stats; }

void orc_set_sample_stride(int stride) { g_stride = stride > 0 ? stride : 1; }

/* Threads of the parallel loops (pair kernels over their targets or grid
 * points, upward pass and downpass over nodes); each target's or grid
 * point's sum keeps its serial order, so results do not depend on it. */
void orc_set_threads(int n)
{
#ifdef _OPENMP
    omp_set_num_threads(n > 0 ? n : 1);
#else
    (void)n;
#endif
}

int orc_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

int orc_source_clusters(int n, const double *x, const double *y, const double *z,
                        const double *q, const double *w, int degree, int max_leaf,
                        int singularity, double *cx, double *cy, double *cz,
                        double *q_out, double *w_out)
{
    return orc_source_clusters_apx(n, x, y, z, q, w, degree, max_leaf, singularity, APX_LAGRANGE, cx, cy,
                                   cz, q_out, w_out);
}

int orc_source_clusters_apx(int n, const double *x, const double *y, const double *z,
                            const double *q, const double *w, int degree, int max_leaf,
                            int singularity, int approximation, double *cx, double *cy, double *cz,
                            double *q_out, double *w_out)
{
    if (n <= 0 || degree < 1 || degree > 30 || max_leaf < 1) return ORC_ERR_ARGS;
    if (approximation != APX_LAGRANGE && approximation != APX_HERMITE) return ORC_ERR_ARGS;
    int planes = approximation == APX_HERMITE ? 8 : 1;
    Prm prm;
    memset(&prm, 0, sizeof(prm));
    prm.singularity = singularity; prm.degree = degree; prm.approximation = approximation;
    prm.P = (degree + 1) * (degree + 1) * (degree + 1);
    Parts S;
    memset(&S, 0, sizeof(S));
    int rc = parts_copy(&S, n, x, y, z, q, w, 1);
    Tree t;
    Clus c;
    memset(&t, 0, sizeof(t)); memset(&c, 0, sizeof(c));
    if (!rc) rc = tree_construct(&t, &S, max_leaf, 0);
    if (!rc && q_out)
        rc = planes == 8 ? clusters_sources_hermite(&c, &t, &S, &prm) : clusters_sources(&c, &t, &S, &prm);
    if (!rc) {
        rc = t.n;
        if (q_out) {
            size_t tot = (size_t)t.n * prm.P;
            memcpy(q_out, c.q, tot * planes * sizeof(double));
            if (cx) memcpy(cx, c.x, tot * sizeof(double));
            if (cy) memcpy(cy, c.y, tot * sizeof(double));
            if (cz) memcpy(cz, c.z, tot * sizeof(double));
            if (w_out && c.w) memcpy(w_out, c.w, tot * planes * sizeof(double));
        }
    }
    clus_free(&c); tree_free(&t); parts_free(&S);
    return rc;
}
