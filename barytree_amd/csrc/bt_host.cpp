// bt_host.cpp -- run parameters, tree/batch construction and interaction
// lists on the host.
//
// The tree must reproduce the reference's node membership and bounding
// boxes exactly, because every later decision (MAC acceptance, Chebyshev
// nodes, exact-node hits) depends on them bit for bit.  The split rules and
// the in-place partition therefore follow src/tree/tree_linked_list.c:16-280
// and src/tree/partition.c:14-391 step for step.
#include "bt_host.h"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace bt {

namespace {
thread_local int g_beta_leaf_mode = 0;
}

void set_beta_leaf_mode(int mode) { g_beta_leaf_mode = mode; }
int beta_leaf_mode() { return g_beta_leaf_mode; }

void params_setup(Params &p, int kernel, int num_kp, const double *kp, int approximation,
                  int singularity, int compute_type, double theta, int degree, int msl, int mtl,
                  double size_check, double beta, int verbosity)
{
    p.kernel = kernel;
    p.singularity = singularity;
    p.approximation = approximation;
    p.compute_type = compute_type;
    p.num_kp = num_kp;
    for (int i = 0; i < 4; i++) p.kp[i] = (kp && i < num_kp) ? kp[i] : 0.0;
    p.verbosity = verbosity;
    if (beta < 0 || beta > 1) {
        // run_params.c:37-49
        p.beta = -1;
        p.theta = theta;
        p.degree = degree;
        p.size_check_factor = size_check;
        p.max_source_leaf = msl;
        p.max_target_leaf = mtl;
    } else {
        // run_params.c:51-120 (leaf sizes of the CPU build by default, so
        // that results agree with the reference CPU run; of the GPU build
        // on request)
        p.beta = beta;
        double tmin = 0.55, tmax = 0.95, nmin = 1, nmax = 12, es = 2, et = 3;
        p.size_check_factor = 1.0;
        if (compute_type == C_PC || compute_type == C_CP) {
            if (approximation != A_LAGRANGE) {
                p.size_check_factor = 4.0;
                nmax = 9;
            }
        } else {
            es = 1.75;
        }
        p.theta = tmax - (tmax - tmin) * std::pow(beta, es);
        p.degree = (int)(nmax - (nmax - nmin) * std::pow(1. - beta, et));
        if (g_beta_leaf_mode == 1) { p.max_source_leaf = 3000; p.max_target_leaf = 3000; }
        else if (compute_type == C_PC) { p.max_source_leaf = 50; p.max_target_leaf = 5; }
        else if (compute_type == C_CP) { p.max_source_leaf = 5; p.max_target_leaf = 50; }
        else { p.max_source_leaf = 50; p.max_target_leaf = 50; }
    }
    int nl = p.degree + 1;
    p.P = nl * nl * nl;
}

std::string params_check(const Params &p)
{
    const bool other = p.kernel == K_REG_YUKAWA || p.kernel == K_ATAN || p.kernel == K_SIN_OVER_R ||
                       p.kernel == K_MQ || p.kernel == K_RBS_U || p.kernel == K_RBS_V || p.kernel == K_USER;
    if (!coulomb_family(p.kernel) && !other)
        return "INVALID KERNEL (supported: COULOMB, YUKAWA, REGULARIZED_COULOMB, REGULARIZED_YUKAWA, "
               "ATAN, SIN_OVER_R, MQ, RBS_U, RBS_V, USER)";
    if (p.singularity != S_SKIPPING && p.singularity != S_SUBTRACTION)
        return "INVALID CHOICE OF SINGULARITY";
    if (other) {
        // the combinations the dispatchers compute as written
        // (interaction_compute_pc.c:259-435, _cp.c:255-384, _cc.c:275-403)
        // Hermite exists for sin-over-r in PC and CP only
        // (interaction_compute_pc.c:338-362, interaction_compute_cp.c:304-328)
        if (p.approximation == A_HERMITE && !(p.kernel == K_SIN_OVER_R && p.compute_type != C_CC))
            return "NOT SET UP FOR HERMITE WITH THIS KERNEL";
        if ((p.kernel == K_ATAN || p.kernel == K_MQ) && p.compute_type != C_PC)
            return "ATAN AND MQ ARE SET UP FOR PARTICLE-CLUSTER ONLY";
        if (p.singularity == S_SUBTRACTION && !(p.kernel == K_REG_YUKAWA && p.compute_type == C_PC))
            return "SUBTRACTION WITH THIS KERNEL IS SET UP FOR PC REGULARIZED_YUKAWA ONLY";
        const int need = (p.kernel == K_REG_YUKAWA || p.kernel == K_ATAN || p.kernel == K_MQ) ? 2 : 1;
        if (p.num_kp < need) return "this kernel needs more kernel parameters";
    }
    if (p.approximation != A_LAGRANGE && p.approximation != A_HERMITE) return "INVALID CHOICE OF APPROXIMATION";
    // Hermite: PC with either singularity, CP with skipping, no CC
    // (interaction_compute_cc.c:146, interaction_compute_cp.c:133, 187,
    // interaction_compute_downpass.c:126-128)
    if (p.approximation == A_HERMITE && p.compute_type == C_CC) return "NOT SET UP FOR CC HERMITE";
    if (p.approximation == A_HERMITE && p.compute_type == C_CP && p.singularity == S_SUBTRACTION)
        return "NOT SET UP FOR CP HERMITE SS";
    if (p.compute_type != C_PC && p.compute_type != C_CP && p.compute_type != C_CC)
        return "INVALID COMPUTE TYPE";
    // interaction_compute_cp.c:218, interaction_compute_cc.c:241
    if (p.kernel == K_REG_COULOMB && p.singularity == S_SUBTRACTION && p.compute_type != C_PC)
        return "NOT SET UP FOR CP/CC REGULARIZED COULOMB SS";
    if (p.degree < 1 || p.degree > 12) return "interpolation degree must be in [1, 12]";
    if (p.max_source_leaf < 1 || p.max_target_leaf < 1) return "leaf sizes must be positive";
    if (p.kernel == K_YUKAWA && p.num_kp < 1) return "YUKAWA needs kappa";
    if (p.singularity == S_SUBTRACTION && p.num_kp < 1) return "SUBTRACTION needs a kernel parameter";
    if (p.kernel == K_REG_COULOMB && p.singularity == S_SUBTRACTION && p.num_kp < 2)
        return "REGULARIZED_COULOMB SUBTRACTION needs (alpha, epsilon)";
    if (p.kernel == K_REG_COULOMB && p.num_kp < 1) return "REGULARIZED_COULOMB needs epsilon";
    return "";
}

int Tree::add_node()
{
    int i = n++;
    ibeg.push_back(0); iend.push_back(0); numpar.push_back(0); nchild.push_back(0);
    for (int k = 0; k < 8; k++) child.push_back(-1);
    parent.push_back(-1); level.push_back(0);
    xmin.push_back(0); xmax.push_back(0); ymin.push_back(0); ymax.push_back(0);
    zmin.push_back(0); zmax.push_back(0); xmid.push_back(0); ymid.push_back(0);
    zmid.push_back(0); radius.push_back(0);
    return i;
}

bool node_split(int numpar, int maxpar, NodeRec &r, int dv[3])
{
    // tree_linked_list.c:53-69
    const double xl = r.hi[0] - r.lo[0], yl = r.hi[1] - r.lo[1], zl = r.hi[2] - r.lo[2];
    for (int d = 0; d < 3; d++) r.mid[d] = (r.hi[d] + r.lo[d]) / 2.0;
    r.radius = std::sqrt(xl * xl + yl * yl + zl * zl) / 2.0;
    dv[0] = dv[1] = dv[2] = 0;
    if (!(numpar > maxpar)) return false;

    // tree_linked_list.c:81-91 and partition.c:178-286
    const int maxch = numpar < 2 * maxpar ? 2 : (numpar < 4 * maxpar ? 4 : 8);
    double crit = xl;
    if (crit < yl) crit = yl;
    if (crit < zl) crit = zl;
    crit = crit / std::sqrt(2.0);
    dv[0] = xl >= crit; dv[1] = yl >= crit; dv[2] = zl >= crit;
    const double len[3] = {xl, yl, zl};
    if (maxch == 4) {
        for (int d = 0; d < 3; d++) {
            const double a = len[d], o1 = len[(d + 1) % 3], o2 = len[(d + 2) % 3];
            if (a < o1 && a < o2) dv[d] = 0;
        }
        if (dv[0] + dv[1] + dv[2] == 3) dv[0] = 0;
    } else if (maxch == 2) {
        for (int d = 0; d < 3; d++) {
            const double a = len[d], o1 = len[(d + 1) % 3], o2 = len[(d + 2) % 3];
            if (a < o1 || a < o2) dv[d] = 0;
        }
        if (dv[0] + dv[1] + dv[2] == 3) { dv[0] = 0; dv[1] = 0; }
        if (dv[0] + dv[1] + dv[2] == 2) {
            if (dv[0] == 0) dv[1] = 0;
            else if (dv[1] == 0) dv[2] = 0;
            else dv[0] = 0;
        }
    }
    return true;
}

namespace {

// Splits [b, e) (0-based, half open) around val along key array `a`,
// permuting every particle array alike.  Returns the first index of the
// upper part.  This is the reference's pivot-hole partition
// (partition.c:94-175, which works on the 1-based inclusive range [b+1, e]):
// the first element is lifted out, its key slot holds val as a sentinel,
// the two cursors move the hole between the ends, and the lifted element
// lands where the cursors meet.
struct Splitter {
    Particles &p;
    bool with_w;
    inline void mv(int dst, int src)
    {
        p.x[dst] = p.x[src]; p.y[dst] = p.y[src]; p.z[dst] = p.z[src]; p.q[dst] = p.q[src];
        if (with_w) p.w[dst] = p.w[src];
        p.order[dst] = p.order[src];
    }
    int split(int dim, int b, int e, double val)
    {
        double *a = dim == 0 ? p.x.data() : (dim == 1 ? p.y.data() : p.z.data());
        int cnt = e - b;
        if (cnt >= 2) {
            const double hx = p.x[b], hy = p.y[b], hz = p.z[b], hq = p.q[b];
            const double hw = with_w ? p.w[b] : 0.0;
            const int ho = p.order[b];
            const double hk = a[b];
            a[b] = val;
            int lo = b, hi = e - 1;  // 0-based cursors
            while (lo != hi) {
                while (lo < hi && val < a[hi]) hi--;
                if (lo != hi) mv(lo, hi);
                while (lo < hi && val >= a[lo]) lo++;
                if (lo != hi) mv(hi, lo);
            }
            p.x[lo] = hx; p.y[lo] = hy; p.z[lo] = hz; p.q[lo] = hq;
            if (with_w) p.w[lo] = hw;
            p.order[lo] = ho;
            return hk > val ? lo : lo + 1;
        }
        if (cnt == 1) return a[b] < val ? e : b;
        return b;
    }
};

// Ranges at least this long are split or descended in parallel tasks.
constexpr int PAR_RANGE = 1 << 17;

struct Builder {
    Particles &p;
    Splitter sp;
    int maxpar;
    bool leaves_only;
    // a split left the whole range in one child: more coincident particles
    // than maxpar, which no split separates (the reference's recursion never
    // ends there)
    bool stuck = false;

    void bounds(int b, int e, double lo[3], double hi[3]) const
    {
        double xn = p.x[b], xx = p.x[b], yn = p.y[b], yx = p.y[b], zn = p.z[b], zx = p.z[b];
        if (e - b >= PAR_RANGE) {
#pragma omp parallel for reduction(min : xn, yn, zn) reduction(max : xx, yx, zx) schedule(static)
            for (int i = b + 1; i < e; i++) {
                xn = std::min(xn, p.x[i]); xx = std::max(xx, p.x[i]);
                yn = std::min(yn, p.y[i]); yx = std::max(yx, p.y[i]);
                zn = std::min(zn, p.z[i]); zx = std::max(zx, p.z[i]);
            }
        } else {
            for (int i = b + 1; i < e; i++) {
                if (xn > p.x[i]) xn = p.x[i];
                if (xx < p.x[i]) xx = p.x[i];
                if (yn > p.y[i]) yn = p.y[i];
                if (yx < p.y[i]) yx = p.y[i];
                if (zn > p.z[i]) zn = p.z[i];
                if (zx < p.z[i]) zx = p.z[i];
            }
        }
        lo[0] = xn; hi[0] = xx; lo[1] = yn; hi[1] = yx; lo[2] = zn; hi[2] = zx;
    }

    // Builds the subtree of [b, e) into `out` (preorder, local ids).  For
    // batches (leaves_only) only leaves are appended.  The min and max are
    // order independent and every split works in place on its own range, so
    // the result is the serial result whatever runs in parallel.
    void build(int b, int e, int level, std::vector<NodeRec> &out, int &max_depth)
    {
        NodeRec r{};
        r.b = b; r.e = e; r.level = level; r.nchild = 0;
        bounds(b, e, r.lo, r.hi);
        int dv[3];
        const bool leaf = !node_split(e - b, maxpar, r, dv);
        const int numpar = e - b;
        max_depth = std::max(max_depth, level + 1);
        if (!leaves_only || leaf) out.push_back(r);
        if (leaf) return;
        const size_t self = out.size() - 1;

        int rb[8], re[8], nr = 1;
        rb[0] = b; re[0] = e;
        for (int d = 0; d < 3; d++) {
            if (!dv[d]) continue;
            int m[8];
            // the nr ranges are disjoint: split them concurrently
#pragma omp taskloop shared(m, rb, re) if (numpar >= PAR_RANGE && nr > 1) grainsize(1)
            for (int i = 0; i < nr; i++) m[i] = sp.split(d, rb[i], re[i], r.mid[d]);
            for (int i = 0; i < nr; i++) {
                rb[nr + i] = m[i]; re[nr + i] = re[i]; re[i] = m[i];
            }
            nr *= 2;
        }
        // children: independent ranges, built into their own vectors
        std::vector<NodeRec> sub[8];
        int depth[8];
        int nc = 0, cb[8], ce[8];
        for (int i = 0; i < nr; i++)
            if (rb[i] < re[i]) { cb[nc] = rb[i]; ce[nc] = re[i]; nc++; }
        if (nc == 1) {
#pragma omp atomic write
            stuck = true;
            return;
        }
        for (int c = 0; c < nc; c++) {
            depth[c] = 0;
#pragma omp task shared(sub, depth) if (ce[c] - cb[c] >= PAR_RANGE)
            build(cb[c], ce[c], level + 1, sub[c], depth[c]);
        }
#pragma omp taskwait
        // splice in preorder: self, child 0's subtree, child 1's subtree, ...
        for (int c = 0; c < nc; c++) {
            max_depth = std::max(max_depth, depth[c]);
            const int off = (int)out.size();
            if (!leaves_only) out[self].child[out[self].nchild++] = off;
            for (NodeRec n : sub[c]) {
                for (int k = 0; k < n.nchild; k++) n.child[k] += off;
                out.push_back(n);
            }
        }
    }
};

void levels_rec(Tree &t, int idx, int level)
{
    if (t.nchild[idx] == 0) {
        t.leaves.push_back(idx);
        return;
    }
    t.levels[level].push_back(idx);
    for (int k = 0; k < t.nchild[idx]; k++) levels_rec(t, t.child[8 * idx + k], level + 1);
}

}  // namespace

void finish_tree(Tree &t, const std::vector<NodeRec> &nodes, int max_depth, bool leaves_only)
{
    t.max_depth = max_depth;
    for (const NodeRec &r : nodes) {
        const int i = t.add_node();
        t.ibeg[i] = r.b; t.iend[i] = r.e; t.numpar[i] = r.e - r.b; t.level[i] = r.level;
        t.xmin[i] = r.lo[0]; t.ymin[i] = r.lo[1]; t.zmin[i] = r.lo[2];
        t.xmax[i] = r.hi[0]; t.ymax[i] = r.hi[1]; t.zmax[i] = r.hi[2];
        t.xmid[i] = r.mid[0]; t.ymid[i] = r.mid[1]; t.zmid[i] = r.mid[2]; t.radius[i] = r.radius;
        t.nchild[i] = r.nchild;
        for (int k = 0; k < r.nchild; k++) t.child[8 * i + k] = r.child[k];
    }
    for (int i = 0; i < t.n; i++)
        for (int k = 0; k < t.nchild[i]; k++) t.parent[t.child[8 * i + k]] = i;
    if (leaves_only) {
        for (int i = 0; i < t.n; i++) t.leaves.push_back(i);
        return;
    }
    t.levels.assign(t.max_depth, {});
    levels_rec(t, 0, 0);
}

std::string build_trees(const std::vector<TreeJob> &jobs)
{
    const size_t nj = jobs.size();
    std::vector<std::vector<NodeRec>> nodes(nj);
    std::vector<int> depth(nj, 0);
    std::vector<Builder> blds;
    for (const TreeJob &j : jobs) {
        *j.t = Tree();
        blds.push_back(Builder{*j.p, Splitter{*j.p, !j.p->w.empty()}, j.max_per_leaf, j.leaves_only});
    }
#pragma omp parallel
#pragma omp single
    {
        for (size_t k = 0; k < nj; k++) {
            if (jobs[k].p->n <= 0) continue;
#pragma omp task shared(nodes, depth, blds)
            blds[k].build(0, jobs[k].p->n, 0, nodes[k], depth[k]);
        }
#pragma omp taskwait
    }
    for (size_t k = 0; k < nj; k++)
        if (blds[k].stuck) return COINCIDENT_ERROR;
    for (size_t k = 0; k < nj; k++)
        if (jobs[k].p->n > 0) finish_tree(*jobs[k].t, nodes[k], depth[k], jobs[k].leaves_only);
    return "";
}

std::string build_tree(Tree &t, Particles &p, int max_per_leaf, bool leaves_only)
{
    return build_trees({TreeJob{&t, &p, max_per_leaf, leaves_only}});
}

namespace {

// interaction_lists.c:464-533
void pc_rec(Lists &L, int key, const Tree &s, int node, double br, double bx, double by, double bz,
            const Params &prm)
{
    const double tx = bx - s.xmid[node];
    const double ty = by - s.ymid[node];
    const double tz = bz - s.zmid[node];
    const double dist = std::sqrt(tx * tx + ty * ty + tz * tz);
    if (((s.radius[node] + br) < dist * prm.theta) && (s.radius[node] != 0.00) &&
        (prm.size_check_factor * prm.P < s.numpar[node])) {
        L.approx[key].push_back(node);
        return;
    }
    if (s.nchild[node] == 0) {
        L.pp[key].push_back(node);
        return;
    }
    for (int i = 0; i < s.nchild[node]; i++) pc_rec(L, key, s, s.child[8 * node + i], br, bx, by, bz, prm);
}

// interaction_lists.c:537-748
void cc_rec(Lists &L, const Tree &s, int sn, const Tree &t, int tn, const Params &prm, int size_check)
{
    const double tx = t.xmid[tn] - s.xmid[sn];
    const double ty = t.ymid[tn] - s.ymid[sn];
    const double tz = t.zmid[tn] - s.zmid[sn];
    const double dist = std::sqrt(tx * tx + ty * ty + tz * tz);
    if ((s.radius[sn] + t.radius[tn]) < dist * prm.theta) {
        const bool ssmall = s.numpar[sn] <= size_check, tsmall = t.numpar[tn] <= size_check;
        if (ssmall && tsmall) L.pp[tn].push_back(sn);
        else if (ssmall) L.cc_cp[tn].push_back(sn);
        else if (tsmall) L.cc_pc[tn].push_back(sn);
        else L.approx[tn].push_back(sn);
        return;
    }
    const bool tleaf = t.nchild[tn] == 0, sleaf = s.nchild[sn] == 0;
    if (tleaf && sleaf) {
        L.pp[tn].push_back(sn);
    } else if (sleaf || (!tleaf && s.numpar[sn] < t.numpar[tn])) {
        for (int i = 0; i < t.nchild[tn]; i++) cc_rec(L, s, sn, t, t.child[8 * tn + i], prm, size_check);
    } else {
        for (int i = 0; i < s.nchild[sn]; i++) cc_rec(L, s, s.child[8 * sn + i], t, tn, prm, size_check);
    }
}

}  // namespace

void make_pc_lists(Lists &L, const Tree &src, const Tree &key, const Params &prm)
{
    L = Lists();
    L.pp.assign(key.n, {});
    L.approx.assign(key.n, {});
    for (int i = 0; i < key.n; i++)
        pc_rec(L, i, src, 0, key.radius[i], key.xmid[i], key.ymid[i], key.zmid[i], prm);
}

void make_cc_lists(Lists &L, const Tree &src, const Tree &tgt, const Params &prm)
{
    L = Lists();
    L.pp.assign(tgt.n, {});
    L.approx.assign(tgt.n, {});
    L.cc_pc.assign(tgt.n, {});
    L.cc_cp.assign(tgt.n, {});
    const int size_check = (int)(prm.size_check_factor * prm.P);
    cc_rec(L, src, 0, tgt, 0, prm, size_check);
}

void hermite_weight_table(const Tree &t, int degree, std::vector<double> &tab)
{
    const int nl = degree + 1;
    tab.assign((size_t)t.n * 3 * nl, 0.0);
    std::vector<double> ww(nl);
    for (int i = 0; i < nl; i++) {
        const double xx = i * M_PI / degree;
        ww[i] = -std::cos(xx) / (2 * std::sin(xx) * std::sin(xx));
    }
    ww[0] = 0.25 * (degree * degree / 3.0 + 1.0 / 6.0);
    ww[degree] = -ww[0];
    for (int nd = 0; nd < t.n; nd++) {
        const double len[3] = {t.xmax[nd] - t.xmin[nd], t.ymax[nd] - t.ymin[nd], t.zmax[nd] - t.zmin[nd]};
        double *o = &tab[(size_t)nd * 3 * nl];
        for (int d = 0; d < 3; d++)
            for (int j = 0; j < nl; j++) o[d * nl + j] = -4.0 * ww[j] / len[d];
    }
}

void node_table(const Tree &t, int degree, std::vector<double> &tab)
{
    const int nl = degree + 1;
    tab.assign((size_t)t.n * 3 * nl, 0.0);
    std::vector<double> tt(nl);
    for (int i = 0; i < nl; i++) tt[i] = std::cos(i * M_PI / degree);
    for (int nd = 0; nd < t.n; nd++) {
        const double lo[3] = {t.xmin[nd], t.ymin[nd], t.zmin[nd]};
        const double hi[3] = {t.xmax[nd], t.ymax[nd], t.zmax[nd]};
        double *o = &tab[(size_t)nd * 3 * nl];
        for (int d = 0; d < 3; d++)
            for (int i = 0; i < nl; i++) o[d * nl + i] = lo[d] + (tt[i] + 1.0) / 2.0 * (hi[d] - lo[d]);
    }
}

}  // namespace bt
