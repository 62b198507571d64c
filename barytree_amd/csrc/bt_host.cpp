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
        // run_params.c:51-120 (the CPU build's leaf sizes, so that results
        // agree with the reference CPU run)
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
        if (compute_type == C_PC) { p.max_source_leaf = 50; p.max_target_leaf = 5; }
        else if (compute_type == C_CP) { p.max_source_leaf = 5; p.max_target_leaf = 50; }
        else { p.max_source_leaf = 50; p.max_target_leaf = 50; }
    }
    int nl = p.degree + 1;
    p.P = nl * nl * nl;
}

std::string params_check(const Params &p)
{
    if (p.kernel != K_COULOMB && p.kernel != K_YUKAWA && p.kernel != K_REG_COULOMB)
        return "INVALID KERNEL (supported: COULOMB, YUKAWA, REGULARIZED_COULOMB)";
    if (p.singularity != S_SKIPPING && p.singularity != S_SUBTRACTION)
        return "INVALID CHOICE OF SINGULARITY";
    if (p.approximation == A_HERMITE) return "HERMITE approximation is not implemented on this path";
    if (p.approximation != A_LAGRANGE) return "INVALID CHOICE OF APPROXIMATION";
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

struct Builder {
    Tree &t;
    Particles &p;
    Splitter sp;
    int maxpar;
    bool leaves_only;

    // returns node id, or -1 for an unrecorded interior node (batches)
    int build(int b, int e, int level, int parent)
    {
        const int numpar = e - b;
        double x_min = p.x[b], x_max = p.x[b], y_min = p.y[b], y_max = p.y[b];
        double z_min = p.z[b], z_max = p.z[b];
        for (int i = b + 1; i < e; i++) {
            if (x_min > p.x[i]) x_min = p.x[i];
            if (x_max < p.x[i]) x_max = p.x[i];
            if (y_min > p.y[i]) y_min = p.y[i];
            if (y_max < p.y[i]) y_max = p.y[i];
            if (z_min > p.z[i]) z_min = p.z[i];
            if (z_max < p.z[i]) z_max = p.z[i];
        }
        // tree_linked_list.c:53-69
        const double xl = x_max - x_min, yl = y_max - y_min, zl = z_max - z_min;
        const double x_mid = (x_max + x_min) / 2.0, y_mid = (y_max + y_min) / 2.0;
        const double z_mid = (z_max + z_min) / 2.0;
        const double radius = std::sqrt(xl * xl + yl * yl + zl * zl) / 2.0;
        const bool leaf = !(numpar > maxpar);
        if (level + 1 > t.max_depth) t.max_depth = level + 1;
        int idx = -1;
        if (!leaves_only || leaf) {
            idx = t.add_node();
            t.ibeg[idx] = b; t.iend[idx] = e; t.numpar[idx] = numpar; t.level[idx] = level;
            t.parent[idx] = parent;
            t.xmin[idx] = x_min; t.xmax[idx] = x_max; t.ymin[idx] = y_min; t.ymax[idx] = y_max;
            t.zmin[idx] = z_min; t.zmax[idx] = z_max;
            t.xmid[idx] = x_mid; t.ymid[idx] = y_mid; t.zmid[idx] = z_mid; t.radius[idx] = radius;
        }
        if (leaf) return idx;

        // tree_linked_list.c:81-91 and partition.c:178-286
        const int maxch = numpar < 2 * maxpar ? 2 : (numpar < 4 * maxpar ? 4 : 8);
        double crit = xl;
        if (crit < yl) crit = yl;
        if (crit < zl) crit = zl;
        crit = crit / std::sqrt(2.0);
        int dv[3] = {xl >= crit, yl >= crit, zl >= crit};
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
        int rb[8], re[8], nr = 1;
        rb[0] = b; re[0] = e;
        const double mid[3] = {x_mid, y_mid, z_mid};
        for (int d = 0; d < 3; d++) {
            if (!dv[d]) continue;
            for (int i = 0; i < nr; i++) {
                int m = sp.split(d, rb[i], re[i], mid[d]);
                rb[nr + i] = m; re[nr + i] = re[i]; re[i] = m;
            }
            nr *= 2;
        }
        for (int i = 0; i < nr; i++) {
            if (rb[i] < re[i]) {
                int c = build(rb[i], re[i], level + 1, idx);
                if (idx >= 0) t.child[8 * idx + t.nchild[idx]++] = c;
            }
        }
        return idx;
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

void build_tree(Tree &t, Particles &p, int max_per_leaf, bool leaves_only)
{
    t = Tree();
    if (p.n <= 0) return;
    Builder bld{t, p, Splitter{p, !p.w.empty()}, max_per_leaf, leaves_only};
    bld.build(0, p.n, 0, -1);
    if (leaves_only) {
        for (int i = 0; i < t.n; i++) t.leaves.push_back(i);
        return;
    }
    t.levels.assign(t.max_depth, {});
    levels_rec(t, 0, 0);
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
