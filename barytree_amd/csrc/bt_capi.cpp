// bt_capi.cpp -- extern "C" entry points of libBaryTree_gpu.so.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/BaryTreeInterface.h"
#include "bt_devmem.h"
#include "bt_learn.h"
#include "bt_model.h"
#include "bt_noalloc.h"
#include "bt_solver.h"

struct BaryTreePlan {
    bt::Plan plan;
};

struct BaryTreeDegreeModel {
    bt::DegreeModel m;
};

namespace {

thread_local std::string g_last_error;
thread_local bool g_have_stats = false;
thread_local BaryTreeStats g_stats;

int fail(const std::string &msg)
{
    g_last_error = msg;
    return -1;
}

void fill_stats(const bt::Plan &p, BaryTreeStats *s)
{
    std::memset(s, 0, sizeof(*s));
    const bt::Counts &c = p.counts();
    s->num_pp = c.num_pp; s->num_pc = c.num_pc; s->num_cp = c.num_cp; s->num_cc = c.num_cc;
    s->pp_ptwise = c.pp_ptwise; s->pc_ptwise = c.pc_ptwise; s->cp_ptwise = c.cp_ptwise;
    s->cc_ptwise = c.cc_ptwise; s->pc_replaced = c.pc_replaced; s->cp_replaced = c.cp_replaced;
    s->cc_replaced = c.cc_replaced;
    s->source_nodes = p.source_nodes(); s->target_nodes = p.target_nodes();
    s->interp_degree = p.params().degree; s->pts_per_cluster = p.params().P;
    s->theta = p.params().theta;
    s->max_per_source_leaf = p.params().max_source_leaf;
    s->max_per_target_leaf = p.params().max_target_leaf;
    s->near_tiles = p.num_near_tiles(); s->far_tiles = p.num_far_tiles();
    const bt::Timings &t = p.timings();
    s->host_copy_s = t.host_copy; s->host_tree_s = t.host_tree; s->host_lists_s = t.host_lists;
    s->host_work_s = t.host_work; s->h2d_s = t.h2d;
    s->gpu_upward_s = t.gpu_upward; s->gpu_near_s = t.gpu_near; s->gpu_far_s = t.gpu_far;
    s->gpu_down_s = t.gpu_down; s->gpu_finish_s = t.gpu_finish; s->gpu_total_s = t.gpu_total;
    s->num_systems = p.num_systems(); s->targets_per_thread = p.tpt();
    s->persistent = p.persistent() ? 1 : 0;
    s->num_far_pairs = (long long)p.far_pairs_list().size();
    s->variable_degree = p.variable_degree() ? 1 : 0;
    s->mean_pair_degree = p.mean_pair_degree();
    s->far_ptwise_variable = p.variable_far_evals();
    s->sym_ptwise = p.sym_pairs();
    s->gpu_near_sym_s = t.gpu_near_sym;
    s->sym_slots = p.sym_slots();
    s->model_fallback_pairs = p.variable_degree() ? p.model_fallback_pairs() : 0;
    s->sym_evals = p.sym_evals();
}

// Degrees from the model; a pair whose features lie outside the model's
// training range keeps the plan's fixed degree (the reference's method).
// The model's degree for every far-field pair of the plan (the fixed degree
// where the pair's features lie outside the model's training range).
std::string model_degrees(const bt::Plan &plan, const bt::DegreeModel &m, float delta, std::vector<int> &deg,
                          long long &fallback)
{
    std::vector<float> f;
    plan.pair_features(f);
    const int n = (int)plan.far_pairs_list().size();
    deg.assign(n, 0);
    std::string e = m.predict(n, f.data(), delta, deg.data(), nullptr);
    if (!e.empty()) return e;
    fallback = 0;
    for (int i = 0; i < n; i++)
        if (m.outside_range(f.data() + (size_t)i * btl::MLP_IN)) {
            deg[i] = plan.params().degree;
            fallback++;
        }
    return "";
}

std::string use_model(bt::Plan &plan, const bt::DegreeModel &m, float delta)
{
    std::vector<int> deg;
    long long fallback = 0;
    std::string e = model_degrees(plan, m, delta, deg, fallback);
    if (e.empty()) e = plan.set_pair_degrees(deg.data());
    if (e.empty()) plan.set_model_fallback_pairs(fallback);
    return e;
}

int plan_create(bt::Plan &plan, int M, int N, const double *tx, const double *ty, const double *tz,
                const double *tq, const double *sx, const double *sy, const double *sz,
                const double *sq, const double *sw, int kernel, int nkp, const double *kp,
                int singularity, int approximation, int compute_type, double theta, int degree,
                int msl, int mtl, double size_check, double beta, int verbosity, int precision,
                bool device = false)
{
    bt::Params prm;
    bt::params_setup(prm, kernel, nkp, kp, approximation, singularity, compute_type, theta, degree,
                     msl, mtl, size_check, beta, verbosity);
    if (precision != BARYTREE_PRECISION_FP64 && precision != BARYTREE_PRECISION_MIXED)
        return fail("unknown precision mode");
    prm.precision = precision;
    if (verbosity > 0) {
        std::printf("[BaryTree]\n[BaryTree] BaryTree has started.\n");
        std::printf("[BaryTree] RunParams: kernel=%d approximation=%d singularity=%d "
                    "compute_type=%d theta=%f degree=%d leaves=%d/%d size_check=%f beta=%f "
                    "precision=%d\n",
                    prm.kernel, prm.approximation, prm.singularity, prm.compute_type, prm.theta,
                    prm.degree, prm.max_source_leaf, prm.max_target_leaf, prm.size_check_factor,
                    prm.beta, prm.precision);
    }
    std::string e = plan.create(prm, M, N, tx, ty, tz, tq, sx, sy, sz, sq, sw, false, device);
    if (!e.empty()) return fail(e);
    return 0;
}

// Diagnostics after a solve, at the reference's verbosity levels
// (treedriver.c:53-86, 989-1136): > 0 start and end lines, > 1 sizes and
// the total time, > 2 the clusterwise and pointwise interaction counts,
// the point-wise interactions each kind of interaction accounts for, and
// the time of every phase.  One GPU, one rank: the per-rank minimum and
// maximum columns equal the totals.
void print_diagnostics(const BaryTreeStats &s, int M, int N, double total_s, int verbosity)
{
    if (verbosity <= 0) return;
    const char *bar = "|------------------------------------------------------------------------------|";
    if (verbosity > 1) {
        std::printf("[BaryTree] %d target and %d source particles on one GPU (one rank).\n", M, N);
        std::printf("[BaryTree] Total BaryTree time: %1.3f seconds.\n", total_s);
    }
    if (verbosity > 2) {
        auto table = [&](const char *title, const long long v[5]) {
            std::printf("[BaryTree]\n[BaryTree] %s:\n[BaryTree]\n[BaryTree] %s\n", title, bar);
            std::printf("[BaryTree] |       |             Cumulative       Minimum per rank       Maximum per rank |\n");
            const char *names[5] = {"Total", "PP   ", "PC   ", "CP   ", "CC   "};
            for (int i = 0; i < 5; i++) {
                std::printf("[BaryTree] %s\n", bar);
                std::printf("[BaryTree] | %s | %22lld %22lld %22lld |\n", names[i], v[i], v[i], v[i]);
            }
            std::printf("[BaryTree] %s\n[BaryTree]\n", bar);
        };
        const long long cw[5] = {s.num_pp + s.num_pc + s.num_cp + s.num_cc, s.num_pp, s.num_pc, s.num_cp,
                                 s.num_cc};
        table("Clusterwise interaction information", cw);
        const long long pw[5] = {s.pp_ptwise + s.pc_ptwise + s.cp_ptwise + s.cc_ptwise, s.pp_ptwise,
                                 s.pc_ptwise, s.cp_ptwise, s.cc_ptwise};
        table("Pointwise interaction information", pw);
        const long long pure = s.pp_ptwise + s.pc_replaced + s.cp_replaced + s.cc_replaced;
        const long long acc[4] = {s.pp_ptwise, s.pc_replaced, s.cp_replaced, s.cc_replaced};
        const char *names[4] = {"PP   ", "PC   ", "CP   ", "CC   "};
        const char *bar2 = "|-------------------------------------------------------|";
        std::printf("[BaryTree]\n[BaryTree] Pointwise interactions accounted for:\n[BaryTree]\n");
        std::printf("[BaryTree] %s\n[BaryTree] |       |                 Number             Percentage |\n", bar2);
        std::printf("[BaryTree] %s\n[BaryTree] | Total | %22lld            100.000000%% |\n", bar2, pure);
        for (int i = 0; i < 4; i++)
            std::printf("[BaryTree] %s\n[BaryTree] | %s | %22lld %21.6f%% |\n", bar2, names[i], acc[i],
                        pure > 0 ? 100.0 * (double)acc[i] / (double)pure : 0.0);
        std::printf("[BaryTree] %s\n[BaryTree]\n", bar2);
        // phase times (this build's phases; the reference's are its own)
        std::printf("[BaryTree] Phase times (seconds):\n");
        const struct { const char *name; double t; } ph[] = {
            {"host: argument checks and copies", s.host_copy_s}, {"trees", s.host_tree_s},
            {"host: interaction lists", s.host_lists_s}, {"host: work lists and tables", s.host_work_s},
            {"uploads", s.h2d_s}, {"GPU: upward pass", s.gpu_upward_s},
            {"GPU: near field (and PC far field)", s.gpu_near_s}, {"GPU: far field", s.gpu_far_s},
            {"GPU: downpass", s.gpu_down_s}, {"GPU: correction and caller order", s.gpu_finish_s}};
        for (const auto &x : ph) std::printf("[BaryTree]   %-38s %10.6f\n", x.name, x.t);
        std::printf("[BaryTree]\n");
    }
    std::printf("[BaryTree] BaryTree has finished.\n");
    std::fflush(stdout);
}

// The output array is zeroed first (host memory, or device memory).
int zero_output(double *out, int M, bool device)
{
    if (M <= 0) return 0;
    if (!device) {
        std::memset(out, 0, (size_t)M * sizeof(double));
        return 0;
    }
    hipError_t e = hipMemset(out, 0, (size_t)M * sizeof(double));
    return e == hipSuccess ? 0 : fail(std::string("outputArray: ") + hipGetErrorString(e));
}

int solve(int M, int N, const double *tx, const double *ty, const double *tz, const double *tq,
          const double *sx, const double *sy, const double *sz, const double *sq, const double *sw,
          double *out, int kernel, int nkp, const double *kp, int singularity, int approximation,
          int compute_type, double theta, int degree, int msl, int mtl, double size_check,
          double beta, int verbosity, int precision, bool device = false)
{
    g_last_error.clear();
    if (!out) return fail("outputArray is NULL");
    const auto t_start = std::chrono::steady_clock::now();
    if (int z = zero_output(out, M, device)) return z;
    bt::Plan plan;
    int rc = plan_create(plan, M, N, tx, ty, tz, tq, sx, sy, sz, sq, sw, kernel, nkp, kp, singularity,
                         approximation, compute_type, theta, degree, msl, mtl, size_check, beta,
                         verbosity, precision, device);
    if (rc) return rc;
    const auto t_created = std::chrono::steady_clock::now();
    std::string e = plan.execute(true);
    const auto t_exec = std::chrono::steady_clock::now();
    if (e.empty()) e = plan.download(out, device);
    if (!e.empty()) return fail(e);
    fill_stats(plan, &g_stats);
    g_have_stats = true;
    const auto t_down = std::chrono::steady_clock::now();
    plan.release();
    const auto t_end = std::chrono::steady_clock::now();
    if (verbosity > 3) {
        auto sec = [](auto a, auto b) { return std::chrono::duration<double>(b - a).count(); };
        std::printf("[BaryTree] call: plan %.6f, execution %.6f, download %.6f, release %.6f s\n",
                    sec(t_start, t_created), sec(t_created, t_exec), sec(t_exec, t_down), sec(t_down, t_end));
    }
    print_diagnostics(g_stats, M, N, std::chrono::duration<double>(t_end - t_start).count(), verbosity);
    return 0;
}

// ---- batching ---------------------------------------------------------

// 2^c <= max(M, N) < 2^(c+1)
int size_class(int M, int N)
{
    const long long n = std::max(M, N);
    int c = 0;
    while (c + 1 < BARYTREE_SIZE_CLASSES && (2LL << c) <= n) c++;
    return c;
}

std::string check_offsets(int nsys, const int *toffs, const int *soffs)
{
    if (nsys < 1) return "numSystems must be positive";
    if (!toffs || !soffs) return "system offsets are NULL";
    if (toffs[0] != 0 || soffs[0] != 0) return "system offsets must start at 0";
    for (int s = 0; s < nsys; s++)
        if (toffs[s + 1] <= toffs[s] || soffs[s + 1] <= soffs[s])
            return "every system needs at least one target and one source (offsets increasing)";
    return "";
}

// leaf size of size class c from a schedule: its own entry, else the nearest
// set class, else 2P
int schedule_leaf(const BaryTreeSchedule &sc, int c, int P)
{
    for (int d = 0; d < BARYTREE_SIZE_CLASSES; d++) {
        if (c - d >= 0 && sc.leaf[c - d] > 0) return sc.leaf[c - d];
        if (c + d < BARYTREE_SIZE_CLASSES && sc.leaf[c + d] > 0) return sc.leaf[c + d];
    }
    return 2 * P;
}

using TuneKey = std::tuple<int, int, int, int, int, double, double, double, int>;
std::mutex g_tune_mu;
std::map<TuneKey, BaryTreeSchedule> g_tune_cache;

TuneKey tune_key(const bt::Params &p)
{
    return TuneKey{p.kernel, p.singularity, p.approximation, p.compute_type, p.degree, p.theta,
                   p.size_check_factor, p.beta, p.precision};
}

struct BatchArrays {
    const double *tx, *ty, *tz, *tq, *sx, *sy, *sz, *sq, *sw;
};

// Minimum GPU time of `reps` timed executions after one warm-up; < 0 on error.
double time_plan(bt::Plan &plan, int reps)
{
    if (!plan.execute(false).empty()) return -1.0;
    double best = 1e300;
    for (int r = 0; r < reps; r++) {
        if (!plan.execute(true).empty()) return -1.0;
        best = std::min(best, plan.timings().gpu_total);
    }
    return best;
}

// A tuning sample: the listed systems, each repeated `rep` times, as one
// concatenated batch with its own offsets.
struct Sample {
    std::vector<int> toffs{0}, soffs{0}, leaf;
    std::vector<double> tx, ty, tz, tq, sx, sy, sz, sq, sw;
    void add(const BatchArrays &a, const int *toffs_in, const int *soffs_in, int s, int lf)
    {
        const int t0 = toffs_in[s], t1 = toffs_in[s + 1], s0 = soffs_in[s], s1 = soffs_in[s + 1];
        tx.insert(tx.end(), a.tx + t0, a.tx + t1); ty.insert(ty.end(), a.ty + t0, a.ty + t1);
        tz.insert(tz.end(), a.tz + t0, a.tz + t1);
        if (a.tq) tq.insert(tq.end(), a.tq + t0, a.tq + t1);
        else tq.insert(tq.end(), (size_t)(t1 - t0), 0.0);
        sx.insert(sx.end(), a.sx + s0, a.sx + s1); sy.insert(sy.end(), a.sy + s0, a.sy + s1);
        sz.insert(sz.end(), a.sz + s0, a.sz + s1); sq.insert(sq.end(), a.sq + s0, a.sq + s1);
        if (a.sw) sw.insert(sw.end(), a.sw + s0, a.sw + s1);
        else sw.insert(sw.end(), (size_t)(s1 - s0), 1.0);
        toffs.push_back(toffs.back() + (t1 - t0));
        soffs.push_back(soffs.back() + (s1 - s0));
        leaf.push_back(lf);
    }
    int nsys() const { return (int)leaf.size(); }
    bt::PlanInput input(int tpt, int persistent) const
    {
        bt::PlanInput in;
        in.nsys = nsys(); in.toffs = toffs.data(); in.soffs = soffs.data();
        in.msl = leaf.data(); in.mtl = leaf.data(); in.tpt = tpt; in.persistent = persistent;
        in.tx = tx.data(); in.ty = ty.data(); in.tz = tz.data(); in.tq = tq.data();
        in.sx = sx.data(); in.sy = sy.data(); in.sz = sz.data(); in.sq = sq.data(); in.sw = sw.data();
        return in;
    }
};

constexpr long long TUNE_FILL = 1 << 20;  // targets per tuning plan
constexpr int TUNE_SAMPLE = 8;            // sampled systems per size class
constexpr int TUNE_REPS = 3;

int autotune(const bt::Params &prm, int nsys, const int *toffs, const int *soffs,
             const BatchArrays &a, BaryTreeSchedule *out)
{
    const auto t_start = std::chrono::steady_clock::now();
    BaryTreeSchedule sc;
    std::memset(&sc, 0, sizeof(sc));
    sc.targets_per_thread = 2;
    sc.persistent = 1;
    std::map<int, std::vector<int>> classes;
    for (int s = 0; s < nsys; s++)
        classes[size_class(toffs[s + 1] - toffs[s], soffs[s + 1] - soffs[s])].push_back(s);
    const int P = prm.P;
    const int cands[5] = {std::max(1, P / 2), P, 2 * P, 4 * P, 8 * P};
    std::map<int, std::vector<int>> picked;  // sampled systems per class
    for (auto &kv : classes) {
        const std::vector<int> &members = kv.second;
        std::vector<int> smp;
        const int k = std::min<int>(TUNE_SAMPLE, (int)members.size());
        for (int i = 0; i < k; i++) smp.push_back(members[(size_t)i * members.size() / k]);
        picked[kv.first] = smp;
        long long nt = 0;
        for (int s : smp) nt += toffs[s + 1] - toffs[s];
        const int rep = (int)std::max<long long>(1, (TUNE_FILL + nt - 1) / nt);
        double best_t = 1e300;
        int best_leaf = 2 * P;
        for (int lf : cands) {
            Sample S;
            for (int r = 0; r < rep; r++)
                for (int s : smp) S.add(a, toffs, soffs, s, lf);
            bt::Plan plan;
            std::string e = plan.create(prm, S.input(2, 1));
            if (!e.empty()) return fail("autotune: " + e);
            const double t = time_plan(plan, TUNE_REPS);
            if (t < 0) return fail("autotune: execution failed");
            sc.trials++;
            // per target of the sample (the same for every candidate)
            if (t < best_t) { best_t = t; best_leaf = lf; }
        }
        sc.leaf[kv.first] = best_leaf;
    }
    // tile shape and launch form on all class samples together
    {
        Sample S;
        long long nt = 0;
        for (auto &kv : picked)
            for (int s : kv.second) nt += toffs[s + 1] - toffs[s];
        const int rep = (int)std::max<long long>(1, (TUNE_FILL + nt - 1) / nt);
        for (int r = 0; r < rep; r++)
            for (auto &kv : picked)
                for (int s : kv.second) S.add(a, toffs, soffs, s, sc.leaf[kv.first]);
        double best_t = 1e300;
        for (int tpt = 1; tpt <= 2; tpt++) {
            bt::Plan plan;
            std::string e = plan.create(prm, S.input(tpt, 0));
            if (!e.empty()) return fail("autotune: " + e);
            for (int pers = 0; pers <= 1; pers++) {
                plan.set_persistent(pers != 0);
                const double t = time_plan(plan, TUNE_REPS);
                if (t < 0) return fail("autotune: execution failed");
                sc.trials++;
                if (t < best_t) { best_t = t; sc.targets_per_thread = tpt; sc.persistent = pers; }
            }
        }
    }
    sc.tune_seconds =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    *out = sc;
    return 0;
}

int batched_params(bt::Params &prm, int kernel, int nkp, const double *kp, int singularity,
                   int approximation, int compute_type, double theta, int degree,
                   double size_check, double beta, int verbosity, int precision)
{
    // leaf sizes are per system; 1 stands in until they are resolved
    bt::params_setup(prm, kernel, nkp, kp, approximation, singularity, compute_type, theta, degree,
                     1, 1, size_check, beta, verbosity);
    if (precision != BARYTREE_PRECISION_FP64 && precision != BARYTREE_PRECISION_MIXED)
        return fail("unknown precision mode");
    prm.precision = precision;
    std::string e = bt::params_check(prm);
    if (!e.empty()) return fail(e);
    return 0;
}

int plan_create_batched(bt::Plan &plan, int nsys, const int *toffs, const int *soffs,
                        const BatchArrays &a, bt::Params prm, const int *msl, const int *mtl,
                        const BaryTreeSchedule *schedule, bool host_only = false)
{
    std::string e = check_offsets(nsys, toffs, soffs);
    if (!e.empty()) return fail(e);
    if (!a.tx || !a.ty || !a.tz || !a.sx || !a.sy || !a.sz || !a.sq)
        return fail("One or more required particle arrays is unset");
    std::vector<int> vs(nsys), vt(nsys);
    BaryTreeSchedule sc;
    std::memset(&sc, 0, sizeof(sc));
    sc.targets_per_thread = 2;
    const bool beta_leaves = prm.beta >= 0;  // run_params.c:51-120 sets the leaf sizes
    if (schedule) {
        sc = *schedule;
        if (sc.targets_per_thread != 1 && sc.targets_per_thread != 2)
            return fail("schedule: targets_per_thread must be 1 or 2");
    } else if (!beta_leaves && (!msl || !mtl) && host_only) {
        return fail("leaf sizes or a schedule are needed (the autotuner runs on the GPU)");
    } else if (!beta_leaves && (!msl || !mtl)) {
        const TuneKey key = tune_key(prm);
        bool have = false;
        {
            std::lock_guard<std::mutex> lk(g_tune_mu);
            auto it = g_tune_cache.find(key);
            if (it != g_tune_cache.end()) {
                sc = it->second;
                have = true;
                for (int s = 0; s < nsys && have; s++)
                    if (sc.leaf[size_class(toffs[s + 1] - toffs[s], soffs[s + 1] - soffs[s])] <= 0)
                        have = false;
            }
        }
        if (!have) {
            BaryTreeSchedule tuned;
            int rc = autotune(prm, nsys, toffs, soffs, a, &tuned);
            if (rc) return rc;
            std::lock_guard<std::mutex> lk(g_tune_mu);
            BaryTreeSchedule &c = g_tune_cache[key];
            const bool fresh = c.targets_per_thread == 0;
            for (int k = 0; k < BARYTREE_SIZE_CLASSES; k++)
                if (tuned.leaf[k] > 0) c.leaf[k] = tuned.leaf[k];
            if (fresh) {
                c.targets_per_thread = tuned.targets_per_thread;
                c.persistent = tuned.persistent;
            }
            c.trials += tuned.trials;
            c.tune_seconds += tuned.tune_seconds;
            sc = c;
        }
    }
    for (int s = 0; s < nsys; s++) {
        if (beta_leaves) {
            vs[s] = prm.max_source_leaf;
            vt[s] = prm.max_target_leaf;
            continue;
        }
        const int c = size_class(toffs[s + 1] - toffs[s], soffs[s + 1] - soffs[s]);
        vs[s] = msl ? msl[s] : schedule_leaf(sc, c, prm.P);
        vt[s] = mtl ? mtl[s] : schedule_leaf(sc, c, prm.P);
    }
    prm.max_source_leaf = vs[0];
    prm.max_target_leaf = vt[0];
    bt::PlanInput in;
    in.nsys = nsys; in.toffs = toffs; in.soffs = soffs; in.msl = vs.data(); in.mtl = vt.data();
    in.tpt = sc.targets_per_thread; in.persistent = sc.persistent;
    in.tx = a.tx; in.ty = a.ty; in.tz = a.tz; in.tq = a.tq;
    in.sx = a.sx; in.sy = a.sy; in.sz = a.sz; in.sq = a.sq; in.sw = a.sw;
    e = plan.create(prm, in, host_only);
    if (!e.empty()) return fail(e);
    return 0;
}

}  // namespace

extern "C" {

void BaryTreeInterface(int numTargets, int numSources, double *targetX, double *targetY,
                       double *targetZ, double *targetValue, double *sourceX, double *sourceY,
                       double *sourceZ, double *sourceValue, double *sourceWeight,
                       double *outputArray, KERNEL kernel, int numKernelParams,
                       double *kernelParams, SINGULARITY singularity, APPROXIMATION approximation,
                       COMPUTE_TYPE compute_type, double theta, int interpOrder,
                       int maxPerSourceLeaf, int maxPerTargetLeaf, double sizeCheck, double beta,
                       int verbosity)
{
    int rc = solve(numTargets, numSources, targetX, targetY, targetZ, targetValue, sourceX, sourceY,
                   sourceZ, sourceValue, sourceWeight, outputArray, (int)kernel, numKernelParams,
                   kernelParams, (int)singularity, (int)approximation, (int)compute_type, theta,
                   interpOrder, maxPerSourceLeaf, maxPerTargetLeaf, sizeCheck, beta, verbosity,
                   BARYTREE_PRECISION_FP64);
    if (rc) {
        std::printf("[BaryTree]\n[BaryTree] ERROR! %s\n[BaryTree]\n", g_last_error.c_str());
        if (outputArray && numTargets > 0) std::memset(outputArray, 0, (size_t)numTargets * sizeof(double));
    }
}

int BaryTree_Solve(int numTargets, int numSources, const double *targetX, const double *targetY,
                   const double *targetZ, const double *targetValue, const double *sourceX,
                   const double *sourceY, const double *sourceZ, const double *sourceValue,
                   const double *sourceWeight, double *outputArray, int kernel,
                   int numKernelParams, const double *kernelParams, int singularity,
                   int approximation, int compute_type, double theta, int interpOrder,
                   int maxPerSourceLeaf, int maxPerTargetLeaf, double sizeCheck, double beta,
                   int verbosity, int precision)
{
    return solve(numTargets, numSources, targetX, targetY, targetZ, targetValue, sourceX, sourceY,
                 sourceZ, sourceValue, sourceWeight, outputArray, kernel, numKernelParams,
                 kernelParams, singularity, approximation, compute_type, theta, interpOrder,
                 maxPerSourceLeaf, maxPerTargetLeaf, sizeCheck, beta, verbosity, precision);
}

const char *BaryTree_GetLastError(void) { return g_last_error.c_str(); }

unsigned long long BaryTree_EmptyDeviceCache(void) { return btm::empty_cache(); }

int BaryTree_SetDeviceCacheLimit(unsigned long long bytes)
{
    btm::set_cache_limit((size_t)bytes);
    return 0;
}

unsigned long long BaryTree_CachedDeviceBytes(void) { return btm::cached_bytes(); }

int BaryTree_SetBetaLeafSizes(int which)
{
    g_last_error.clear();
    if (which != BARYTREE_BETA_LEAVES_CPU && which != BARYTREE_BETA_LEAVES_GPU)
        return fail("which must be BARYTREE_BETA_LEAVES_CPU or BARYTREE_BETA_LEAVES_GPU");
    bt::set_beta_leaf_mode(which);
    return 0;
}

int BaryTree_GetLastStats(BaryTreeStats *stats)
{
    if (!stats || !g_have_stats) return -1;
    *stats = g_stats;
    return 0;
}

int BaryTreePlan_Create(BaryTreePlan **plan, int numTargets, int numSources, const double *targetX,
                        const double *targetY, const double *targetZ, const double *targetValue,
                        const double *sourceX, const double *sourceY, const double *sourceZ,
                        const double *sourceValue, const double *sourceWeight, int kernel,
                        int numKernelParams, const double *kernelParams, int singularity,
                        int approximation, int compute_type, double theta, int interpOrder,
                        int maxPerSourceLeaf, int maxPerTargetLeaf, double sizeCheck, double beta,
                        int precision)
{
    g_last_error.clear();
    if (!plan) return fail("plan is NULL");
    *plan = nullptr;
    BaryTreePlan *p = new (std::nothrow) BaryTreePlan;
    if (!p) return fail("out of host memory");
    int rc = plan_create(p->plan, numTargets, numSources, targetX, targetY, targetZ, targetValue,
                         sourceX, sourceY, sourceZ, sourceValue, sourceWeight, kernel,
                         numKernelParams, kernelParams, singularity, approximation, compute_type,
                         theta, interpOrder, maxPerSourceLeaf, maxPerTargetLeaf, sizeCheck, beta, 0,
                         precision);
    if (rc) {
        delete p;
        return rc;
    }
    *plan = p;
    return 0;
}

int BaryTreePlan_Execute(BaryTreePlan *plan, double *outputArray)
{
    g_last_error.clear();
    if (!plan) return fail("plan is NULL");
    std::string e = plan->plan.execute(true);
    if (e.empty() && outputArray) e = plan->plan.download(outputArray);
    if (!e.empty()) return fail(e);
    return 0;
}

int BaryTreePlan_SetPrecision(BaryTreePlan *plan, int precision)
{
    g_last_error.clear();
    if (!plan) return fail("plan is NULL");
    if (precision != BARYTREE_PRECISION_FP64 && precision != BARYTREE_PRECISION_MIXED)
        return fail("unknown precision mode");
    plan->plan.set_precision(precision);
    return 0;
}

int BaryTreePlan_SetSymmetric(BaryTreePlan *plan, int enable)
{
    g_last_error.clear();
    if (!plan) return fail("plan is NULL");
    if (enable != 0 && enable != 1) return fail("enable must be 0 or 1");
    plan->plan.set_symmetric(enable != 0);
    return 0;
}

int BaryTreePlan_GetStats(const BaryTreePlan *plan, BaryTreeStats *stats)
{
    if (!plan || !stats) return fail("NULL argument");
    fill_stats(plan->plan, stats);
    return 0;
}

int BaryTreePlan_DirectSample(BaryTreePlan *plan, int m, const int *targetIndex, double *out)
{
    g_last_error.clear();
    if (!plan || !targetIndex || !out || m < 0) return fail("bad argument");
    std::vector<int> idx(targetIndex, targetIndex + m);
    std::vector<double> res;
    std::string e = plan->plan.direct(idx, res);
    if (!e.empty()) return fail(e);
    std::memcpy(out, res.data(), (size_t)m * sizeof(double));
    return 0;
}

int BaryTreePlan_DirectSampleNoAlloc(BaryTreePlan *plan, int m, const int *targetIndex, double *out)
{
    g_last_error.clear();
    if (!plan || !targetIndex || !out || m < 0) return fail("bad argument");
    if (m > btk::NOALLOC_MAX_SAMPLE)
        return fail("BaryTreePlan_DirectSampleNoAlloc takes at most 8192 targets (the size of its scratch)");
    void *scratch = nullptr;
    size_t bytes = 0;
    hipError_t he = btk::noalloc_scratch(&scratch, &bytes);
    if (he != hipSuccess) return fail(std::string("scratch lookup failed: ") + hipGetErrorString(he));
    std::string e = plan->plan.direct_noalloc(m, targetIndex, out, scratch, bytes);
    if (!e.empty()) return fail(e);
    return 0;
}

void BaryTreePlan_Destroy(BaryTreePlan *plan) { delete plan; }

int BaryTree_Analyze(int numTargets, int numSources, const double *targetX, const double *targetY,
                     const double *targetZ, const double *sourceX, const double *sourceY,
                     const double *sourceZ, int kernel, int numKernelParams,
                     const double *kernelParams, int singularity, int approximation,
                     int compute_type, double theta, int interpOrder, int maxPerSourceLeaf,
                     int maxPerTargetLeaf, double sizeCheck, double beta, BaryTreeStats *stats)
{
    g_last_error.clear();
    if (!stats) return fail("stats is NULL");
    bt::Params prm;
    bt::params_setup(prm, kernel, numKernelParams, kernelParams, approximation, singularity,
                     compute_type, theta, interpOrder, maxPerSourceLeaf, maxPerTargetLeaf, sizeCheck,
                     beta, 0);
    bt::Plan plan;
    // charges do not influence trees or lists
    std::vector<double> zs(numSources > 0 ? numSources : 0, 0.0);
    std::string e = plan.create(prm, numTargets, numSources, targetX, targetY, targetZ, nullptr,
                                sourceX, sourceY, sourceZ, zs.data(), nullptr, true);
    if (!e.empty()) return fail(e);
    fill_stats(plan, stats);
    return 0;
}


int BaryTree_SolveBatched(int numSystems, const int *targetOffsets, const int *sourceOffsets,
                          const double *targetX, const double *targetY, const double *targetZ,
                          const double *targetValue, const double *sourceX,
                          const double *sourceY, const double *sourceZ,
                          const double *sourceValue, const double *sourceWeight,
                          double *outputArray, int kernel, int numKernelParams,
                          const double *kernelParams, int singularity, int approximation,
                          int compute_type, double theta, int interpOrder,
                          const int *maxPerSourceLeaf, const int *maxPerTargetLeaf,
                          const BaryTreeSchedule *schedule, double sizeCheck, double beta,
                          int verbosity, int precision)
{
    g_last_error.clear();
    if (!outputArray) return fail("outputArray is NULL");
    const auto t_start = std::chrono::steady_clock::now();
    std::string e = check_offsets(numSystems, targetOffsets, sourceOffsets);
    if (!e.empty()) return fail(e);
    std::memset(outputArray, 0, (size_t)targetOffsets[numSystems] * sizeof(double));
    bt::Params prm;
    int rc = batched_params(prm, kernel, numKernelParams, kernelParams, singularity, approximation,
                            compute_type, theta, interpOrder, sizeCheck, beta, verbosity, precision);
    if (rc) return rc;
    const BatchArrays a{targetX, targetY, targetZ, targetValue, sourceX, sourceY, sourceZ,
                        sourceValue, sourceWeight};
    bt::Plan plan;
    rc = plan_create_batched(plan, numSystems, targetOffsets, sourceOffsets, a, prm,
                             maxPerSourceLeaf, maxPerTargetLeaf, schedule);
    if (rc) return rc;
    e = plan.execute(true);
    if (e.empty()) e = plan.download(outputArray);
    if (!e.empty()) return fail(e);
    fill_stats(plan, &g_stats);
    g_have_stats = true;
    if (verbosity > 1) std::printf("[BaryTree] batched: %d independent systems.\n", numSystems);
    print_diagnostics(g_stats, targetOffsets[numSystems], sourceOffsets[numSystems],
                      std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count(), verbosity);
    return 0;
}

int BaryTreePlan_CreateBatched(BaryTreePlan **plan, int numSystems, const int *targetOffsets,
                               const int *sourceOffsets, const double *targetX,
                               const double *targetY, const double *targetZ,
                               const double *targetValue, const double *sourceX,
                               const double *sourceY, const double *sourceZ,
                               const double *sourceValue, const double *sourceWeight,
                               int kernel, int numKernelParams, const double *kernelParams,
                               int singularity, int approximation, int compute_type,
                               double theta, int interpOrder, const int *maxPerSourceLeaf,
                               const int *maxPerTargetLeaf, const BaryTreeSchedule *schedule,
                               double sizeCheck, double beta, int precision)
{
    g_last_error.clear();
    if (!plan) return fail("plan is NULL");
    *plan = nullptr;
    bt::Params prm;
    int rc = batched_params(prm, kernel, numKernelParams, kernelParams, singularity, approximation,
                            compute_type, theta, interpOrder, sizeCheck, beta, 0, precision);
    if (rc) return rc;
    BaryTreePlan *p = new (std::nothrow) BaryTreePlan;
    if (!p) return fail("out of host memory");
    const BatchArrays a{targetX, targetY, targetZ, targetValue, sourceX, sourceY, sourceZ,
                        sourceValue, sourceWeight};
    rc = plan_create_batched(p->plan, numSystems, targetOffsets, sourceOffsets, a, prm,
                             maxPerSourceLeaf, maxPerTargetLeaf, schedule);
    if (rc) {
        delete p;
        return rc;
    }
    *plan = p;
    return 0;
}

int BaryTree_AnalyzeBatched(int numSystems, const int *targetOffsets, const int *sourceOffsets,
                            const double *targetX, const double *targetY, const double *targetZ,
                            const double *sourceX, const double *sourceY, const double *sourceZ,
                            int kernel, int numKernelParams, const double *kernelParams,
                            int singularity, int approximation, int compute_type, double theta,
                            int interpOrder, const int *maxPerSourceLeaf,
                            const int *maxPerTargetLeaf, const BaryTreeSchedule *schedule,
                            double sizeCheck, double beta, BaryTreeStats *stats)
{
    g_last_error.clear();
    if (!stats) return fail("stats is NULL");
    std::string e = check_offsets(numSystems, targetOffsets, sourceOffsets);
    if (!e.empty()) return fail(e);
    bt::Params prm;
    int rc = batched_params(prm, kernel, numKernelParams, kernelParams, singularity, approximation,
                            compute_type, theta, interpOrder, sizeCheck, beta, 0,
                            BARYTREE_PRECISION_FP64);
    if (rc) return rc;
    // charges do not influence trees or lists
    std::vector<double> zs((size_t)sourceOffsets[numSystems], 0.0);
    const BatchArrays a{targetX, targetY, targetZ, nullptr, sourceX, sourceY, sourceZ, zs.data(),
                        nullptr};
    bt::Plan plan;
    rc = plan_create_batched(plan, numSystems, targetOffsets, sourceOffsets, a, prm,
                             maxPerSourceLeaf, maxPerTargetLeaf, schedule, true);
    if (rc) return rc;
    fill_stats(plan, stats);
    return 0;
}

int BaryTree_Autotune(int numSystems, const int *targetOffsets, const int *sourceOffsets,
                      const double *targetX, const double *targetY, const double *targetZ,
                      const double *targetValue, const double *sourceX, const double *sourceY,
                      const double *sourceZ, const double *sourceValue,
                      const double *sourceWeight, int kernel, int numKernelParams,
                      const double *kernelParams, int singularity, int approximation,
                      int compute_type, double theta, int interpOrder, double sizeCheck,
                      double beta, int precision, BaryTreeSchedule *schedule)
{
    g_last_error.clear();
    if (!schedule) return fail("schedule is NULL");
    std::string e = check_offsets(numSystems, targetOffsets, sourceOffsets);
    if (!e.empty()) return fail(e);
    if (!targetX || !targetY || !targetZ || !sourceX || !sourceY || !sourceZ || !sourceValue)
        return fail("One or more required particle arrays is unset");
    bt::Params prm;
    int rc = batched_params(prm, kernel, numKernelParams, kernelParams, singularity, approximation,
                            compute_type, theta, interpOrder, sizeCheck, beta, 0, precision);
    if (rc) return rc;
    if (prm.singularity == bt::S_SUBTRACTION && !sourceWeight)
        return fail("Singularity subtraction requires weights for the sources");
    const BatchArrays a{targetX, targetY, targetZ, targetValue, sourceX, sourceY, sourceZ,
                        sourceValue, sourceWeight};
    return autotune(prm, numSystems, targetOffsets, sourceOffsets, a, schedule);
}


int BaryTreePlan_NumPairs(const BaryTreePlan *plan)
{
    if (!plan) return fail("plan is NULL");
    return (int)plan->plan.far_pairs_list().size();
}

int BaryTreePlan_PairInfo(const BaryTreePlan *plan, int *kind, int *system)
{
    if (!plan) return fail("plan is NULL");
    const auto &pr = plan->plan.far_pairs_list();
    for (size_t i = 0; i < pr.size(); i++) {
        if (kind) kind[i] = pr[i].kind;
        if (system) system[i] = pr[i].sys;
    }
    return 0;
}

int BaryTreePlan_PairFeatures(const BaryTreePlan *plan, float *features)
{
    if (!plan || !features) return fail("NULL argument");
    std::vector<float> f;
    plan->plan.pair_features(f);
    std::memcpy(features, f.data(), f.size() * sizeof(float));
    return 0;
}

int BaryTreePlan_LabelPairs(BaryTreePlan *plan, int first, int count, double tol, int *labels,
                            double *errors)
{
    g_last_error.clear();
    if (!plan || !labels) return fail("NULL argument");
    std::string e = plan->plan.label_pairs(first, count, tol, labels, errors);
    if (!e.empty()) return fail(e);
    return 0;
}

int BaryTreePlan_SetPairDegrees(BaryTreePlan *plan, const int *degrees)
{
    g_last_error.clear();
    if (!plan) return fail("plan is NULL");
    std::string e = plan->plan.set_pair_degrees(degrees);
    if (!e.empty()) return fail(e);
    return 0;
}

int BaryTreePlan_UseDegreeModel(BaryTreePlan *plan, const BaryTreeDegreeModel *model, float delta)
{
    g_last_error.clear();
    if (!plan || !model) return fail("NULL argument");
    std::string e = use_model(plan->plan, model->m, delta);
    if (!e.empty()) return fail(e);
    return 0;
}

// Labels of explicitly given pairs; ss: the subtraction kernel (target
// charges tq, source weights sw, p1 the Gaussian's 1/width^2).
static int label_explicit(int numPairs, const int *kind, const int *targetBegin, const int *targetCount,
                          const int *sourceBegin, const int *sourceCount, const double *targetBox,
                          const double *sourceBox, int numTargets, const double *targetX, const double *targetY,
                          const double *targetZ, const double *tq, int numSources, const double *sourceX,
                          const double *sourceY, const double *sourceZ, const double *sourceValue,
                          const double *sw, int kernel, double kernelParam, int ss, double p1, double tol,
                          int *labels, double *errors)
{
    g_last_error.clear();
    if (numPairs < 0 || !labels) return fail("bad argument");
    if (numPairs == 0) return 0;
    if (!kind || !targetBegin || !targetCount || !sourceBegin || !sourceCount || !targetBox ||
        !sourceBox || !targetX || !targetY || !targetZ || !sourceX || !sourceY || !sourceZ ||
        !sourceValue || (ss && (!tq || !sw)))
        return fail("NULL argument");
    if (kernel < 1 || kernel > 3) return fail("INVALID KERNEL (supported: COULOMB, YUKAWA, REGULARIZED_COULOMB)");
    for (int i = 0; i < numPairs; i++) {
        if (kind[i] < 1 || kind[i] > 3) return fail("pair kind must be 1, 2 or 3");
        if (targetCount[i] <= 0 || sourceCount[i] <= 0 || targetBegin[i] < 0 || sourceBegin[i] < 0 ||
            (long long)targetBegin[i] + targetCount[i] > numTargets ||
            (long long)sourceBegin[i] + sourceCount[i] > numSources)
            return fail("pair particle range out of bounds");
    }
    std::vector<void *> tmp;
    auto cleanup = [&]() { for (void *q : tmp) (void)hipFree(q); };
    auto put = [&](const void *h, size_t bytes) -> void * {
        void *d = nullptr;
        if (hipMalloc(&d, std::max<size_t>(bytes, 8)) != hipSuccess) return nullptr;
        tmp.push_back(d);
        if (hipMemcpy(d, h, bytes, hipMemcpyHostToDevice) != hipSuccess) return nullptr;
        return d;
    };
    btl::LabelArgs a{};
    a.npairs = numPairs;
    const size_t np = numPairs;
    a.kind = (const int *)put(kind, np * sizeof(int));
    a.tbeg = (const int *)put(targetBegin, np * sizeof(int));
    a.tcnt = (const int *)put(targetCount, np * sizeof(int));
    a.sbeg = (const int *)put(sourceBegin, np * sizeof(int));
    a.scnt = (const int *)put(sourceCount, np * sizeof(int));
    a.tbox = (const double *)put(targetBox, 6 * np * sizeof(double));
    a.sbox = (const double *)put(sourceBox, 6 * np * sizeof(double));
    a.tx = (const double *)put(targetX, (size_t)numTargets * sizeof(double));
    a.ty = (const double *)put(targetY, (size_t)numTargets * sizeof(double));
    a.tz = (const double *)put(targetZ, (size_t)numTargets * sizeof(double));
    a.sx = (const double *)put(sourceX, (size_t)numSources * sizeof(double));
    a.sy = (const double *)put(sourceY, (size_t)numSources * sizeof(double));
    a.sz = (const double *)put(sourceZ, (size_t)numSources * sizeof(double));
    a.sq = (const double *)put(sourceValue, (size_t)numSources * sizeof(double));
    if (ss) {
        a.ss = 1;
        a.tq = (const double *)put(tq, (size_t)numTargets * sizeof(double));
        a.sw = (const double *)put(sw, (size_t)numSources * sizeof(double));
        a.p1 = p1;
    }
    void *dl = nullptr, *de = nullptr;
    if (hipMalloc(&dl, np * sizeof(int)) == hipSuccess) tmp.push_back(dl); else dl = nullptr;
    if (hipMalloc(&de, np * btl::LABEL_ND * sizeof(double)) == hipSuccess) tmp.push_back(de); else de = nullptr;
    if (!a.kind || !a.tbeg || !a.tcnt || !a.sbeg || !a.scnt || !a.tbox || !a.sbox || !a.tx || !a.ty ||
        !a.tz || !a.sx || !a.sy || !a.sz || !a.sq || !dl || !de || (ss && (!a.tq || !a.sw))) {
        cleanup();
        return fail("device buffers for labels failed");
    }
    a.kernel = kernel;
    a.p0 = kernelParam;
    a.tol = tol;
    a.labels = (int *)dl;
    a.errors = (double *)de;
    hipError_t he = btl::launch_label(a, nullptr);
    if (he == hipSuccess) he = hipMemcpy(labels, dl, np * sizeof(int), hipMemcpyDeviceToHost);
    if (he == hipSuccess && errors)
        he = hipMemcpy(errors, de, np * btl::LABEL_ND * sizeof(double), hipMemcpyDeviceToHost);
    cleanup();
    if (he != hipSuccess) return fail(std::string("label kernel: ") + hipGetErrorString(he));
    return 0;
}

int BaryTree_LabelPairs(int numPairs, const int *kind, const int *targetBegin,
                        const int *targetCount, const int *sourceBegin, const int *sourceCount,
                        const double *targetBox, const double *sourceBox, int numTargets,
                        const double *targetX, const double *targetY, const double *targetZ,
                        int numSources, const double *sourceX, const double *sourceY,
                        const double *sourceZ, const double *sourceValue, int kernel,
                        double kernelParam, double tol, int *labels, double *errors)
{
    return label_explicit(numPairs, kind, targetBegin, targetCount, sourceBegin, sourceCount, targetBox,
                          sourceBox, numTargets, targetX, targetY, targetZ, nullptr, numSources, sourceX,
                          sourceY, sourceZ, sourceValue, nullptr, kernel, kernelParam, 0, 0.0, tol, labels,
                          errors);
}

int BaryTree_LabelPairsSubtraction(int numPairs, const int *kind, const int *targetBegin,
                                   const int *targetCount, const int *sourceBegin, const int *sourceCount,
                                   const double *targetBox, const double *sourceBox, int numTargets,
                                   const double *targetX, const double *targetY, const double *targetZ,
                                   const double *targetValue, int numSources, const double *sourceX,
                                   const double *sourceY, const double *sourceZ, const double *sourceValue,
                                   const double *sourceWeight, int kernel, int numKernelParams,
                                   const double *kernelParams, double tol, int *labels, double *errors)
{
    g_last_error.clear();
    if (!kernelParams || numKernelParams < (kernel == 3 ? 2 : 1)) return fail("kernel parameters missing");
    double p0 = 0.0, p1 = 0.0;
    if (kernel == 1) {
        p1 = 1.0 / (kernelParams[0] * kernelParams[0]);
    } else if (kernel == 2) {
        p0 = kernelParams[0];
    } else if (kernel == 3) {
        p0 = kernelParams[1];
        p1 = 1.0 / (kernelParams[0] * kernelParams[0]);
    } else {
        return fail("kernel must be 1, 2 or 3");
    }
    return label_explicit(numPairs, kind, targetBegin, targetCount, sourceBegin, sourceCount, targetBox,
                          sourceBox, numTargets, targetX, targetY, targetZ, targetValue, numSources, sourceX,
                          sourceY, sourceZ, sourceValue, sourceWeight, kernel, p0, 1, p1, tol, labels, errors);
}

int BaryTreeModel_Create(BaryTreeDegreeModel **model, unsigned long long seed)
{
    g_last_error.clear();
    if (!model) return fail("model is NULL");
    *model = new (std::nothrow) BaryTreeDegreeModel;
    if (!*model) return fail("out of host memory");
    (*model)->m.init(seed);
    return 0;
}

int BaryTreeModel_Load(BaryTreeDegreeModel **model, const char *path)
{
    g_last_error.clear();
    if (!model || !path) return fail("NULL argument");
    *model = nullptr;
    BaryTreeDegreeModel *m = new (std::nothrow) BaryTreeDegreeModel;
    if (!m) return fail("out of host memory");
    std::string e = m->m.load(path);
    if (!e.empty()) {
        delete m;
        return fail(e);
    }
    *model = m;
    return 0;
}

int BaryTreeModel_Save(const BaryTreeDegreeModel *model, const char *path)
{
    g_last_error.clear();
    if (!model || !path) return fail("NULL argument");
    std::string e = model->m.save(path);
    if (!e.empty()) return fail(e);
    return 0;
}

int BaryTreeModel_GetParams(const BaryTreeDegreeModel *model, float *params, float *featureMean,
                            float *featureScale)
{
    if (!model) return fail("model is NULL");
    if (params) std::memcpy(params, model->m.params.data(), btl::MLP_PARAMS * sizeof(float));
    if (featureMean) std::memcpy(featureMean, model->m.mean, sizeof(model->m.mean));
    if (featureScale) std::memcpy(featureScale, model->m.scale, sizeof(model->m.scale));
    return 0;
}

int BaryTreeModel_SetParams(BaryTreeDegreeModel *model, const float *params,
                            const float *featureMean, const float *featureScale)
{
    if (!model) return fail("model is NULL");
    if (params) std::memcpy(model->m.params.data(), params, btl::MLP_PARAMS * sizeof(float));
    if (featureMean) std::memcpy(model->m.mean, featureMean, sizeof(model->m.mean));
    if (featureScale) std::memcpy(model->m.scale, featureScale, sizeof(model->m.scale));
    return 0;
}

int BaryTreeModel_GetFeatureRange(const BaryTreeDegreeModel *model, float *lo, float *hi)
{
    if (!model) return fail("model is NULL");
    if (lo) std::memcpy(lo, model->m.lo, sizeof(model->m.lo));
    if (hi) std::memcpy(hi, model->m.hi, sizeof(model->m.hi));
    return model->m.has_range() ? 1 : 0;
}

int BaryTreeModel_SetFeatureRange(BaryTreeDegreeModel *model, const float *lo, const float *hi)
{
    g_last_error.clear();
    if (!model) return fail("model is NULL");
    if (!lo || !hi) {
        model->m.clear_range();
        return 0;
    }
    for (int f = 0; f < btl::MLP_IN; f++)
        if (!(lo[f] <= hi[f])) return fail("each feature range needs lo <= hi");
    std::memcpy(model->m.lo, lo, sizeof(model->m.lo));
    std::memcpy(model->m.hi, hi, sizeof(model->m.hi));
    return 0;
}

int BaryTreeModel_Gradient(const BaryTreeDegreeModel *model, int n, const float *features,
                           const int *labels, float *grad, double *loss)
{
    g_last_error.clear();
    if (!model || !features || !labels || !grad) return fail("NULL argument");
    std::vector<float> g;
    double l = 0.0;
    std::string e = model->m.gradient(n, features, labels, g, l);
    if (!e.empty()) return fail(e);
    std::memcpy(grad, g.data(), g.size() * sizeof(float));
    if (loss) *loss = l;
    return 0;
}

int BaryTreeModel_Train(BaryTreeDegreeModel *model, int n, const float *features,
                        const int *labels, int epochs, int batch, float learningRate,
                        unsigned long long seed, float *epochLoss)
{
    g_last_error.clear();
    if (!model || !features || !labels) return fail("NULL argument");
    std::vector<float> el;
    std::string e = model->m.train(n, features, labels, epochs, batch, learningRate, seed, &el);
    if (!e.empty()) return fail(e);
    if (epochLoss) std::memcpy(epochLoss, el.data(), el.size() * sizeof(float));
    return 0;
}

int BaryTreeModel_Predict(const BaryTreeDegreeModel *model, int n, const float *features,
                          float delta, int *degrees, float *logits)
{
    g_last_error.clear();
    if (!model || (n > 0 && !features)) return fail("NULL argument");
    std::string e = model->m.predict(n, features, delta, degrees, logits);
    if (!e.empty()) return fail(e);
    return 0;
}

void BaryTreeModel_Destroy(BaryTreeDegreeModel *model) { delete model; }

static int solve_learned(int numTargets, int numSources, const double *targetX,
                          const double *targetY, const double *targetZ, const double *targetValue,
                          const double *sourceX, const double *sourceY, const double *sourceZ,
                          const double *sourceValue, const double *sourceWeight,
                          double *outputArray, int kernel, int numKernelParams,
                          const double *kernelParams, int singularity, int approximation,
                          int compute_type, double theta, int interpOrder, int maxPerSourceLeaf,
                          int maxPerTargetLeaf, double sizeCheck, double beta, int verbosity,
                          int precision, const BaryTreeDegreeModel *model, float delta, bool device)
{
    g_last_error.clear();
    if (!outputArray) return fail("outputArray is NULL");
    const auto t_start = std::chrono::steady_clock::now();
    if (int z = zero_output(outputArray, numTargets, device)) return z;
    bt::Plan plan;
    int rc = plan_create(plan, numTargets, numSources, targetX, targetY, targetZ, targetValue,
                         sourceX, sourceY, sourceZ, sourceValue, sourceWeight, kernel,
                         numKernelParams, kernelParams, singularity, approximation, compute_type,
                         theta, interpOrder, maxPerSourceLeaf, maxPerTargetLeaf, sizeCheck, beta,
                         verbosity, precision, device);
    if (rc) return rc;
    std::string e;
    const auto t_created = std::chrono::steady_clock::now();
    if (model) {
        // the degrees (the model's inference runs on the GPU), then the near
        // field's item launch -- most of the GPU time -- while the host
        // builds the per-degree work
        std::vector<int> deg;
        long long fallback = 0;
        e = model_degrees(plan, model->m, delta, deg, fallback);
        if (e.empty()) e = plan.start_items();
        if (e.empty()) {
            // the per-degree work's uploads on the stream of the phases
            // beside the items (they would otherwise wait for the items)
            plan.side_begin();
            e = plan.set_pair_degrees(deg.data());
            plan.side_end();
        }
        if (e.empty()) plan.set_model_fallback_pairs(fallback);
    }
    const auto t_model = std::chrono::steady_clock::now();
    if (e.empty()) e = plan.execute(true);
    const auto t_exec = std::chrono::steady_clock::now();
    if (e.empty()) e = plan.download(outputArray, device);
    if (!e.empty()) return fail(e);
    fill_stats(plan, &g_stats);
    g_have_stats = true;
    const auto t_down = std::chrono::steady_clock::now();
    plan.release();
    const auto t_end = std::chrono::steady_clock::now();
    if (verbosity > 3) {
        auto sec = [](auto a, auto b) { return std::chrono::duration<double>(b - a).count(); };
        std::printf("[BaryTree] call: plan %.6f, degree model %.6f, execution %.6f, download %.6f, release %.6f s\n",
                    sec(t_start, t_created), sec(t_created, t_model), sec(t_model, t_exec), sec(t_exec, t_down),
                    sec(t_down, t_end));
    }
    print_diagnostics(g_stats, numTargets, numSources, std::chrono::duration<double>(t_end - t_start).count(),
                      verbosity);
    return 0;
}

int BaryTree_SolveLearned(int numTargets, int numSources, const double *targetX,
                          const double *targetY, const double *targetZ, const double *targetValue,
                          const double *sourceX, const double *sourceY, const double *sourceZ,
                          const double *sourceValue, const double *sourceWeight,
                          double *outputArray, int kernel, int numKernelParams,
                          const double *kernelParams, int singularity, int approximation,
                          int compute_type, double theta, int interpOrder, int maxPerSourceLeaf,
                          int maxPerTargetLeaf, double sizeCheck, double beta, int verbosity,
                          int precision, const BaryTreeDegreeModel *model, float delta)
{
    return solve_learned(numTargets, numSources, targetX, targetY, targetZ, targetValue, sourceX, sourceY,
                         sourceZ, sourceValue, sourceWeight, outputArray, kernel, numKernelParams, kernelParams,
                         singularity, approximation, compute_type, theta, interpOrder, maxPerSourceLeaf,
                         maxPerTargetLeaf, sizeCheck, beta, verbosity, precision, model, delta, false);
}

int BaryTree_SolveLearnedDevice(int numTargets, int numSources, const double *targetX,
                                const double *targetY, const double *targetZ, const double *targetValue,
                                const double *sourceX, const double *sourceY, const double *sourceZ,
                                const double *sourceValue, const double *sourceWeight,
                                double *outputArray, int kernel, int numKernelParams,
                                const double *kernelParams, int singularity, int approximation,
                                int compute_type, double theta, int interpOrder, int maxPerSourceLeaf,
                                int maxPerTargetLeaf, double sizeCheck, double beta, int verbosity,
                                int precision, const BaryTreeDegreeModel *model, float delta)
{
    return solve_learned(numTargets, numSources, targetX, targetY, targetZ, targetValue, sourceX, sourceY,
                         sourceZ, sourceValue, sourceWeight, outputArray, kernel, numKernelParams, kernelParams,
                         singularity, approximation, compute_type, theta, interpOrder, maxPerSourceLeaf,
                         maxPerTargetLeaf, sizeCheck, beta, verbosity, precision, model, delta, true);
}

int BaryTree_SolveDevice(int numTargets, int numSources, const double *targetX, const double *targetY,
                         const double *targetZ, const double *targetValue, const double *sourceX,
                         const double *sourceY, const double *sourceZ, const double *sourceValue,
                         const double *sourceWeight, double *outputArray, int kernel,
                         int numKernelParams, const double *kernelParams, int singularity,
                         int approximation, int compute_type, double theta, int interpOrder,
                         int maxPerSourceLeaf, int maxPerTargetLeaf, double sizeCheck, double beta,
                         int verbosity, int precision)
{
    return solve(numTargets, numSources, targetX, targetY, targetZ, targetValue, sourceX, sourceY,
                 sourceZ, sourceValue, sourceWeight, outputArray, kernel, numKernelParams,
                 kernelParams, singularity, approximation, compute_type, theta, interpOrder,
                 maxPerSourceLeaf, maxPerTargetLeaf, sizeCheck, beta, verbosity, precision, true);
}


int BaryTree_AnalyzeDegrees(int numTargets, int numSources, const double *targetX,
                            const double *targetY, const double *targetZ, const double *sourceX,
                            const double *sourceY, const double *sourceZ, int kernel,
                            int numKernelParams, const double *kernelParams, int singularity,
                            int approximation, int compute_type, double theta, int interpOrder,
                            int maxPerSourceLeaf, int maxPerTargetLeaf, double sizeCheck,
                            double beta, int numPairs, const int *degrees, float *features,
                            BaryTreeStats *stats)
{
    g_last_error.clear();
    if (!stats) return fail("stats is NULL");
    bt::Params prm;
    bt::params_setup(prm, kernel, numKernelParams, kernelParams, approximation, singularity,
                     compute_type, theta, interpOrder, maxPerSourceLeaf, maxPerTargetLeaf, sizeCheck,
                     beta, 0);
    bt::Plan plan;
    std::vector<double> zs(numSources > 0 ? numSources : 0, 0.0);
    std::string e = plan.create(prm, numTargets, numSources, targetX, targetY, targetZ, nullptr,
                                sourceX, sourceY, sourceZ, zs.data(), nullptr, true);
    if (!e.empty()) return fail(e);
    const int n = (int)plan.far_pairs_list().size();
    // features and degrees both hold numPairs entries: a count that differs
    // from the plan's would overrun the caller's buffer or read past it
    if ((features || degrees) && numPairs != n)
        return fail("numPairs does not match the plan's far-field pairs (call with features and "
                    "degrees NULL to learn the count)");
    if (features) {
        std::vector<float> f;
        plan.pair_features(f);
        std::memcpy(features, f.data(), f.size() * sizeof(float));
    }
    if (degrees) {
        e = plan.set_pair_degrees(degrees);
        if (!e.empty()) return fail(e);
    }
    fill_stats(plan, stats);
    return 0;
}

}  // extern "C"
