// bt_capi.cpp -- extern "C" entry points of libBaryTree_gpu.so.
#include <cstdio>
#include <new>
#include <vector>
#include <cstring>
#include <string>

#include "../../include/BaryTreeInterface.h"
#include "bt_solver.h"

struct BaryTreePlan {
    bt::Plan plan;
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
}

int plan_create(bt::Plan &plan, int M, int N, const double *tx, const double *ty, const double *tz,
                const double *tq, const double *sx, const double *sy, const double *sz,
                const double *sq, const double *sw, int kernel, int nkp, const double *kp,
                int singularity, int approximation, int compute_type, double theta, int degree,
                int msl, int mtl, double size_check, double beta, int verbosity, int precision)
{
    bt::Params prm;
    bt::params_setup(prm, kernel, nkp, kp, approximation, singularity, compute_type, theta, degree,
                     msl, mtl, size_check, beta, verbosity);
    if (precision != BARYTREE_PRECISION_FP64 && precision != BARYTREE_PRECISION_MIXED)
        return fail("unknown precision mode");
    prm.precision = precision;
    if (verbosity > 0) {
        std::printf("[BaryTree]\n[BaryTree] RunParams: kernel=%d approximation=%d singularity=%d "
                    "compute_type=%d theta=%f degree=%d leaves=%d/%d size_check=%f beta=%f "
                    "precision=%d\n",
                    prm.kernel, prm.approximation, prm.singularity, prm.compute_type, prm.theta,
                    prm.degree, prm.max_source_leaf, prm.max_target_leaf, prm.size_check_factor,
                    prm.beta, prm.precision);
    }
    std::string e = plan.create(prm, M, N, tx, ty, tz, tq, sx, sy, sz, sq, sw);
    if (!e.empty()) return fail(e);
    return 0;
}

int solve(int M, int N, const double *tx, const double *ty, const double *tz, const double *tq,
          const double *sx, const double *sy, const double *sz, const double *sq, const double *sw,
          double *out, int kernel, int nkp, const double *kp, int singularity, int approximation,
          int compute_type, double theta, int degree, int msl, int mtl, double size_check,
          double beta, int verbosity, int precision)
{
    g_last_error.clear();
    if (!out) return fail("outputArray is NULL");
    if (M > 0) std::memset(out, 0, (size_t)M * sizeof(double));
    bt::Plan plan;
    int rc = plan_create(plan, M, N, tx, ty, tz, tq, sx, sy, sz, sq, sw, kernel, nkp, kp, singularity,
                         approximation, compute_type, theta, degree, msl, mtl, size_check, beta,
                         verbosity, precision);
    if (rc) return rc;
    std::string e = plan.execute(true);
    if (e.empty()) e = plan.download(out);
    if (!e.empty()) return fail(e);
    fill_stats(plan, &g_stats);
    g_have_stats = true;
    if (verbosity > 0) {
        std::printf("[BaryTree] host tree %.3fs lists %.3fs; gpu total %.3fs (near %.3fs far %.3fs)\n",
                    g_stats.host_tree_s, g_stats.host_lists_s, g_stats.gpu_total_s,
                    g_stats.gpu_near_s, g_stats.gpu_far_s);
    }
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

}  // extern "C"
