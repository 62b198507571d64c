/*
 * BaryTreeInterface.h -- C ABI of the MI355X BaryTree interaction path
 * (libBaryTree_gpu.so).
 *
 * BaryTreeInterface() has the reference's prototype and enum numbering and
 * replaces src/interface/BaryTreeInterface.h:52-59 (enums :7-47), so the
 * reference's ctypes binding (interfaces/python/BaryTreeInterface.py:47-55,
 * 134-152) and Fortran interface (interfaces/fortran/BaryTreeInterface.fh:
 * 14-39) bind to it unchanged.  The remaining entry points have no
 * counterpart in the reference; each says what it is for.
 *
 * All arrays are host pointers owned by the caller.  Nothing here modifies
 * caller input arrays (the reference permutes them during the call and
 * restores them on return; the observable result is the same).
 */
#ifndef BARYTREE_INTERFACE_H
#define BARYTREE_INTERFACE_H

#ifdef __cplusplus
extern "C" {
#endif

#ifndef H_BARYTREE_TYPES_H
#define H_BARYTREE_TYPES_H

typedef enum KERNEL {
    NO_KERNEL,
    COULOMB,
    YUKAWA,
    REGULARIZED_COULOMB,
    REGULARIZED_YUKAWA,
    ATAN,
    TCF,
    DCF,
    SIN_OVER_R,
    MQ,
    RBS_U,
    RBS_V,
    USER
} KERNEL;

typedef enum SINGULARITY { NO_SINGULARITY, SKIPPING, SUBTRACTION } SINGULARITY;

typedef enum APPROXIMATION { NO_APPROX, LAGRANGE, HERMITE } APPROXIMATION;

typedef enum COMPUTE_TYPE {
    NO_COMPUTE_TYPE,
    PARTICLE_CLUSTER,
    CLUSTER_PARTICLE,
    CLUSTER_CLUSTER,
} COMPUTE_TYPE;

#endif /* H_BARYTREE_TYPES_H */

/* Reference entry point (src/interface/BaryTreeInterface.h:52-59,
 * src/interface/BaryTreeInterface.c:18-62): zeroes outputArray, then adds the
 * treecode potential at every target, in the caller's target order.
 * Supported, in the combinations the reference dispatchers accept
 * (src/interaction_compute/interaction_compute_{pc,cp,cc}.c):
 *   LAGRANGE: COULOMB, YUKAWA, REGULARIZED_COULOMB with SKIPPING or
 *     SUBTRACTION in PARTICLE_CLUSTER, CLUSTER_PARTICLE, CLUSTER_CLUSTER
 *     (REGULARIZED_COULOMB SUBTRACTION in PARTICLE_CLUSTER only);
 *     REGULARIZED_YUKAWA, SIN_OVER_R, RBS_U, RBS_V, USER in all three
 *     compute types with SKIPPING (REGULARIZED_YUKAWA also PARTICLE_CLUSTER
 *     SUBTRACTION); ATAN and MQ in PARTICLE_CLUSTER with SKIPPING;
 *   HERMITE: COULOMB, YUKAWA, REGULARIZED_COULOMB in PARTICLE_CLUSTER (either
 *     singularity) and CLUSTER_PARTICLE (SKIPPING); SIN_OVER_R in
 *     PARTICLE_CLUSTER and CLUSTER_PARTICLE with SKIPPING; no CLUSTER_CLUSTER;
 * degree 1..12.  Where the reference prints an error and calls exit(1), this
 * prints the error (with the "[BaryTree]" prefix), leaves outputArray zero
 * and records the message for BaryTree_GetLastError(); the process keeps
 * running.  More coincident particles than a leaf size are refused the same
 * way (the reference's tree recursion does not end there).  verbosity > 0
 * prints start and finish, > 1 the total time, > 2 the reference's
 * interaction tables and this build's phase times; > 3 (the one-call
 * BaryTree_Solve* forms) also the wall time of the call's parts: plan,
 * degree model, execution, download, release. */
void BaryTreeInterface(int numTargets, int numSources,
                       double *targetX, double *targetY, double *targetZ, double *targetValue,
                       double *sourceX, double *sourceY, double *sourceZ, double *sourceValue,
                       double *sourceWeight, double *outputArray,
                       KERNEL kernel, int numKernelParams, double *kernelParams,
                       SINGULARITY singularity, APPROXIMATION approximation,
                       COMPUTE_TYPE compute_type, double theta, int interpOrder,
                       int maxPerSourceLeaf, int maxPerTargetLeaf, double sizeCheck,
                       double beta, int verbosity);

/* ---- Extensions (no counterpart in the reference) ------------------- */

/* Arithmetic of the pair kernels. */
#define BARYTREE_PRECISION_FP64 0  /* plain mode: FP64 throughout */
#define BARYTREE_PRECISION_MIXED 1 /* FP32 pair evaluation with compensated
                                      per-target accumulation; cluster charges,
                                      interpolation weights and totals FP64 */

/* Interaction counts (the reference's verbosity>2 diagnostics,
 * src/drivers/treedriver.c:191-200, 472-483, 763-792) and phase times. */
typedef struct BaryTreeStats {
    long long num_pp, num_pc, num_cp, num_cc;
    long long pp_ptwise, pc_ptwise, cp_ptwise, cc_ptwise;
    long long pc_replaced, cp_replaced, cc_replaced;
    int source_nodes, target_nodes, interp_degree, pts_per_cluster;
    double theta;
    int max_per_source_leaf, max_per_target_leaf;
    int near_tiles, far_tiles;
    double host_copy_s, host_tree_s, host_lists_s, host_work_s, h2d_s;
    /* GPU phases of the last execution: upward pass; near field (with the
       Lagrange particle-cluster far field, which shares its launch); the
       other far field (grids of target clusters, or the Hermite
       particle-cluster launch); downpass; correction and scatter.  When the
       symmetric near field's launch runs (sym_ptwise > 0, Lagrange), the
       upward pass, the far field and the downpass run on a second stream
       beside it: their times are then elapsed times while sharing the GPU
       with it, and gpu_total_s is the execution's elapsed time (less than
       the sum of the phases) */
    double gpu_upward_s, gpu_near_s, gpu_far_s, gpu_down_s, gpu_finish_s, gpu_total_s;
    int num_systems;         /* independent systems in the plan (1 unless batched) */
    int targets_per_thread;  /* near-field tile = 256 * targets_per_thread targets
                                (single-system plans: 1 below 262144 targets, else 2;
                                batched plans: the schedule's) */
    int persistent;          /* 1: persistent interaction launches */
    long long num_far_pairs; /* entries of the approximation lists */
    int variable_degree;     /* 1: far-field pairs carry their own degree */
    double mean_pair_degree; /* mean degree over far-field pairs (variable degree) */
    long long far_ptwise_variable; /* point-wise evaluations of the far field (variable degree) */
    long long sym_ptwise;    /* near-field particle pairs (ordered) that the
                                symmetric near field delivers (either precision):
                                a single system whose targets are its sources in
                                the same tree order (Coulomb, Yukawa, or
                                regularized Coulomb with skipping; Lagrange or
                                Hermite); pairs of
                                two leaves in each other's near list are
                                evaluated once for both particles, the rest once
                                for their target; 0 otherwise or when switched
                                off (BaryTreePlan_SetSymmetric) */
    double gpu_near_sym_s;   /* the symmetric near-field launch alone (within gpu_near_s) */
    long long sym_slots;     /* FP64 partial sums the symmetric launch writes (8 bytes
                                of device memory each, allocated at the first
                                execution that uses them); 0 when sym_ptwise is 0 */
    long long model_fallback_pairs; /* far-field pairs a degree model left at the
                                fixed degree: their features lie outside the
                                model's training range (variable degree only) */
    long long sym_evals;     /* pair evaluations the symmetric near field
                                executes for its sym_ptwise pairs (between half
                                of sym_ptwise and sym_ptwise) */
} BaryTreeStats;

/* For callers that want an error code instead of a printed message:
 * BaryTreeInterface semantics plus a precision choice.  Returns 0 on
 * success, a negative value on error (message in BaryTree_GetLastError). */
int BaryTree_Solve(int numTargets, int numSources,
                   const double *targetX, const double *targetY, const double *targetZ,
                   const double *targetValue, const double *sourceX, const double *sourceY,
                   const double *sourceZ, const double *sourceValue, const double *sourceWeight,
                   double *outputArray, int kernel, int numKernelParams,
                   const double *kernelParams, int singularity, int approximation,
                   int compute_type, double theta, int interpOrder, int maxPerSourceLeaf,
                   int maxPerTargetLeaf, double sizeCheck, double beta, int verbosity,
                   int precision);

/* BaryTree_Solve for callers whose particles already live in GPU memory:
 * every array argument (targets, sources, outputArray) is a device pointer
 * on the current GPU; nothing crosses PCIe but the trees' small node
 * records.  Results equal BaryTree_Solve's bit for bit. */
int BaryTree_SolveDevice(int numTargets, int numSources, const double *targetX, const double *targetY,
                         const double *targetZ, const double *targetValue, const double *sourceX,
                         const double *sourceY, const double *sourceZ, const double *sourceValue,
                         const double *sourceWeight, double *outputArray, int kernel,
                         int numKernelParams, const double *kernelParams, int singularity,
                         int approximation, int compute_type, double theta, int interpOrder,
                         int maxPerSourceLeaf, int maxPerTargetLeaf, double sizeCheck, double beta,
                         int verbosity, int precision);

/* Message of the last failed call on this thread ("" if none). */
const char *BaryTree_GetLastError(void);

/* Leaf sizes the beta auto-tune (0 <= beta <= 1) picks for later calls on
 * this thread, as src/run_params/run_params.c:100-117 sets them:
 * BARYTREE_BETA_LEAVES_CPU (the default) those of the reference CPU build
 * (PC 50/5, CP 5/50, CC 50/50), so beta runs reproduce the reference CPU
 * results; BARYTREE_BETA_LEAVES_GPU those of the reference GPU build
 * (3000/3000), far faster here.  Returns 0, or -1 for another value. */
#define BARYTREE_BETA_LEAVES_CPU 0
#define BARYTREE_BETA_LEAVES_GPU 1
int BaryTree_SetBetaLeafSizes(int which);

/* Device memory cache.  The buffers of a plan (and of a one-call solve's
 * internal plan) go back to a per-GPU cache when it is destroyed, and later
 * plans take blocks of a close size from it instead of allocating: a
 * one-call solve at N = 16M otherwise spends milliseconds unmapping and
 * mapping several GB per call.  At most the limit (default 32 GiB) stays
 * cached per GPU.  BaryTree_EmptyDeviceCache frees every cached block and
 * returns the bytes freed; BaryTree_SetDeviceCacheLimit sets the limit (0:
 * nothing is kept) and frees what lies beyond it; BaryTree_CachedDeviceBytes
 * reports the current GPU's cached bytes.  No counterpart in the reference
 * (its GPU build allocates through OpenACC data regions per call). */
unsigned long long BaryTree_EmptyDeviceCache(void);
int BaryTree_SetDeviceCacheLimit(unsigned long long bytes);
unsigned long long BaryTree_CachedDeviceBytes(void);

/* Statistics of the last successful BaryTreeInterface/BaryTree_Solve call on
 * this thread.  Returns 0, or -1 if there was none. */
int BaryTree_GetLastStats(BaryTreeStats *stats);

/* For callers that size a run before launching it: builds the trees and
 * interaction lists on the host only (no GPU work) and reports the counts
 * and host times of BaryTree_Solve with the same arguments.  Returns 0 or a
 * negative value. */
int BaryTree_Analyze(int numTargets, int numSources,
                     const double *targetX, const double *targetY, const double *targetZ,
                     const double *sourceX, const double *sourceY, const double *sourceZ,
                     int kernel, int numKernelParams, const double *kernelParams,
                     int singularity, int approximation, int compute_type, double theta,
                     int interpOrder, int maxPerSourceLeaf, int maxPerTargetLeaf,
                     double sizeCheck, double beta, BaryTreeStats *stats);

/* For callers that re-run one particle configuration (and for timing the GPU
 * path with its inputs resident in HBM): Create does the host work (trees,
 * lists) and uploads everything; Execute runs the GPU path and, if
 * outputArray is not NULL, copies the potential to it (caller order). */
typedef struct BaryTreePlan BaryTreePlan;

int BaryTreePlan_Create(BaryTreePlan **plan, int numTargets, int numSources,
                        const double *targetX, const double *targetY, const double *targetZ,
                        const double *targetValue, const double *sourceX,
                        const double *sourceY, const double *sourceZ,
                        const double *sourceValue, const double *sourceWeight, int kernel,
                        int numKernelParams, const double *kernelParams, int singularity,
                        int approximation, int compute_type, double theta, int interpOrder,
                        int maxPerSourceLeaf, int maxPerTargetLeaf, double sizeCheck,
                        double beta, int precision);
int BaryTreePlan_Execute(BaryTreePlan *plan, double *outputArray);
/* Switches the pair arithmetic of later Execute calls (BARYTREE_PRECISION_*)
 * without rebuilding the plan. */
int BaryTreePlan_SetPrecision(BaryTreePlan *plan, int precision);
/* The symmetric near field of later Execute calls: 1 (the default) evaluates
 * each near-field pair of two distinct leaves once for both particles where
 * the plan allows it (see sym_ptwise); 0 keeps one evaluation per ordered
 * pair, as src/interaction_compute/interaction_compute_cc.c:1019-1200 does.
 * Results differ only by the order of the sums (FP64: rounding). */
int BaryTreePlan_SetSymmetric(BaryTreePlan *plan, int enable);
int BaryTreePlan_GetStats(const BaryTreePlan *plan, BaryTreeStats *stats);
/* FP64 direct sum on the GPU (directdriver semantics,
 * src/drivers/directdriver.c:23-122) at m targets given by caller-order
 * indices; for error checks of the plan's result. */
int BaryTreePlan_DirectSample(BaryTreePlan *plan, int m, const int *targetIndex, double *out);
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  BaryTreePlan_DirectSample without any device
 * allocation: its device buffers live in one fixed scratch array per GPU
 * (5 MiB, held until the process ends), so m is at most 8192 (larger m is
 * an error, nothing computed).  Results equal BaryTreePlan_DirectSample's
 * bit for bit.  All calls on one GPU share that array: calls that overlap
 * on one GPU (other host threads) race and must be serialised by the
 * caller. */
int BaryTreePlan_DirectSampleNoAlloc(BaryTreePlan *plan, int m, const int *targetIndex, double *out);
void BaryTreePlan_Destroy(BaryTreePlan *plan);

/* ---- Batching: many independent systems in one set of launches -------
 *
 * System s owns targets [targetOffsets[s], targetOffsets[s+1]) and sources
 * [sourceOffsets[s], sourceOffsets[s+1]) of the particle arrays
 * (offsets[0] = 0, numSystems+1 entries, every range non-empty); its targets
 * see only its own sources.  Each system is solved as BaryTreeInterface
 * would solve it alone with its leaf sizes; all systems share the kernel,
 * approximation, theta, degree and precision, and go through every GPU phase
 * in one launch per phase (one interaction launch walks all their lists).
 *
 * Leaf sizes: maxPerSourceLeaf/maxPerTargetLeaf arrays (numSystems entries)
 * if given, else the schedule's leaf size for each system's size class,
 * else a schedule from BaryTree_Autotune (cached per kernel, singularity,
 * compute type, theta, degree and precision for the life of the process). */
#define BARYTREE_SIZE_CLASSES 32
typedef struct BaryTreeSchedule {
    int targets_per_thread;            /* 1 or 2 (near-field tiles of 256 or 512 targets) */
    int persistent;                    /* 1: resident blocks pop tiles from per-XCD queues */
    int leaf[BARYTREE_SIZE_CLASSES];   /* leaf size of systems with
                                          2^c <= max(targets, sources) < 2^(c+1); 0 = unset */
    int trials;                        /* plans the autotuner timed */
    double tune_seconds;               /* wall time the autotuner spent */
} BaryTreeSchedule;

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
                          int verbosity, int precision);

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
                               double sizeCheck, double beta, int precision);

/* BaryTree_Analyze for a batch: host-only trees and lists of every system,
 * counts summed over the systems.  Leaf sizes as BaryTree_SolveBatched, but
 * without the autotuner (it needs the GPU): arrays or a schedule are
 * required unless beta is in [0, 1]. */
int BaryTree_AnalyzeBatched(int numSystems, const int *targetOffsets, const int *sourceOffsets,
                            const double *targetX, const double *targetY, const double *targetZ,
                            const double *sourceX, const double *sourceY, const double *sourceZ,
                            int kernel, int numKernelParams, const double *kernelParams,
                            int singularity, int approximation, int compute_type, double theta,
                            int interpOrder, const int *maxPerSourceLeaf,
                            const int *maxPerTargetLeaf, const BaryTreeSchedule *schedule,
                            double sizeCheck, double beta, BaryTreeStats *stats);

/* ---- Learned degree selection ---------------------------------------
 *
 * Far-field pairs are the entries of the approximation lists: PC (a source
 * cluster's charges act on target particles), CP (source particles act on a
 * target grid) and CC (a source cluster acts on a target grid).  Each pair
 * may carry its own interpolation degree: a source node used at degree p
 * gets modified charges of degree p from its particles, a target node used
 * at degree p gets a grid of degree p interpolated to its particles.  A
 * degree model (an MLP over pair features) picks the degrees; without one,
 * every pair uses the plan's degree (the reference's method).
 *
 * Pair order: per system, PC: target batches in order, each batch's
 * approximation list; CP: source batches in order, each batch's list; CC:
 * target nodes in order, each node's CC, then PC, then CP entries.
 * Pair kinds: 1 PC, 2 CP, 3 CC. */
#define BARYTREE_PAIR_FEATURES 8
#define BARYTREE_LABEL_DEGREES 11  /* degrees 2..12 */
#define BARYTREE_MODEL_PARAMS 9546 /* 8 -> 64 -> 64 -> 64 -> 10 */

int BaryTreePlan_NumPairs(const BaryTreePlan *plan);
/* kind (1 PC, 2 CP, 3 CC) and system of each pair (either may be NULL) */
int BaryTreePlan_PairInfo(const BaryTreePlan *plan, int *kind, int *system);
/* BARYTREE_PAIR_FEATURES floats per pair: r_t/d, r_s/d, log2(n_t)/16,
 * log2(n_s)/16, [PC], [CP], kernel parameter at scale d (Yukawa kappa*d,
 * regularized Coulomb epsilon/d, Coulomb 0), largest half-width of the
 * interpolated box(es)/d; d = distance of the node centres. */
int BaryTreePlan_PairFeatures(const BaryTreePlan *plan, float *features);
/* Labels of pairs [first, first+count): the least degree in 2..12 whose pair
 * error is <= tol, 13 if none.  Pair error at degree p: relative L2 error,
 * over up to 32 stride-sampled targets of the target node, of the degree-p
 * approximation of the potential of up to 128 stride-sampled sources of the
 * source node (FP64; the plan's kernel with skipping, or with subtraction
 * both of its terms with the targets' charges and the sources' weights).
 * errors (optional): 11 per pair, degrees 2..12, NaN beyond the label. */
int BaryTreePlan_LabelPairs(BaryTreePlan *plan, int first, int count, double tol, int *labels,
                            double *errors);
/* One degree (1..12) per pair for later Execute calls; NULL restores the
 * plan's fixed degree. */
int BaryTreePlan_SetPairDegrees(BaryTreePlan *plan, const int *degrees);

/* Host only (no GPU work): the far-field pairs of the plan BaryTree_Analyze
 * would build, their features (optional, BARYTREE_PAIR_FEATURES per pair),
 * and, if degrees is not NULL (numPairs entries), the variable-degree work
 * built and checked for those degrees (stats: variable_degree,
 * mean_pair_degree, far_ptwise_variable).  Call first with degrees and
 * features NULL to learn the count (stats->num_far_pairs); when either is
 * given, numPairs must equal that count (else an error, nothing written):
 * features then holds numPairs * BARYTREE_PAIR_FEATURES floats, degrees
 * numPairs ints. */
int BaryTree_AnalyzeDegrees(int numTargets, int numSources, const double *targetX,
                            const double *targetY, const double *targetZ, const double *sourceX,
                            const double *sourceY, const double *sourceZ, int kernel,
                            int numKernelParams, const double *kernelParams, int singularity,
                            int approximation, int compute_type, double theta, int interpOrder,
                            int maxPerSourceLeaf, int maxPerTargetLeaf, double sizeCheck,
                            double beta, int numPairs, const int *degrees, float *features,
                            BaryTreeStats *stats);

/* The same labels for explicitly given pairs: pair i has targets
 * [targetBegin[i], +targetCount[i]) and sources [sourceBegin[i],
 * +sourceCount[i]) of the arrays, and node boxes {xmin,xmax,ymin,ymax,zmin,
 * zmax} (6 doubles per pair).  kernel 1/2/3 with kernelParam kappa or
 * epsilon. */
int BaryTree_LabelPairs(int numPairs, const int *kind, const int *targetBegin,
                        const int *targetCount, const int *sourceBegin, const int *sourceCount,
                        const double *targetBox, const double *sourceBox, int numTargets,
                        const double *targetX, const double *targetY, const double *targetZ,
                        int numSources, const double *sourceX, const double *sourceY,
                        const double *sourceZ, const double *sourceValue, int kernel,
                        double kernelParam, double tol, int *labels, double *errors);

/* BaryTree_LabelPairs with the subtraction kernel (both of its terms, as
 * BaryTreePlan_LabelPairs labels subtraction plans): the targets' charges
 * targetValue and the sources' weights sourceWeight; kernel 1 or 2 with
 * kernelParams {kappa}, 3 (regularized Coulomb) with {alpha, epsilon}. */
int BaryTree_LabelPairsSubtraction(int numPairs, const int *kind, const int *targetBegin,
                                   const int *targetCount, const int *sourceBegin, const int *sourceCount,
                                   const double *targetBox, const double *sourceBox, int numTargets,
                                   const double *targetX, const double *targetY, const double *targetZ,
                                   const double *targetValue, int numSources, const double *sourceX,
                                   const double *sourceY, const double *sourceZ, const double *sourceValue,
                                   const double *sourceWeight, int kernel, int numKernelParams,
                                   const double *kernelParams, double tol, int *labels, double *errors);

/* The degree model: MLP 8 -> 64 -> 64 -> 64 -> 10 (ReLU, FP32).  Output k
 * is the logit of "the pair needs a degree above 2+k".  Training (Adam on
 * the batch-mean binary cross-entropy of the 10 outputs) and inference run
 * in HIP kernels on the current GPU. */
typedef struct BaryTreeDegreeModel BaryTreeDegreeModel;
int BaryTreeModel_Create(BaryTreeDegreeModel **model, unsigned long long seed);
int BaryTreeModel_Load(BaryTreeDegreeModel **model, const char *path);
int BaryTreeModel_Save(const BaryTreeDegreeModel *model, const char *path);
/* parameters (BARYTREE_MODEL_PARAMS floats: W1[64][8] b1 W2[64][64] b2
 * W3[64][64] b3 W4[10][64] b4) and feature normalisation (x - mean)*scale */
int BaryTreeModel_GetParams(const BaryTreeDegreeModel *model, float *params, float *featureMean,
                            float *featureScale);
int BaryTreeModel_SetParams(BaryTreeDegreeModel *model, const float *params,
                            const float *featureMean, const float *featureScale);
/* batch-mean loss and gradient over n samples (one batch) */
int BaryTreeModel_Gradient(const BaryTreeDegreeModel *model, int n, const float *features,
                           const int *labels, float *grad, double *loss);
/* epochs of shuffled mini-batches; sets the feature normalisation from the
 * data if the model has none; epochLoss (optional): mean loss per epoch */
int BaryTreeModel_Train(BaryTreeDegreeModel *model, int n, const float *features,
                        const int *labels, int epochs, int batch, float learningRate,
                        unsigned long long seed, float *epochLoss);
/* degree = 3+k for the largest k with sigmoid(logit_k) >= delta, else 2
 * (0 < delta < 1, else an error); logits (optional): 10 per sample */
int BaryTreeModel_Predict(const BaryTreeDegreeModel *model, int n, const float *features,
                          float delta, int *degrees, float *logits);
/* The range of every feature over the model's training data
 * (BARYTREE_PAIR_FEATURES floats each).  Training widens it to cover its
 * data; BaryTreeModel_Save/Load keep it (files written before it was
 * recorded load without one).  A pair whose features lie outside it by more
 * than 5% of its width keeps the plan's fixed degree
 * (BaryTreePlan_UseDegreeModel, BaryTree_SolveLearned; counted in
 * model_fallback_pairs).  Get returns 1 if the model has a range, else 0
 * (then there is no guard); Set with lo or hi NULL removes it. */
int BaryTreeModel_GetFeatureRange(const BaryTreeDegreeModel *model, float *lo, float *hi);
int BaryTreeModel_SetFeatureRange(BaryTreeDegreeModel *model, const float *lo, const float *hi);
void BaryTreeModel_Destroy(BaryTreeDegreeModel *model);
/* features -> model -> SetPairDegrees */
int BaryTreePlan_UseDegreeModel(BaryTreePlan *plan, const BaryTreeDegreeModel *model, float delta);

/* BaryTree_Solve with the far-field degrees picked by a model (NULL: the
 * fixed degree interpOrder, i.e. BaryTree_Solve). */
int BaryTree_SolveLearned(int numTargets, int numSources, const double *targetX,
                          const double *targetY, const double *targetZ, const double *targetValue,
                          const double *sourceX, const double *sourceY, const double *sourceZ,
                          const double *sourceValue, const double *sourceWeight,
                          double *outputArray, int kernel, int numKernelParams,
                          const double *kernelParams, int singularity, int approximation,
                          int compute_type, double theta, int interpOrder, int maxPerSourceLeaf,
                          int maxPerTargetLeaf, double sizeCheck, double beta, int verbosity,
                          int precision, const BaryTreeDegreeModel *model, float delta);

/* BaryTree_SolveLearned with device pointers for every particle array and
 * outputArray (as BaryTree_SolveDevice); kernelParams stays a host array. */
int BaryTree_SolveLearnedDevice(int numTargets, int numSources, const double *targetX,
                                const double *targetY, const double *targetZ, const double *targetValue,
                                const double *sourceX, const double *sourceY, const double *sourceZ,
                                const double *sourceValue, const double *sourceWeight,
                                double *outputArray, int kernel, int numKernelParams,
                                const double *kernelParams, int singularity, int approximation,
                                int compute_type, double theta, int interpOrder, int maxPerSourceLeaf,
                                int maxPerTargetLeaf, double sizeCheck, double beta, int verbosity,
                                int precision, const BaryTreeDegreeModel *model, float delta);

/* Autotuner: picks the leaf size of every size class present in the batch
 * (candidates P/2, P, 2P, 4P, 8P with P = (interpOrder+1)^3), then the tile
 * shape (targets per thread) and the launch form (one-shot or persistent),
 * by timing GPU plans of up to 8 sampled systems per class, replicated to
 * about 2^20 targets.  Particles and charges are those of the batch.
 * Returns 0 or a negative value. */
int BaryTree_Autotune(int numSystems, const int *targetOffsets, const int *sourceOffsets,
                      const double *targetX, const double *targetY, const double *targetZ,
                      const double *targetValue, const double *sourceX, const double *sourceY,
                      const double *sourceZ, const double *sourceValue,
                      const double *sourceWeight, int kernel, int numKernelParams,
                      const double *kernelParams, int singularity, int approximation,
                      int compute_type, double theta, int interpOrder, double sizeCheck,
                      double beta, int precision, BaryTreeSchedule *schedule);


#ifdef __cplusplus
}
#endif

#endif /* BARYTREE_INTERFACE_H */
