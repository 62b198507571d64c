// bt_probe.hip -- issue cost of the VALU instruction kinds the pair kernels
// are made of, measured on the running GPU.
//
// Each probe kernel runs 8 independent dependency chains per lane of one
// instruction kind at full occupancy (8 workgroups of 256 lanes per CU), so
// the time per wave instruction and SIMD is the issue cost of that kind.
// Multiplying a kernel's per-kind instruction counts (SQ_INSTS_VALU_*
// counters) by these costs gives the time the VALU needs to issue them:
// the attainable ceiling of that kernel's instruction mix.
//
// Also: the FP64 reciprocal square root of the pair kernels (bt_rsq.h)
// evaluated on given inputs, for the accuracy test of that function.
//
// Measurement tooling for the benchmark: built into its own library
// (libbt_probe.so), not into the drop-in libBaryTree_gpu.so.
#include <hip/hip_runtime.h>

#include "bt_rsq.h"

namespace btk {

// FP64 fma, FP64 add, FP32 fma, FP32 mul, FP32 exp2, FP32 rsq and
// FP64<->FP32 conversion, in that order
constexpr int PROBE_KINDS = 7;

namespace {

constexpr int PROBE_ITER = 2048;
constexpr int PROBE_CHAINS = 8;

// kinds: 0 FP64 fma, 1 FP64 add, 2 FP32 fma, 3 FP32 mul, 4 FP32 exp2,
//        5 FP32 rsq, 6 FP64 -> FP32 -> FP64 conversions and an FP64 fma (the
//        fma keeps the chain from folding; its cost is taken off on the host)
template <int K>
__global__ __launch_bounds__(256) void k_probe(float *outf, double *outd, float seed)
{
    float f[PROBE_CHAINS];
    double d[PROBE_CHAINS];
#pragma unroll
    for (int i = 0; i < PROBE_CHAINS; i++) {
        f[i] = seed + 0.01f * i + 1e-6f * threadIdx.x;
        d[i] = f[i];
    }
    for (int it = 0; it < PROBE_ITER; it++) {
#pragma unroll
        for (int i = 0; i < PROBE_CHAINS; i++) {
            if constexpr (K == 0) d[i] = fma(d[i], 0.999, 0.001);
            if constexpr (K == 1) d[i] = d[i] + 0.001;
            if constexpr (K == 2) f[i] = fmaf(f[i], 0.999f, 0.001f);
            if constexpr (K == 3) f[i] = f[i] * 0.999f;
            if constexpr (K == 4) f[i] = -__builtin_amdgcn_exp2f(f[i]);   // fixed point near -0.64
            if constexpr (K == 5) f[i] = __builtin_amdgcn_rsqf(f[i]);     // fixed point 1
            if constexpr (K == 6) { f[i] = (float)d[i]; d[i] = fma((double)f[i], 0.999, 0.001); }
        }
    }
    float s = 0.0f;
    double t = 0.0;
#pragma unroll
    for (int i = 0; i < PROBE_CHAINS; i++) { s += f[i]; t += d[i]; }
    const size_t g = (size_t)blockIdx.x * 256 + threadIdx.x;
    outf[g] = s;
    outd[g] = t;
}

template <int K>
hipError_t probe_one(int nb, float *of, double *od, hipEvent_t a, hipEvent_t b, float *ms)
{
    k_probe<K><<<nb, 256>>>(of, od, 0.5f);  // warm-up
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if ((e = hipEventRecord(a)) != hipSuccess) return e;
    for (int r = 0; r < 4; r++) k_probe<K><<<nb, 256>>>(of, od, 0.5f);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = hipEventRecord(b)) != hipSuccess) return e;
    if ((e = hipEventSynchronize(b)) != hipSuccess) return e;
    if ((e = hipEventElapsedTime(ms, a, b)) != hipSuccess) return e;
    *ms /= 4.0f;
    return hipSuccess;
}

}  // namespace

hipError_t probe_valu_issue(double *ns_per_instr, int n)
{
    int dev = 0, cus = 0;
    hipError_t e;
    if ((e = hipGetDevice(&dev)) != hipSuccess) return e;
    if ((e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess) return e;
    const int nb = cus * 8;
    float *of = nullptr;
    double *od = nullptr;
    hipEvent_t a = nullptr, b = nullptr;
    float ms[PROBE_KINDS] = {};
    if ((e = hipMalloc(&of, (size_t)nb * 256 * sizeof(float))) != hipSuccess) return e;
    if ((e = hipMalloc(&od, (size_t)nb * 256 * sizeof(double))) == hipSuccess &&
        (e = hipEventCreate(&a)) == hipSuccess && (e = hipEventCreate(&b)) == hipSuccess) {
        if (e == hipSuccess) e = probe_one<0>(nb, of, od, a, b, &ms[0]);
        if (e == hipSuccess) e = probe_one<1>(nb, of, od, a, b, &ms[1]);
        if (e == hipSuccess) e = probe_one<2>(nb, of, od, a, b, &ms[2]);
        if (e == hipSuccess) e = probe_one<3>(nb, of, od, a, b, &ms[3]);
        if (e == hipSuccess) e = probe_one<4>(nb, of, od, a, b, &ms[4]);
        if (e == hipSuccess) e = probe_one<5>(nb, of, od, a, b, &ms[5]);
        if (e == hipSuccess) e = probe_one<6>(nb, of, od, a, b, &ms[6]);
    }
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    (void)hipFree(of);
    if (od) (void)hipFree(od);
    if (e != hipSuccess) return e;
    // wave instructions per SIMD (4 SIMDs per CU, 4 waves per workgroup)
    const double per_kind_steps = (double)PROBE_ITER * PROBE_CHAINS;
    const double waves_per_simd = (double)nb * 4 / (cus * 4.0);
    double ns[PROBE_KINDS];
    for (int k = 0; k < PROBE_KINDS; k++) ns[k] = ms[k] * 1e6 / (waves_per_simd * per_kind_steps);
    ns[6] = 0.5 * (ns[6] - ns[0]);  // two conversions per step besides the fma
    for (int k = 0; k < n && k < PROBE_KINDS; k++) ns_per_instr[k] = ns[k];
    return hipSuccess;
}

__global__ void k_rsq64_eval(int n, const double *x, double *y)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = rsq64(x[i]);
}

}  // namespace btk

// ns per wave instruction and SIMD of each kind on the current GPU; returns
// the number of kinds filled, or -1 on a HIP error.
extern "C" int bt_probe_valu(double *ns_per_instr, int n)
{
    if (!ns_per_instr || n <= 0) return -1;
    const int k = n < btk::PROBE_KINDS ? n : btk::PROBE_KINDS;
    return btk::probe_valu_issue(ns_per_instr, k) == hipSuccess ? k : -1;
}

// y[i] = rsq64(x[i]) on the current GPU for host arrays x, y of n values;
// returns 0, or -1 on a HIP error or bad arguments.
extern "C" int bt_probe_rsq64(const double *x, double *y, int n)
{
    if (!x || !y || n <= 0) return -1;
    double *dx = nullptr, *dy = nullptr;
    const size_t bytes = (size_t)n * sizeof(double);
    hipError_t e = hipMalloc(&dx, bytes);
    if (e == hipSuccess) e = hipMalloc(&dy, bytes);
    if (e == hipSuccess) e = hipMemcpy(dx, x, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        btk::k_rsq64_eval<<<(n + 255) / 256, 256>>>(n, dx, dy);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(y, dy, bytes, hipMemcpyDeviceToHost);
    if (dx) (void)hipFree(dx);
    if (dy) (void)hipFree(dy);
    return e == hipSuccess ? 0 : -1;
}
