// bt_degree.hip -- change of degree of tensor Chebyshev grids, for plans
// whose far-field pairs carry their own degrees.
//
// A node's grid of degree D (nd = D+1 points per dimension on the node's
// box) holds modified charges Q^D_m = sum_i q_i L^D_m(x_i).  The grid of
// degree p <= D on the same box follows as
//   Q^p_k = sum_m E_p[k1][m1] E_p[k2][m2] E_p[k3][m3] Q^D_m,
// E_p[k][m] = l^p_k(t^D_m), the 1D Lagrange basis of degree p at the
// degree-D Chebyshev points of the reference interval (both grids are the
// same affine image of it).  This is exact in exact arithmetic: degree-D
// interpolation reproduces the degree-p basis polynomials.  The transpose
// carries degree-p grid values onto the degree-D grid of the same node,
// whose interpolant at any point then equals the degree-p interpolant there.
// Grids are stored x fastest: j = k1 + n k2 + n^2 k3.
#include <hip/hip_runtime.h>

#include "bt_kernels.h"

namespace btk {
namespace {

constexpr int DTPB = 256;
constexpr int NLMAX = MAX_DEGREE + 1;
constexpr int PMAX = NLMAX * NLMAX * NLMAX;

// out = (E x E x E) in, in of size nd^3, out of size np^3 (E: np x nd)
__device__ void contract_up(const double *E, int np, int nd, const double *in, double *t1, double *t2,
                            double *out)
{
    const int tid = threadIdx.x;
    for (int i = tid; i < np * nd * nd; i += DTPB) {  // t1[k1, m2, m3]
        const int k1 = i % np, r = i / np;
        double s = 0.0;
        for (int m1 = 0; m1 < nd; m1++) s = fma(E[k1 * nd + m1], in[m1 + nd * r], s);
        t1[i] = s;
    }
    __syncthreads();
    for (int i = tid; i < np * np * nd; i += DTPB) {  // t2[k1, k2, m3]
        const int k1 = i % np, k2 = (i / np) % np, m3 = i / (np * np);
        double s = 0.0;
        for (int m2 = 0; m2 < nd; m2++) s = fma(E[k2 * nd + m2], t1[k1 + np * (m2 + nd * m3)], s);
        t2[i] = s;
    }
    __syncthreads();
    for (int i = tid; i < np * np * np; i += DTPB) {  // out[k1, k2, k3]
        const int k12 = i % (np * np), k3 = i / (np * np);
        double s = 0.0;
        for (int m3 = 0; m3 < nd; m3++) s = fma(E[k3 * nd + m3], t2[k12 + np * np * m3], s);
        out[i] = s;
    }
    __syncthreads();
}

// acc += (E^T x E^T x E^T) phi, phi of size np^3, acc of size nd^3
__device__ void contract_down(const double *E, int np, int nd, const double *phi, double *u1, double *u2,
                              double *acc)
{
    const int tid = threadIdx.x;
    for (int i = tid; i < nd * np * np; i += DTPB) {  // u1[m1, k2, k3]
        const int m1 = i % nd, r = i / nd;
        double s = 0.0;
        for (int k1 = 0; k1 < np; k1++) s = fma(E[k1 * nd + m1], phi[k1 + np * r], s);
        u1[i] = s;
    }
    __syncthreads();
    for (int i = tid; i < nd * nd * np; i += DTPB) {  // u2[m1, m2, k3]
        const int m1 = i % nd, m2 = (i / nd) % nd, k3 = i / (nd * nd);
        double s = 0.0;
        for (int k2 = 0; k2 < np; k2++) s = fma(E[k2 * nd + m2], u1[m1 + nd * (k2 + np * k3)], s);
        u2[i] = s;
    }
    __syncthreads();
    for (int i = tid; i < nd * nd * nd; i += DTPB) {  // acc[m1, m2, m3]
        const int m12 = i % (nd * nd), m3 = i / (nd * nd);
        double s = acc[i];
        for (int k3 = 0; k3 < np; k3++) s = fma(E[k3 * nd + m3], u2[m12 + nd * nd * k3], s);
        acc[i] = s;
    }
    __syncthreads();
}

__global__ __launch_bounds__(DTPB) void k_degree_up(DegreeUpArgs a)
{
    __shared__ double in[PMAX], t1[PMAX], t2[PMAX], E[NLMAX * NLMAX];
    const int slot = blockIdx.x, tid = threadIdx.x;
    const int np = a.p + 1, nd = a.D + 1;
    const int PD = nd * nd * nd, Pp = np * np * np;
    for (int i = tid; i < np * nd; i += DTPB) E[i] = a.E[i];
    const size_t node = (size_t)a.node[slot];
    for (int plane = 0; plane < (a.subtraction ? 2 : 1); plane++) {
        const double *src = plane == 0 ? a.QD : a.WD;
        double *dst = plane == 0 ? a.Qp : a.Wp;
        __syncthreads();
        for (int i = tid; i < PD; i += DTPB) in[i] = src[node * PD + i];
        __syncthreads();
        contract_up(E, np, nd, in, t1, t2, dst + (size_t)slot * Pp);
    }
}

__global__ __launch_bounds__(DTPB) void k_degree_down(DegreeDownArgs a)
{
    __shared__ double acc[PMAX], phi[PMAX], u1[PMAX], u2[PMAX], E[NLMAX * NLMAX];
    const int b = blockIdx.x, tid = threadIdx.x;
    const int nd = a.D + 1, PD = nd * nd * nd;
    const size_t node = (size_t)a.node[b];
    for (int plane = 0; plane < (a.subtraction ? 2 : 1); plane++) {
        for (int i = tid; i < PD; i += DTPB) acc[i] = 0.0;
        for (int k = a.off[b]; k < a.off[b + 1]; k++) {
            const int p = a.item[k] >> DEG_SHIFT, slot = a.item[k] & SLOT_MASK;
            const int np = p + 1, Pp = np * np * np;
            const double *src = (plane == 0 ? a.Qs[p] : a.Ws[p]) + (size_t)slot * Pp;
            __syncthreads();
            for (int i = tid; i < np * nd; i += DTPB) E[i] = a.E[p][i];
            for (int i = tid; i < Pp; i += DTPB) phi[i] = src[i];
            __syncthreads();
            contract_down(E, np, nd, phi, u1, u2, acc);
        }
        double *dst = (plane == 0 ? a.QD : a.WD) + node * PD;
        for (int i = tid; i < PD; i += DTPB) dst[i] = acc[i];
        __syncthreads();
    }
}

}  // namespace

hipError_t launch_degree_up(const DegreeUpArgs &a, hipStream_t st)
{
    if (a.nslots <= 0) return hipSuccess;
    if (a.p < 1 || a.p > a.D || a.D > MAX_DEGREE) return hipErrorInvalidValue;
    k_degree_up<<<dim3(a.nslots), dim3(DTPB), 0, st>>>(a);
    return hipGetLastError();
}

hipError_t launch_degree_down(const DegreeDownArgs &a, hipStream_t st)
{
    if (a.nnodes <= 0) return hipSuccess;
    if (a.D < 1 || a.D > MAX_DEGREE) return hipErrorInvalidValue;
    k_degree_down<<<dim3(a.nnodes), dim3(DTPB), 0, st>>>(a);
    return hipGetLastError();
}

}  // namespace btk
