// bt_learn.hip -- gfx950 kernels of the learned degree selection:
//   k_label            : minimal degree of a far-field pair (FP64, one block per pair)
//   k_input, k_dense   : MLP forward, one layer per launch (FP32, LDS-tiled)
//   k_loss, k_dense_t  : loss, output gradient and backward through a layer
//   k_mlp_wgrad        : weight gradients D^T A as LDS-tiled products
//   k_decide           : degrees from the logits
//   k_adam             : Adam update
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>

#include "bt_learn.h"

namespace btl {

namespace {

// ---------------------------------------------------------------------
// Pair labels.
// ---------------------------------------------------------------------
constexpr int LTPB = 256;
constexpr int NLMAX = LABEL_DMAX + 1;
constexpr int PMAX = NLMAX * NLMAX * NLMAX;
constexpr int LPT = LTPB / LABEL_NT;  // lanes per sampled target (8)

// SK pair kernel: 1/r, exp(-kappa r)/r, 1/sqrt(r^2 + eps^2)
__device__ __forceinline__ double green(int kernel, double p0, double dx, double dy, double dz)
{
    const double r2 = dx * dx + dy * dy + dz * dz;
    if (kernel == 3) return 1.0 / sqrt(r2 + p0 * p0);
    const double r = sqrt(r2);
    if (kernel == 2) return exp(-p0 * r) / r;
    return 1.0 / r;
}

// Lagrange basis at Chebyshev points (barycentric form, second kind):
// l_k(x) = (w_k/(x-x_k)) / sum_j w_j/(x-x_j), w_k = (-1)^k d_k with
// d_0 = d_n = 1/2; an exact hit |x - x_k| < DBL_MIN gives the Kronecker row.
__device__ void lagrange_row(double x, const double *nodes, int nl, double *out)
{
    const int n = nl - 1;
    double sum = 0.0;
    int hit = -1;
    for (int k = 0; k < nl; k++) {
        const double c = x - nodes[k];
        if (fabs(c) < DBL_MIN) hit = k;
        const double d = (k == 0 || k == n) ? 0.5 : 1.0;
        const double v = ((k & 1) ? -d : d) / c;
        out[k] = v;
        sum += v;
    }
    for (int k = 0; k < nl; k++) out[k] = hit >= 0 ? (k == hit ? 1.0 : 0.0) : out[k] / sum;
}

// sum over the LPT lanes of a group (lanes tid & ~(LPT-1) .. +LPT-1), fixed order
__device__ __forceinline__ double group_sum(double v)
{
#pragma unroll
    for (int o = LPT / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, LPT);
    return v;
}

// Subtraction kernels: the second term's kernel, which multiplies the
// source weight and, at the target, minus the target's charge
// (coulomb_ss_pc.c:47, yukawa_ss_pc.c:46, regularized-coulomb_ss_pc.c,
// coulomb_ss_cp.c:54-55): Coulomb e^{-r^2/k^2}/r, Yukawa e^{-kappa r}/r (the
// same kernel), regularized Coulomb e^{-r^2/a^2}/sqrt(r^2 + eps^2);
// p1 = 1/k^2 or 1/a^2.
__device__ __forceinline__ double green_w(int kernel, double p0, double p1, double dx, double dy, double dz)
{
    if (kernel == 2) return green(kernel, p0, dx, dy, dz);
    const double r2 = dx * dx + dy * dy + dz * dz;
    return exp(-r2 * p1) * green(kernel, p0, dx, dy, dz);
}

__global__ __launch_bounds__(LTPB) void k_label(LabelArgs A)
{
    constexpr int NT = LABEL_NT, NS = LABEL_NS;
    __shared__ double Ltx[NT], Lty[NT], Ltz[NT], Ltq[NT], Lex[NT], Lred[2 * NT];
    __shared__ double Lsx[NS], Lsy[NS], Lsz[NS], Lsq[NS], Lsw[NS];
    __shared__ double Lnt[3 * NLMAX], Lns[3 * NLMAX];
    __shared__ double Bs[NS * 3 * NLMAX];
    __shared__ double Bt[NT * 3 * NLMAX];
    // grid values of the two terms (the second only with subtraction)
    __shared__ double G1[PMAX], G2[PMAX], G1w[PMAX], G2w[PMAX];
    __shared__ int Ldone;

    const bool ss = A.ss != 0;
    const int pi = blockIdx.x, tid = threadIdx.x;
    const int kind = A.kind[pi];
    const int tb0 = A.tbeg[pi], tc = A.tcnt[pi], sb0 = A.sbeg[pi], sc = A.scnt[pi];
    const int nt = min(tc, NT), ns = min(sc, NS);
    const double *tbox = A.tbox + 6 * (size_t)pi, *sbox = A.sbox + 6 * (size_t)pi;
    if (tid < nt) {
        const int i = tb0 + (int)(((long long)tid * tc) / nt);
        Ltx[tid] = A.tx[i]; Lty[tid] = A.ty[i]; Ltz[tid] = A.tz[i];
        Ltq[tid] = ss ? A.tq[i] : 0.0;
    }
    if (tid < ns) {
        const int s = sb0 + (int)(((long long)tid * sc) / ns);
        Lsx[tid] = A.sx[s]; Lsy[tid] = A.sy[s]; Lsz[tid] = A.sz[s];
        // subtraction: the first term's charge is q w, the second's w
        const double w = ss ? A.sw[s] : 1.0;
        Lsq[tid] = ss ? A.sq[s] * w : A.sq[s];
        Lsw[tid] = w;
    }
    if (tid == 0) Ldone = 0;
    __syncthreads();

    const int g = tid / LPT, l = tid % LPT;  // target of this lane group, lane in group
    // exact sums at the sampled targets
    {
        double v = 0.0, vw = 0.0;
        if (g < nt)
            for (int s = l; s < ns; s += LPT) {
                const double dx = Ltx[g] - Lsx[s], dy = Lty[g] - Lsy[s], dz = Ltz[g] - Lsz[s];
                v += Lsq[s] * green(A.kernel, A.p0, dx, dy, dz);
                if (ss) vw += Lsw[s] * green_w(A.kernel, A.p0, A.p1, dx, dy, dz);
            }
        v = group_sum(v);
        if (ss) vw = group_sum(vw);
        if (l == 0 && g < nt) Lex[g] = ss ? v - Ltq[g] * vw : v;
    }
    __syncthreads();

    int label = LABEL_DMAX + 1;
    for (int p = LABEL_DMIN; p <= LABEL_DMAX; p++) {
        const int nl = p + 1, P = nl * nl * nl;
        if (tid < 3 * nl) {
            const int d = tid / nl, k = tid % nl;
            const double u = (cos(k * M_PI / p) + 1.0) * 0.5;
            Lnt[d * nl + k] = tbox[2 * d] + u * (tbox[2 * d + 1] - tbox[2 * d]);
            Lns[d * nl + k] = sbox[2 * d] + u * (sbox[2 * d + 1] - sbox[2 * d]);
        }
        __syncthreads();
        if (kind != 2) {  // PC, CC: source-side basis
            for (int idx = tid; idx < ns * 3; idx += LTPB) {
                const int s = idx / 3, d = idx % 3;
                const double c = d == 0 ? Lsx[s] : (d == 1 ? Lsy[s] : Lsz[s]);
                lagrange_row(c, Lns + d * nl, nl, Bs + (s * 3 + d) * NLMAX);
            }
        }
        if (kind != 1) {  // CP, CC: target-side basis
            for (int idx = tid; idx < nt * 3; idx += LTPB) {
                const int i = idx / 3, d = idx % 3;
                const double c = d == 0 ? Ltx[i] : (d == 1 ? Lty[i] : Ltz[i]);
                lagrange_row(c, Lnt + d * nl, nl, Bt + (i * 3 + d) * NLMAX);
            }
        }
        __syncthreads();
        if (kind != 2) {
            // modified charges of the source grid: Q_k = sum_s q_s l_k(y_s)
            for (int j = tid; j < P; j += LTPB) {
                const int k1 = j % nl, k2 = (j / nl) % nl, k3 = j / (nl * nl);
                double q = 0.0, qw = 0.0;
                for (int s = 0; s < ns; s++) {
                    const double b = Bs[(s * 3) * NLMAX + k1] * Bs[(s * 3 + 1) * NLMAX + k2] *
                                     Bs[(s * 3 + 2) * NLMAX + k3];
                    q += Lsq[s] * b;
                    if (ss) qw += Lsw[s] * b;
                }
                G1[j] = q;
                G1w[j] = qw;
            }
            __syncthreads();
        }
        if (kind == 2) {
            // target grid values from the sources
            for (int j = tid; j < P; j += LTPB) {
                const int k1 = j % nl, k2 = (j / nl) % nl, k3 = j / (nl * nl);
                const double gx = Lnt[k1], gy = Lnt[nl + k2], gz = Lnt[2 * nl + k3];
                double v = 0.0, vw = 0.0;
                for (int s = 0; s < ns; s++) {
                    const double dx = gx - Lsx[s], dy = gy - Lsy[s], dz = gz - Lsz[s];
                    v += Lsq[s] * green(A.kernel, A.p0, dx, dy, dz);
                    if (ss) vw += Lsw[s] * green_w(A.kernel, A.p0, A.p1, dx, dy, dz);
                }
                G2[j] = v;
                G2w[j] = vw;
            }
            __syncthreads();
        } else if (kind == 3) {
            // target grid values from the source grid
            for (int j = tid; j < P; j += LTPB) {
                const int k1 = j % nl, k2 = (j / nl) % nl, k3 = j / (nl * nl);
                const double gx = Lnt[k1], gy = Lnt[nl + k2], gz = Lnt[2 * nl + k3];
                double v = 0.0, vw = 0.0;
                for (int k = 0; k < P; k++) {
                    const int m1 = k % nl, m2 = (k / nl) % nl, m3 = k / (nl * nl);
                    const double dx = gx - Lns[m1], dy = gy - Lns[nl + m2], dz = gz - Lns[2 * nl + m3];
                    v += G1[k] * green(A.kernel, A.p0, dx, dy, dz);
                    if (ss) vw += G1w[k] * green_w(A.kernel, A.p0, A.p1, dx, dy, dz);
                }
                G2[j] = v;
                G2w[j] = vw;
            }
            __syncthreads();
        }
        // approximation at the sampled targets
        double v = 0.0, vw = 0.0;
        if (g < nt) {
            if (kind == 1) {
                for (int k = l; k < P; k += LPT) {
                    const int m1 = k % nl, m2 = (k / nl) % nl, m3 = k / (nl * nl);
                    const double dx = Ltx[g] - Lns[m1], dy = Lty[g] - Lns[nl + m2], dz = Ltz[g] - Lns[2 * nl + m3];
                    v += G1[k] * green(A.kernel, A.p0, dx, dy, dz);
                    if (ss) vw += G1w[k] * green_w(A.kernel, A.p0, A.p1, dx, dy, dz);
                }
            } else {
                const double *b = Bt + (g * 3) * NLMAX;
                for (int j = l; j < P; j += LPT) {
                    const int k1 = j % nl, k2 = (j / nl) % nl, k3 = j / (nl * nl);
                    const double bb = b[k1] * b[NLMAX + k2] * b[2 * NLMAX + k3];
                    v += bb * G2[j];
                    if (ss) vw += bb * G2w[j];
                }
            }
        }
        v = group_sum(v);
        if (ss) vw = group_sum(vw);
        if (l == 0 && g < NT) {
            const double e = g < nt ? (ss ? v - Ltq[g] * vw : v) - Lex[g] : 0.0;
            Lred[g] = e * e;
            Lred[NT + g] = g < nt ? Lex[g] * Lex[g] : 0.0;
        }
        __syncthreads();
        if (tid == 0) {
            double num = 0.0, den = 0.0;
            for (int i = 0; i < NT; i++) { num += Lred[i]; den += Lred[NT + i]; }
            const double err = den > 0.0 ? sqrt(num / den) : (num > 0.0 ? INFINITY : 0.0);
            A.errors[(size_t)pi * LABEL_ND + (p - LABEL_DMIN)] = err;
            if (err <= A.tol) Ldone = 1;
        }
        __syncthreads();
        if (Ldone) { label = p; break; }
    }
    if (tid == 0) {
        A.labels[pi] = label;
        for (int p = label + 1; p <= LABEL_DMAX; p++)
            A.errors[(size_t)pi * LABEL_ND + (p - LABEL_DMIN)] = NAN;
    }
}

// ---------------------------------------------------------------------
// MLP.  Activations are feature-major: H[f][b] for feature f of sample b of
// a batch of nb samples.  A layer is a product with the (<= 64 x 64) weight
// matrix held in LDS: a block covers all outputs for 64 samples; thread t
// takes sample column t % 64 and 16 consecutive output rows, so every lane
// of a wave reads the same weight (an LDS broadcast) and its own sample.
// ---------------------------------------------------------------------
constexpr int MTPB = 256;
constexpr int CS = 64;  // samples per block
constexpr int RPT = 16; // output rows per thread

__device__ __forceinline__ float sigmoidf_(float z) { return 1.0f / (1.0f + expf(-z)); }

// Hout[i][b] = act(bias[i] + sum_j W[i][j] Hin[j][b]), i < fout, j < fin
__global__ __launch_bounds__(MTPB) void k_dense(const float *W, const float *bias, const float *Hin, int fin,
                                               int fout, int nb, float *Hout, int relu)
{
    __shared__ float Lw[MLP_H * MLP_H + MLP_H];
    __shared__ float Lh[MLP_H][CS + 1];
    const int tid = threadIdx.x, c = tid % CS, r0 = (tid / CS) * RPT;
    const int b0 = blockIdx.x * CS;
    for (int e = tid; e < fout * fin; e += MTPB) Lw[e] = W[e];
    for (int e = tid; e < fout; e += MTPB) Lw[MLP_H * MLP_H + e] = bias[e];
    for (int e = tid; e < fin * CS; e += MTPB) {
        const int j = e / CS, cc = e % CS;
        Lh[j][cc] = b0 + cc < nb ? Hin[(size_t)j * nb + b0 + cc] : 0.0f;
    }
    __syncthreads();
    if (r0 >= fout) return;
    float acc[RPT];
#pragma unroll
    for (int r = 0; r < RPT; r++) acc[r] = r0 + r < fout ? Lw[MLP_H * MLP_H + r0 + r] : 0.0f;
    for (int j = 0; j < fin; j++) {
        const float h = Lh[j][c];
#pragma unroll
        for (int r = 0; r < RPT; r++)
            if (r0 + r < fout) acc[r] = fmaf(Lw[(r0 + r) * fin + j], h, acc[r]);
    }
    if (b0 + c < nb) {
#pragma unroll
        for (int r = 0; r < RPT; r++)
            if (r0 + r < fout) Hout[(size_t)(r0 + r) * nb + b0 + c] = relu ? fmaxf(acc[r], 0.0f) : acc[r];
    }
}

// Backward through a layer: G[j][b] = [A[j][b] > 0] * sum_i W[i][j] D[i][b]
// (W is fout x fin; j < fin, i < fout)
__global__ __launch_bounds__(MTPB) void k_dense_t(const float *W, const float *D, const float *A, int fin,
                                                 int fout, int nb, float *G)
{
    __shared__ float Lw[MLP_H * MLP_H];
    __shared__ float Ld[MLP_H][CS + 1];
    const int tid = threadIdx.x, c = tid % CS, r0 = (tid / CS) * RPT;
    const int b0 = blockIdx.x * CS;
    for (int e = tid; e < fout * fin; e += MTPB) Lw[e] = W[e];
    for (int e = tid; e < fout * CS; e += MTPB) {
        const int i = e / CS, cc = e % CS;
        Ld[i][cc] = b0 + cc < nb ? D[(size_t)i * nb + b0 + cc] : 0.0f;
    }
    __syncthreads();
    if (r0 >= fin) return;
    float acc[RPT];
#pragma unroll
    for (int r = 0; r < RPT; r++) acc[r] = 0.0f;
    for (int i = 0; i < fout; i++) {
        const float d = Ld[i][c];
#pragma unroll
        for (int r = 0; r < RPT; r++)
            if (r0 + r < fin) acc[r] = fmaf(Lw[i * fin + r0 + r], d, acc[r]);
    }
    if (b0 + c < nb) {
#pragma unroll
        for (int r = 0; r < RPT; r++) {
            if (r0 + r >= fin) continue;
            const size_t o = (size_t)(r0 + r) * nb + b0 + c;
            G[o] = A[o] > 0.0f ? acc[r] : 0.0f;
        }
    }
}

// Normalised inputs: A0[f][b] = (x[s][f] - mean[f]) * scale[f], s = idx[b] (or b)
__global__ void k_input(const float *x, const int *idx, const float *mean, const float *scale, int nb,
                        float *A0)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const size_t s = idx ? (size_t)idx[b] : (size_t)b;
#pragma unroll
    for (int f = 0; f < MLP_IN; f++) A0[(size_t)f * nb + b] = (x[s * MLP_IN + f] - mean[f]) * scale[f];
}

// Loss and output gradient, in place: Z[k][b] (logits) -> dZ[k][b]
__global__ void k_loss(const int *labels, const int *idx, int nb, float *Z, float *loss)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const int lab = labels[idx ? idx[b] : b];
    const float inv_b = 1.0f / (float)nb;
    float l = 0.0f;
#pragma unroll
    for (int k = 0; k < MLP_OUT; k++) {
        const size_t o = (size_t)k * nb + b;
        const float z = Z[o];
        const float y = lab > LABEL_DMIN + k ? 1.0f : 0.0f;
        l += fmaxf(z, 0.0f) + log1pf(expf(-fabsf(z))) - y * z;
        Z[o] = (sigmoidf_(z) - y) * inv_b;
    }
    loss[b] = l;
}

// Degrees (and logits in sample-major order) from Z[k][b]
__global__ void k_decide(const float *Z, int nb, float delta, float *logits, int *degree)
{
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    int deg = LABEL_DMIN;
#pragma unroll
    for (int k = 0; k < MLP_OUT; k++) {
        const float z = Z[(size_t)k * nb + b];
        if (logits) logits[(size_t)b * MLP_OUT + k] = z;
        if (sigmoidf_(z) >= delta) deg = LABEL_DMIN + 1 + k;
    }
    if (degree) degree[b] = deg;
}

hipError_t forward_pass(const float *prm, const float *mean, const float *scale, const float *x,
                        const int *idx, int nb, const TrainBufs &tb, hipStream_t st)
{
    const int g1 = (nb + 255) / 256, gc = (nb + CS - 1) / CS;
    k_input<<<g1, 256, 0, st>>>(x, idx, mean, scale, nb, tb.a0);
    k_dense<<<gc, MTPB, 0, st>>>(prm + OFF_W1, prm + OFF_B1, tb.a0, MLP_IN, MLP_H, nb, tb.a1, 1);
    k_dense<<<gc, MTPB, 0, st>>>(prm + OFF_W2, prm + OFF_B2, tb.a1, MLP_H, MLP_H, nb, tb.a2, 1);
    k_dense<<<gc, MTPB, 0, st>>>(prm + OFF_W3, prm + OFF_B3, tb.a2, MLP_H, MLP_H, nb, tb.a3, 1);
    k_dense<<<gc, MTPB, 0, st>>>(prm + OFF_W4, prm + OFF_B4, tb.a3, MLP_H, MLP_OUT, nb, tb.dz, 0);
    return hipGetLastError();
}

// out[i][j] = sum_b D[i][b] A[j][b] (j < J), bias[i] = sum_b D[i][b]
// (column J); a 16x16 output tile per block, batch chunks of 64 in LDS.
constexpr int WT = 16, WC = 64;
__global__ __launch_bounds__(256) void k_mlp_wgrad(const float *D, const float *A, int I, int J, int nb,
                                                   float *outW, float *outB)
{
    __shared__ float Ld[WT][WC + 1], La[WT][WC + 1];
    const int tid = threadIdx.x, ti = tid / WT, tj = tid % WT;
    const int i0 = blockIdx.x * WT, j0 = blockIdx.y * WT;
    float acc = 0.0f;
    for (int c0 = 0; c0 < nb; c0 += WC) {
        for (int e = tid; e < WT * WC; e += 256) {
            const int r = e / WC, c = e % WC, b = c0 + c;
            const int i = i0 + r, jj = j0 + r;
            Ld[r][c] = (i < I && b < nb) ? D[(size_t)i * nb + b] : 0.0f;
            La[r][c] = b < nb ? (jj < J ? A[(size_t)jj * nb + b] : (jj == J ? 1.0f : 0.0f)) : 0.0f;
        }
        __syncthreads();
#pragma unroll 8
        for (int c = 0; c < WC; c++) acc = fmaf(Ld[ti][c], La[tj][c], acc);
        __syncthreads();
    }
    const int i = i0 + ti, jj = j0 + tj;
    if (i < I) {
        if (jj < J) outW[(size_t)i * J + jj] = acc;
        else if (jj == J) outB[i] = acc;
    }
}

__global__ void k_adam(float *p, const float *g, float *m, float *v, int n, float lr, float b1, float b2,
                       float eps, float c1, float c2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

}  // namespace

hipError_t launch_label(const LabelArgs &a, hipStream_t st)
{
    if (a.npairs <= 0) return hipSuccess;
    k_label<<<a.npairs, LTPB, 0, st>>>(a);
    return hipGetLastError();
}

hipError_t launch_mlp_infer(const float *params, const float *mean, const float *scale, int n,
                            const float *x, float delta, float *logits, int *degree, const TrainBufs &tb,
                            hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipError_t e = forward_pass(params, mean, scale, x, nullptr, n, tb, st);
    if (e != hipSuccess) return e;
    k_decide<<<(n + 255) / 256, 256, 0, st>>>(tb.dz, n, delta, logits, degree);
    return hipGetLastError();
}

hipError_t launch_mlp_fwdbwd(const float *params, const float *mean, const float *scale,
                             const float *x, const int *labels, const int *idx, int b,
                             const TrainBufs &tb, hipStream_t st)
{
    if (b <= 0) return hipSuccess;
    hipError_t e = forward_pass(params, mean, scale, x, idx, b, tb, st);
    if (e != hipSuccess) return e;
    const int g1 = (b + 255) / 256, gc = (b + CS - 1) / CS;
    k_loss<<<g1, 256, 0, st>>>(labels, idx, b, tb.dz, tb.loss);
    k_dense_t<<<gc, MTPB, 0, st>>>(params + OFF_W4, tb.dz, tb.a3, MLP_H, MLP_OUT, b, tb.d3);
    k_dense_t<<<gc, MTPB, 0, st>>>(params + OFF_W3, tb.d3, tb.a2, MLP_H, MLP_H, b, tb.d2);
    k_dense_t<<<gc, MTPB, 0, st>>>(params + OFF_W2, tb.d2, tb.a1, MLP_H, MLP_H, b, tb.d1);
    return hipGetLastError();
}

hipError_t launch_mlp_wgrad(const TrainBufs &tb, int b, float *grad, hipStream_t st)
{
    auto one = [&](const float *D, const float *A, int I, int J, int offW, int offB) {
        dim3 g((I + WT - 1) / WT, (J + 1 + WT - 1) / WT);
        k_mlp_wgrad<<<g, 256, 0, st>>>(D, A, I, J, b, grad + offW, grad + offB);
    };
    one(tb.d1, tb.a0, MLP_H, MLP_IN, OFF_W1, OFF_B1);
    one(tb.d2, tb.a1, MLP_H, MLP_H, OFF_W2, OFF_B2);
    one(tb.d3, tb.a2, MLP_H, MLP_H, OFF_W3, OFF_B3);
    one(tb.dz, tb.a3, MLP_OUT, MLP_H, OFF_W4, OFF_B4);
    return hipGetLastError();
}

hipError_t launch_adam(float *params, const float *grad, float *m, float *v, int n, float lr,
                       float beta1, float beta2, float eps, int t, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    const float c1 = 1.0f / (1.0f - powf(beta1, (float)t));
    const float c2 = 1.0f / (1.0f - powf(beta2, (float)t));
    k_adam<<<(n + 255) / 256, 256, 0, st>>>(params, grad, m, v, n, lr, beta1, beta2, eps, c1, c2);
    return hipGetLastError();
}

}  // namespace btl
