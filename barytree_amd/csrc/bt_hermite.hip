// bt_hermite.hip -- the Hermite approximation on gfx950, in FP64.
//
// A Hermite cluster carries 8 planes per grid point (value, the three first,
// the three mixed second and the mixed third derivative terms, in the order
// 0, x, y, z, xy, yz, xz, xyz); the planes of node c sit at q[8*P*c + m*P + j]
// and its grid coordinates at x[P*c + j].  Per dimension, a point at
// displacement d from node k has the factors
//   A_k = dj_k / d^2 + w_k / d,   B_k = dj_k / d
// (dj = 1/4 at the two ends, 1 inside; w from the node's box), and plane m of
// the tensor grid point (k1,k2,k3) weighs the point by the product of one
// factor per dimension: A for a dimension that plane m does not
// differentiate, B for one it does.  A point on a node coordinate (|d| <
// DBL_MIN) takes A = [k == hit], B = 0 in that dimension and leaves that
// dimension out of its denominator (src/clusters/clusters.c:1224-1836,
// src/interaction_compute/interaction_compute_downpass.c:860-1109).
#include <hip/hip_runtime.h>

#include <cfloat>

#include "bt_kernels.h"

namespace btk {
namespace {

constexpr int HTPB = 256;
constexpr int HCHUNK = 64;  // particles per LDS chunk of the upward pass
constexpr int HKC = 1, HKY = 2, HKR = 3, HKS = KERNEL_SIN_OVER_R;

__device__ __forceinline__ double herm_dj(int k, int n) { return (k == 0 || k == n) ? 0.25 : 1.0; }

// ---------------------------------------------------------------------
// Upward pass: block (part, slice) = particles [beg, end) of one node part
// against grid points slice*HTPB + tid; the 8 plane sums (8 more of w for
// subtraction) are stored to the part's row (partial rows are summed per
// node afterwards, in part order).
// ---------------------------------------------------------------------
template <int NL, bool SS>
__global__ __launch_bounds__(HTPB) void k_herm_anterp(HermAnterpArgs a)
{
    constexpr int n = NL - 1, P = NL * NL * NL;
    const int tid = threadIdx.x;
    const int part = blockIdx.x;
    const int node = a.part_node[part];
    const int beg = a.part_beg[part], end = a.part_end[part];
    const int j = blockIdx.y * HTPB + tid;
    const int jc = j < P ? j : 0;
    const int k1 = jc % NL, k2 = (jc / NL) % NL, k3 = jc / (NL * NL);

    __shared__ double Lnode[3][NL], Lw[3][NL];
    __shared__ double FA[3][NL][HCHUNK], FB[3][NL][HCHUNK];
    __shared__ double Lsum[3][HCHUNK], F1[HCHUNK], F2[HCHUNK];
    __shared__ int Lex[3][HCHUNK];
    if (tid < 3 * NL) {
        Lnode[tid / NL][tid % NL] = a.tab[(size_t)node * 3 * NL + tid];
        Lw[tid / NL][tid % NL] = a.wtab[(size_t)node * 3 * NL + tid];
    }

    double t[8], u[8];
#pragma unroll
    for (int m = 0; m < 8; m++) { t[m] = 0.0; u[m] = 0.0; }

    for (int base = beg; base < end; base += HCHUNK) {
        const int cnt = min(HCHUNK, end - base);
        __syncthreads();
        // factors of particle p in dimension d (thread d*HCHUNK + p)
        if (tid < 3 * HCHUNK) {
            const int d = tid / HCHUNK, p = tid % HCHUNK;
            if (p < cnt) {
                const double *pc = d == 0 ? a.px : (d == 1 ? a.py : a.pz);
                const double c = pc[base + p];
                int ex = -1;
                double sum = 0.0;
#pragma unroll
                for (int k = 0; k < NL; k++) {
                    const double dx = c - Lnode[d][k];
                    if (fabs(dx) < DBL_MIN) ex = k;
                    const double dj = herm_dj(k, n);
                    const double A = dj / (dx * dx) + Lw[d][k] / dx;
                    sum += A;
                    FA[d][k][p] = A;
                    FB[d][k][p] = dj / dx;
                }
                if (ex >= 0) {
#pragma unroll
                    for (int k = 0; k < NL; k++) { FA[d][k][p] = k == ex ? 1.0 : 0.0; FB[d][k][p] = 0.0; }
                }
                Lsum[d][p] = sum;
                Lex[d][p] = ex;
            } else {
#pragma unroll
                for (int k = 0; k < NL; k++) { FA[d][k][p] = 0.0; FB[d][k][p] = 0.0; }
            }
        }
        __syncthreads();
        // modified charges (clusters.c:1334-1370)
        if (tid < HCHUNK) {
            const int p = tid;
            double f1 = 0.0, f2 = 0.0;
            if (p < cnt) {
                double den = 1.0;
                if (Lex[0][p] == -1) den *= Lsum[0][p];
                if (Lex[1][p] == -1) den *= Lsum[1][p];
                if (Lex[2][p] == -1) den *= Lsum[2][p];
                const double q = a.pq[base + p];
                if constexpr (SS) {
                    const double w = a.pw[base + p];
                    f1 = q * w / den;
                    f2 = w / den;
                } else {
                    f1 = q / den;
                }
            }
            F1[p] = f1;
            F2[p] = f2;
        }
        __syncthreads();
        for (int p = 0; p < HCHUNK; p++) {
            const double ay = FA[1][k2][p], by = FB[1][k2][p], az = FA[2][k3][p], bz = FB[2][k3][p];
            const double yz00 = ay * az, yz10 = by * az, yz01 = ay * bz, yz11 = by * bz;
            const double xa = FA[0][k1][p] * F1[p], xb = FB[0][k1][p] * F1[p];
            t[0] = fma(xa, yz00, t[0]); t[1] = fma(xb, yz00, t[1]);
            t[2] = fma(xa, yz10, t[2]); t[3] = fma(xa, yz01, t[3]);
            t[4] = fma(xb, yz10, t[4]); t[5] = fma(xa, yz11, t[5]);
            t[6] = fma(xb, yz01, t[6]); t[7] = fma(xb, yz11, t[7]);
            if constexpr (SS) {
                const double wa = FA[0][k1][p] * F2[p], wb = FB[0][k1][p] * F2[p];
                u[0] = fma(wa, yz00, u[0]); u[1] = fma(wb, yz00, u[1]);
                u[2] = fma(wa, yz10, u[2]); u[3] = fma(wa, yz01, u[3]);
                u[4] = fma(wb, yz10, u[4]); u[5] = fma(wa, yz11, u[5]);
                u[6] = fma(wb, yz01, u[6]); u[7] = fma(wb, yz11, u[7]);
            }
        }
    }
    if (j < P) {
        double *q = a.Qp + (size_t)part * 8 * P;
#pragma unroll
        for (int m = 0; m < 8; m++) q[m * P + j] = t[m];
        if constexpr (SS) {
            double *w = a.Wp + (size_t)part * 8 * P;
#pragma unroll
            for (int m = 0; m < 8; m++) w[m * P + j] = u[m];
        }
    }
}

// ---------------------------------------------------------------------
// Particle-cluster far field: tile {key, first target, count} of up to HTPB
// targets (one per thread) against the 8-plane clusters of its key, added to
// the potential (coulomb_pc.c:66-135, yukawa_pc.c:61-140,
// regularized-coulomb_pc.c:67-139, coulomb_ss_pc.c:63-163,
// yukawa_ss_pc.c:61-167, regularized-coulomb_ss_pc.c:63-160).
// ---------------------------------------------------------------------
template <int KER, bool SS>
__device__ __forceinline__ double herm_pc_pair(double dx, double dy, double dz, const double *q, const double *w,
                                               double tq, const KParams &kp)
{
    if constexpr (KER == HKS) {
        // derivative expansion of sin(k r)/r; coincident points are skipped
        // (sin-over-r_pc.c:68-155)
        const double k = kp.p0, k2 = k * k, k3 = k * k2;
        const double r = sqrt(dx * dx + dy * dy + dz * dz);
        if (!(r > DBL_MIN)) return 0.0;
        const double rinv = 1.0 / r;
        const double r2inv = rinv * rinv, r3inv = rinv * r2inv, r4inv = r2inv * r2inv;
        const double r5inv = r3inv * r2inv, r6inv = r3inv * r3inv, r7inv = r4inv * r3inv;
        double sn, cs;
        sincos(k * r, &sn, &cs);
        return sn * rinv * q[0] + (sn * r3inv - k * cs * r2inv) * (q[1] * dx + q[2] * dy + q[3] * dz) +
               (sn * (3 * r5inv - k2 * r3inv) - 3 * k * cs * r4inv) * (q[4] * dx * dy + q[5] * dy * dz + q[6] * dx * dz) +
               (sn * (15 * r7inv - 6 * k2 * r5inv) + cs * (k3 * r4inv - 15 * k * r6inv)) * q[7] * dx * dy * dz;
    } else if constexpr (KER == HKY) {
        const double k = kp.p0, k2 = k * k, k3 = k * k2;
        const double r2 = dx * dx + dy * dy + dz * dz;
        const double rinv = rsqrt(r2), r = r2 * rinv, r3 = r2 * r;
        const double r3inv = rinv * rinv * rinv, r5inv = r3inv * rinv * rinv, r7inv = r5inv * rinv * rinv;
        double c[8];
#pragma unroll
        for (int m = 0; m < 8; m++) c[m] = SS ? q[m] - w[m] * tq : q[m];
        return exp(-k * r) *
               (rinv * c[0] + r3inv * (1 + k * r) * (c[1] * dx + c[2] * dy + c[3] * dz) +
                r5inv * (3 + 3 * k * r + k2 * r2) * (c[4] * dx * dy + c[5] * dy * dz + c[6] * dx * dz) +
                r7inv * (15 + 15 * k * r + 6 * k2 * r2 + k3 * r3) * c[7] * dx * dy * dz);
    } else {
        // Coulomb and regularized Coulomb: derivative expansion of 1/r (r with
        // eps^2 for the regularized kernel); subtraction adds the one of
        // exp(-r^2/k^2)/r weighted by -tq
        double r2 = dx * dx + dy * dy + dz * dz;
        if constexpr (KER == HKR) r2 += kp.eps2;
        const double rinv = rsqrt(r2);
        const double r3inv = rinv * rinv * rinv, r5inv = r3inv * rinv * rinv, r7inv = r5inv * rinv * rinv;
        double v = rinv * q[0] + r3inv * (q[1] * dx + q[2] * dy + q[3] * dz) +
                   3 * r5inv * (q[4] * dx * dy + q[5] * dy * dz + q[6] * dx * dz) + 15 * r7inv * q[7] * dx * dy * dz;
        if constexpr (SS) {
            const double rk2 = r2 * kp.inv_a2, rk4 = rk2 * rk2, rk6 = rk4 * rk2;
            v -= tq * exp(-rk2) *
                 (rinv * w[0] + r3inv * (1 + 2 * rk2) * (w[1] * dx + w[2] * dy + w[3] * dz) +
                  r5inv * (3 + 4 * rk2 + 4 * rk4) * (w[4] * dx * dy + w[5] * dy * dz + w[6] * dx * dz) +
                  r7inv * (15 + 18 * rk2 + 12 * rk4 + 8 * rk6) * w[7] * dx * dy * dz);
        }
        return v;
    }
}

template <int KER, bool SS>
__global__ __launch_bounds__(HTPB) void k_herm_pc(HermPCArgs a)
{
    const Tile tile = a.tiles[blockIdx.x];
    if (tile.z == 0) return;
    const int tid = threadIdx.x;
    const int P = a.P;
    __shared__ double Lx[HTPB], Ly[HTPB], Lz[HTPB];
    __shared__ double Lq[8][HTPB];
    __shared__ double Lw[SS ? 8 : 1][HTPB];
    const int i = tile.y + (tid < tile.z ? tid : 0);
    const double tx = a.tx[i], ty = a.ty[i], tz = a.tz[i];
    const double tq = SS ? a.tq[i] : 0.0;
    double acc = 0.0;
    for (int c = a.clu_off[tile.x]; c < a.clu_off[tile.x + 1]; c++) {
        const size_t node = (size_t)a.clu_id[c];
        const double *cq = a.cq + node * 8 * P;
        const double *cw = SS ? a.cw + node * 8 * P : nullptr;
        for (int off = 0; off < P; off += HTPB) {
            const int n = min(HTPB, P - off);
            __syncthreads();
            if (tid < n) {
                const size_t g = node * P + off + tid;
                Lx[tid] = a.cx[g]; Ly[tid] = a.cy[g]; Lz[tid] = a.cz[g];
#pragma unroll
                for (int m = 0; m < 8; m++) Lq[m][tid] = cq[m * P + off + tid];
                if constexpr (SS) {
#pragma unroll
                    for (int m = 0; m < 8; m++) Lw[m][tid] = cw[m * P + off + tid];
                }
            }
            __syncthreads();
            for (int jj = 0; jj < n; jj++) {
                double q[8], w[8];
#pragma unroll
                for (int m = 0; m < 8; m++) {
                    q[m] = Lq[m][jj];
                    w[m] = SS ? Lw[m][jj] : 0.0;
                }
                acc += herm_pc_pair<KER, SS>(tx - Lx[jj], ty - Ly[jj], tz - Lz[jj], q, w, tq, a.kp);
            }
        }
    }
    if (tid < tile.z) a.pot[i] += acc;
}

// ---------------------------------------------------------------------
// Cluster-particle far field: tile {key, first grid point, count, node} of
// up to HTPB grid points (one per thread) of a target node against the
// source ranges of its key; the 8 plane sums go to the key's output row
// (coulomb_cp.c:66-184, yukawa_cp.c:68-194, regularized-coulomb_cp.c:68-188,
// sin-over-r_cp.c:68-198).
// ---------------------------------------------------------------------
template <int KER>
__global__ __launch_bounds__(HTPB) void k_herm_cp(HermCPArgs a)
{
    const Tile tile = a.tiles[blockIdx.x];
    if (tile.z == 0) return;
    const int tid = threadIdx.x;
    const int nl = a.nl, P = nl * nl * nl;
    __shared__ double Lx[HTPB], Ly[HTPB], Lz[HTPB], Lq[HTPB];
    const int jl = tile.y + (tid < tile.z ? tid : 0);
    const int k1 = jl % nl, k2 = (jl / nl) % nl, k3 = jl / (nl * nl);
    const double *tab = a.tab + (size_t)tile.w * 3 * nl;
    const double cx = tab[k1], cy = tab[nl + k2], cz = tab[2 * nl + k3];
    double t[8];
#pragma unroll
    for (int m = 0; m < 8; m++) t[m] = 0.0;
    for (int s = a.seg_off[tile.x]; s < a.seg_off[tile.x + 1]; s++) {
        const Seg sg = a.segs[s];
        const int cnt = (int)((unsigned)sg.y & SEG_COUNT_MASK);
        for (int off = 0; off < cnt; off += HTPB) {
            const int n = min(HTPB, cnt - off);
            __syncthreads();
            if (tid < n) {
                const int g = sg.x + off + tid;
                Lx[tid] = a.sx[g]; Ly[tid] = a.sy[g]; Lz[tid] = a.sz[g]; Lq[tid] = a.sq[g];
            }
            __syncthreads();
            for (int jj = 0; jj < n; jj++) {
                const double dx = Lx[jj] - cx, dy = Ly[jj] - cy, dz = Lz[jj] - cz;
                if constexpr (KER == HKS) {
                    const double k = a.kp.p0, k2 = k * k, k3 = k * k2;
                    const double r = sqrt(dx * dx + dy * dy + dz * dz);
                    const double rinv = 1.0 / r;
                    const double r2inv = rinv * rinv, r3inv = rinv * r2inv, r4inv = r2inv * r2inv;
                    const double r5inv = r3inv * r2inv, r6inv = r3inv * r3inv, r7inv = r4inv * r3inv;
                    double sn, cs;
                    sincos(k * r, &sn, &cs);
                    sn *= Lq[jj]; cs *= Lq[jj];
                    const double d1 = sn * r3inv - k * cs * r2inv;
                    const double d2 = sn * (3 * r5inv - k2 * r3inv) - 3 * k * cs * r4inv;
                    const double d3 = sn * (15 * r7inv - 6 * k2 * r5inv) + cs * (k3 * r4inv - 15 * k * r6inv);
                    t[0] += sn * rinv;
                    t[1] += d1 * dx; t[2] += d1 * dy; t[3] += d1 * dz;
                    t[4] += d2 * dx * dy; t[5] += d2 * dy * dz; t[6] += d2 * dx * dz;
                    t[7] += d3 * dx * dy * dz;
                } else if constexpr (KER == HKY) {
                    const double k = a.kp.p0, k2 = k * k, k3 = k * k2;
                    const double r2 = dx * dx + dy * dy + dz * dz;
                    const double rinv = rsqrt(r2), r = r2 * rinv, r3 = r2 * r;
                    const double r2inv = rinv * rinv;
                    const double rinvq = Lq[jj] * rinv * exp(-k * r);
                    const double r3inv = rinvq * r2inv, r5inv = r3inv * r2inv, r7inv = r5inv * r2inv;
                    const double d1 = r3inv * (1 + k * r);
                    const double d2 = r5inv * (3 + 3 * k * r + k2 * r2);
                    const double d3 = r7inv * (15 + 15 * k * r + 6 * k2 * r2 + k3 * r3);
                    t[0] += rinvq;
                    t[1] += d1 * dx; t[2] += d1 * dy; t[3] += d1 * dz;
                    t[4] += d2 * dx * dy; t[5] += d2 * dy * dz; t[6] += d2 * dx * dz;
                    t[7] += d3 * dx * dy * dz;
                } else {
                    double r2 = dx * dx + dy * dy + dz * dz;
                    if constexpr (KER == HKR) r2 += a.kp.eps2;
                    const double rinv = rsqrt(r2);
                    const double r2inv = rinv * rinv;
                    const double rinvq = Lq[jj] * rinv;
                    const double r3inv = rinvq * r2inv, r5inv = 3.0 * (r3inv * r2inv), r7inv = r3inv * r2inv * r2inv;
                    t[0] += rinvq;
                    t[1] += r3inv * dx; t[2] += r3inv * dy; t[3] += r3inv * dz;
                    t[4] += r5inv * dx * dy; t[5] += r5inv * dy * dz; t[6] += r5inv * dx * dz;
                    t[7] += r7inv * dx * dy * dz * 15.0;
                }
            }
        }
    }
    if (tid < tile.z) {
        double *o = a.rows + a.out_row[tile.x];
#pragma unroll
        for (int m = 0; m < 8; m++) o[(size_t)m * P + jl] = t[m];
    }
}

// One node's 8-plane interpolant at the thread's target p (the node's
// Chebyshev nodes and Hermite weights in Lnode/Lw; planes q).
template <int NL>
__device__ double herm_down_node(const double p[3], const double (*Lnode)[NL], const double (*Lw)[NL],
                                 double (*Lq)[NL * NL], const double *q, int tid)
{
    constexpr int n = NL - 1, P = NL * NL * NL, S = NL * NL;
    // factors per dimension; exact hits (downpass.c:962-975)
    double A[3][NL], B[3][NL];
    double den = 1.0;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        int e = -1;
        double sum = 0.0;
#pragma unroll
        for (int k = 0; k < NL; k++) {
            const double dx = p[d] - Lnode[d][k];
            if (fabs(dx) < DBL_MIN) e = k;
            const double dj = herm_dj(k, n);
            A[d][k] = dj / (dx * dx) + Lw[d][k] / dx;
            B[d][k] = dj / dx;
            sum += A[d][k];
        }
        if (e == -1) {
            den /= sum;
        } else {
#pragma unroll
            for (int k = 0; k < NL; k++) { A[d][k] = k == e ? 1.0 : 0.0; B[d][k] = 0.0; }
        }
    }
    double temp = 0.0;
#pragma unroll 1
    for (int k3 = 0; k3 < NL; k3++) {
        __syncthreads();
        for (int g = tid; g < 8 * S; g += HTPB) Lq[g / S][g % S] = q[(size_t)(g / S) * P + k3 * S + g % S];
        __syncthreads();
        double az = 0.0, bz = 0.0;
#pragma unroll
        for (int k = 0; k < NL; k++)
            if (k == k3) { az = A[2][k]; bz = B[2][k]; }
#pragma unroll
        for (int k2 = 0; k2 < NL; k2++) {
            const double yz00 = A[1][k2] * az, yz10 = B[1][k2] * az, yz01 = A[1][k2] * bz, yz11 = B[1][k2] * bz;
#pragma unroll
            for (int k1 = 0; k1 < NL; k1++) {
                const int j = k2 * NL + k1;
                const double va = yz00 * Lq[0][j] + yz10 * Lq[2][j] + yz01 * Lq[3][j] + yz11 * Lq[5][j];
                const double vb = yz00 * Lq[1][j] + yz10 * Lq[4][j] + yz01 * Lq[6][j] + yz11 * Lq[7][j];
                temp = fma(A[0][k1], va, fma(B[0][k1], vb, temp));
            }
        }
    }
    return den * temp;
}

// ---------------------------------------------------------------------
// Downpass: tile {leaf, first target, count} of up to HTPB targets (one per
// thread).  Every node containing a target adds its 8-plane interpolant to
// it (interaction_compute_downpass.c:117-120, 860-1109); here a target
// walks from its leaf to the root, summing the interpolants of the nodes
// that received planes, and adds the sum once (no two blocks touch one
// target).  A node's planes stream through LDS one z-slab (8 x NL^2 values)
// at a time.
// ---------------------------------------------------------------------
template <int NL>
__global__ __launch_bounds__(HTPB) void k_herm_down(HermDownArgs a)
{
    constexpr int P = NL * NL * NL, S = NL * NL;
    const Tile tile = a.tiles[blockIdx.x];
    if (tile.z == 0) return;
    const int tid = threadIdx.x;
    __shared__ double Lnode[3][NL], Lw[3][NL];
    __shared__ double Lq[8][S];
    const int i = tile.y + (tid < tile.z ? tid : 0);
    const double p[3] = {a.tx[i], a.ty[i], a.tz[i]};
    double total = 0.0;
    for (int node = tile.x; node >= 0; node = a.parent[node]) {
        if (a.rows_off[node + 1] == a.rows_off[node]) continue;  // no planes (uniform in the block)
        __syncthreads();
        if (tid < 3 * NL) {
            Lnode[tid / NL][tid % NL] = a.tab[(size_t)node * 3 * NL + tid];
            Lw[tid / NL][tid % NL] = a.wtab[(size_t)node * 3 * NL + tid];
        }
        __syncthreads();
        total += herm_down_node<NL>(p, Lnode, Lw, Lq, a.Q + (size_t)node * 8 * P, tid);
    }
    if (tid < tile.z) a.pot[i] += total;
}


template <bool SS>
hipError_t herm_anterp_dispatch(const HermAnterpArgs &a, hipStream_t st)
{
    const int nl = a.degree + 1, P = nl * nl * nl;
    dim3 g(a.nparts, (P + HTPB - 1) / HTPB), b(HTPB);
    switch (a.degree) {
#define CASE(D) case D: k_herm_anterp<D + 1, SS><<<g, b, 0, st>>>(a); break;
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8) CASE(9) CASE(10) CASE(11) CASE(12)
#undef CASE
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace

hipError_t launch_herm_anterp(const HermAnterpArgs &a, hipStream_t st)
{
    if (a.nparts <= 0) return hipSuccess;
    return a.subtraction ? herm_anterp_dispatch<true>(a, st) : herm_anterp_dispatch<false>(a, st);
}

hipError_t launch_herm_pc(const HermPCArgs &a, hipStream_t st)
{
    if (a.ntiles <= 0) return hipSuccess;
    dim3 g(a.ntiles), b(HTPB);
    const bool ss = a.singularity == 2;
#define L(K, S) k_herm_pc<K, S><<<g, b, 0, st>>>(a)
    switch (a.kernel) {
        case HKC: if (ss) L(HKC, true); else L(HKC, false); break;
        case HKY: if (ss) L(HKY, true); else L(HKY, false); break;
        case HKR: if (ss) L(HKR, true); else L(HKR, false); break;
        case HKS: if (ss) return hipErrorInvalidValue; L(HKS, false); break;
        default: return hipErrorInvalidValue;
    }
#undef L
    return hipGetLastError();
}

hipError_t launch_herm_cp(const HermCPArgs &a, hipStream_t st)
{
    if (a.ntiles <= 0) return hipSuccess;
    dim3 g(a.ntiles), b(HTPB);
    switch (a.kernel) {
        case HKC: k_herm_cp<HKC><<<g, b, 0, st>>>(a); break;
        case HKY: k_herm_cp<HKY><<<g, b, 0, st>>>(a); break;
        case HKR: k_herm_cp<HKR><<<g, b, 0, st>>>(a); break;
        case HKS: k_herm_cp<HKS><<<g, b, 0, st>>>(a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_herm_down(const HermDownArgs &a, hipStream_t st)
{
    if (a.ntiles <= 0) return hipSuccess;
    dim3 g(a.ntiles), b(HTPB);
    switch (a.degree) {
#define CASE(D) case D: k_herm_down<D + 1><<<g, b, 0, st>>>(a); break;
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8) CASE(9) CASE(10) CASE(11) CASE(12)
#undef CASE
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace btk
