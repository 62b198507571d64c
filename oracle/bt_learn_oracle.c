/*
 * bt_learn_oracle.c -- CPU restatement of the learned-degree parts.
 *
 * TEST INFRASTRUCTURE ONLY, like bt_oracle.c: the checker for the HIP
 * kernels of barytree_amd/csrc/bt_learn.hip.  Only tests/ may load it.
 *
 * These parts have no counterpart in the reference (the degree model is an
 * extension); the definitions restated here are:
 *   - the pair label: least degree p in [2, 12] whose pair error is <= tol
 *     (13 if none), with the pair error the relative L2 error, over up to 32
 *     stride-sampled targets, of the degree-p approximation of the
 *     potential of up to 128 stride-sampled sources.  The approximations are
 *     the treecode's own: source-side interpolation (PC), target-side
 *     interpolation (CP) or both (CC) on Chebyshev points of the second kind
 *     over the node boxes, nodes x0 + (cos(k pi/p)+1)/2 (x1-x0)
 *     (src/clusters/clusters.c:436-440), barycentric Lagrange basis with
 *     weights (-1)^k d_k, d_0 = d_p = 1/2, and the exact-hit Kronecker row
 *     when |x - x_k| < DBL_MIN (clusters.c:446-520); with subtraction the
 *     potential is that of the subtraction kernel (both terms: the charges
 *     q w with the kernel, the weights w with its Gaussian-weighted form
 *     times minus the target's charge, each interpolated as the treecode
 *     does);
 *   - the degree model's loss and gradient: MLP 8 -> 64 -> 64 -> 64 -> 10
 *     with ReLU, batch-mean binary cross-entropy of the 10 "degree above
 *     2+k" outputs, in FP64.
 * Parity pins (tests/test_learn_oracle.py): finite differences of the loss,
 * closed forms of the zero-weight network, the tensor interpolation against
 * scipy's BarycentricInterpolator, convergence of the pair error with p.
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "bt_learn_oracle.h"

#define NT 32
#define NS 128
#define DMIN 2
#define DMAX 12
#define ND (DMAX - DMIN + 1)
#define NLMAX (DMAX + 1)

static double green(int kernel, double p0, double dx, double dy, double dz)
{
    double r2 = dx * dx + dy * dy + dz * dz;
    if (kernel == 3) return 1.0 / sqrt(r2 + p0 * p0);
    double r = sqrt(r2);
    if (kernel == 2) return exp(-p0 * r) / r;
    return 1.0 / r;
}

static void lagrange_row(double x, const double *nodes, int nl, double *out)
{
    int n = nl - 1, hit = -1;
    double sum = 0.0;
    for (int k = 0; k < nl; k++) {
        double c = x - nodes[k];
        if (fabs(c) < DBL_MIN) hit = k;
        double d = (k == 0 || k == n) ? 0.5 : 1.0;
        out[k] = ((k & 1) ? -d : d) / c;
        sum += out[k];
    }
    for (int k = 0; k < nl; k++) out[k] = hit >= 0 ? (k == hit ? 1.0 : 0.0) : out[k] / sum;
}

static void cheb_nodes(const double *box, int p, double *nodes /* 3*(p+1) */)
{
    int nl = p + 1;
    for (int d = 0; d < 3; d++)
        for (int k = 0; k < nl; k++) {
            double u = (cos(k * M_PI / p) + 1.0) * 0.5;
            nodes[d * nl + k] = box[2 * d] + u * (box[2 * d + 1] - box[2 * d]);
        }
}

/* Subtraction kernels' second term (the weight's kernel, multiplied at the
 * target by minus the target's charge): Coulomb e^{-r^2/k^2}/r
 * (coulomb_ss_pc.c:47), Yukawa e^{-kappa r}/r (yukawa_ss_pc.c:46: the same
 * kernel), regularized Coulomb e^{-r^2/a^2}/sqrt(r^2+eps^2); p1 = 1/k^2 or
 * 1/a^2. */
static double green_w(int kernel, double p0, double p1, double dx, double dy, double dz)
{
    if (kernel == 2) return green(kernel, p0, dx, dy, dz);
    double r2 = dx * dx + dy * dy + dz * dz;
    return exp(-r2 * p1) * green(kernel, p0, dx, dy, dz);
}

/* The pair potential at target t: sum_s c0_s g(t - s) - ct_t sum_s c1_s g_w(t - s)
 * with, for subtraction, c0 = q w, c1 = w and ct = the target's charge (the
 * second sum absent without subtraction); the approximations interpolate
 * each sum as the treecode does. */
int orc_label_pairs_ss(int npairs, const int *kind, const int *tbeg, const int *tcnt, const int *sbeg,
                       const int *scnt, const double *tbox, const double *sbox, const double *tx,
                       const double *ty, const double *tz, const double *tq, const double *sx,
                       const double *sy, const double *sz, const double *sq, const double *sw, int kernel,
                       double p0, int ss, double p1, double tol, int *labels, double *errors)
{
    size_t pm = (size_t)NLMAX * NLMAX * NLMAX;
    double *G1 = malloc(sizeof(double) * pm), *G2 = malloc(sizeof(double) * pm);
    double *G1w = malloc(sizeof(double) * pm), *G2w = malloc(sizeof(double) * pm);
    if (!G1 || !G2 || !G1w || !G2w) { free(G1); free(G2); free(G1w); free(G2w); return -1; }
    if (ss && (!tq || !sw)) { free(G1); free(G2); free(G1w); free(G2w); return -1; }
    for (int pi = 0; pi < npairs; pi++) {
        int nt = tcnt[pi] < NT ? tcnt[pi] : NT, ns = scnt[pi] < NS ? scnt[pi] : NS;
        double X[NT][3], CT[NT], Y[NS][3], Q[NS], W[NS], ex[NT];
        for (int k = 0; k < nt; k++) {
            int i = tbeg[pi] + (int)(((long long)k * tcnt[pi]) / nt);
            X[k][0] = tx[i]; X[k][1] = ty[i]; X[k][2] = tz[i];
            CT[k] = ss ? tq[i] : 0.0;
        }
        for (int k = 0; k < ns; k++) {
            int s = sbeg[pi] + (int)(((long long)k * scnt[pi]) / ns);
            Y[k][0] = sx[s]; Y[k][1] = sy[s]; Y[k][2] = sz[s];
            W[k] = ss ? sw[s] : 1.0;
            Q[k] = ss ? sq[s] * W[k] : sq[s];
        }
        for (int i = 0; i < nt; i++) {
            double v = 0.0, vw = 0.0;
            for (int s = 0; s < ns; s++) {
                double dx = X[i][0] - Y[s][0], dy = X[i][1] - Y[s][1], dz = X[i][2] - Y[s][2];
                v += Q[s] * green(kernel, p0, dx, dy, dz);
                if (ss) vw += W[s] * green_w(kernel, p0, p1, dx, dy, dz);
            }
            ex[i] = ss ? v - CT[i] * vw : v;
        }
        int label = DMAX + 1;
        for (int p = DMIN; p <= DMAX; p++) errors[(size_t)pi * ND + p - DMIN] = NAN;
        for (int p = DMIN; p <= DMAX; p++) {
            int nl = p + 1, P = nl * nl * nl;
            double nt_nodes[3 * NLMAX], ns_nodes[3 * NLMAX];
            cheb_nodes(tbox + 6 * (size_t)pi, p, nt_nodes);
            cheb_nodes(sbox + 6 * (size_t)pi, p, ns_nodes);
            if (kind[pi] != 2) {  /* source grid charges of each term */
                double B[NS][3][NLMAX];
                for (int s = 0; s < ns; s++)
                    for (int d = 0; d < 3; d++) lagrange_row(Y[s][d], ns_nodes + d * nl, nl, B[s][d]);
                for (int j = 0; j < P; j++) {
                    int k1 = j % nl, k2 = (j / nl) % nl, k3 = j / (nl * nl);
                    double q = 0.0, qw = 0.0;
                    for (int s = 0; s < ns; s++) {
                        double b = B[s][0][k1] * B[s][1][k2] * B[s][2][k3];
                        q += Q[s] * b;
                        qw += W[s] * b;
                    }
                    G1[j] = q;
                    G1w[j] = qw;
                }
            }
            if (kind[pi] == 2) {
                for (int j = 0; j < P; j++) {
                    int k1 = j % nl, k2 = (j / nl) % nl, k3 = j / (nl * nl);
                    double v = 0.0, vw = 0.0;
                    for (int s = 0; s < ns; s++) {
                        double dx = nt_nodes[k1] - Y[s][0], dy = nt_nodes[nl + k2] - Y[s][1],
                               dz = nt_nodes[2 * nl + k3] - Y[s][2];
                        v += Q[s] * green(kernel, p0, dx, dy, dz);
                        if (ss) vw += W[s] * green_w(kernel, p0, p1, dx, dy, dz);
                    }
                    G2[j] = v;
                    G2w[j] = vw;
                }
            } else if (kind[pi] == 3) {
                for (int j = 0; j < P; j++) {
                    int k1 = j % nl, k2 = (j / nl) % nl, k3 = j / (nl * nl);
                    double v = 0.0, vw = 0.0;
                    for (int k = 0; k < P; k++) {
                        int m1 = k % nl, m2 = (k / nl) % nl, m3 = k / (nl * nl);
                        double dx = nt_nodes[k1] - ns_nodes[m1], dy = nt_nodes[nl + k2] - ns_nodes[nl + m2],
                               dz = nt_nodes[2 * nl + k3] - ns_nodes[2 * nl + m3];
                        v += G1[k] * green(kernel, p0, dx, dy, dz);
                        if (ss) vw += G1w[k] * green_w(kernel, p0, p1, dx, dy, dz);
                    }
                    G2[j] = v;
                    G2w[j] = vw;
                }
            }
            double num = 0.0, den = 0.0;
            for (int i = 0; i < nt; i++) {
                double v = 0.0, vw = 0.0;
                if (kind[pi] == 1) {
                    for (int k = 0; k < P; k++) {
                        int m1 = k % nl, m2 = (k / nl) % nl, m3 = k / (nl * nl);
                        double dx = X[i][0] - ns_nodes[m1], dy = X[i][1] - ns_nodes[nl + m2],
                               dz = X[i][2] - ns_nodes[2 * nl + m3];
                        v += G1[k] * green(kernel, p0, dx, dy, dz);
                        if (ss) vw += G1w[k] * green_w(kernel, p0, p1, dx, dy, dz);
                    }
                } else {
                    double b[3][NLMAX];
                    for (int d = 0; d < 3; d++) lagrange_row(X[i][d], nt_nodes + d * nl, nl, b[d]);
                    for (int j = 0; j < P; j++) {
                        int k1 = j % nl, k2 = (j / nl) % nl, k3 = j / (nl * nl);
                        double bb = b[0][k1] * b[1][k2] * b[2][k3];
                        v += bb * G2[j];
                        vw += bb * G2w[j];
                    }
                }
                double a = ss ? v - CT[i] * vw : v;
                num += (a - ex[i]) * (a - ex[i]);
                den += ex[i] * ex[i];
            }
            double err = den > 0.0 ? sqrt(num / den) : (num > 0.0 ? INFINITY : 0.0);
            errors[(size_t)pi * ND + p - DMIN] = err;
            if (err <= tol) { label = p; break; }
        }
        labels[pi] = label;
    }
    free(G1); free(G2); free(G1w); free(G2w);
    return 0;
}

int orc_label_pairs(int npairs, const int *kind, const int *tbeg, const int *tcnt, const int *sbeg,
                    const int *scnt, const double *tbox, const double *sbox, const double *tx,
                    const double *ty, const double *tz, const double *sx, const double *sy,
                    const double *sz, const double *sq, int kernel, double p0, double tol,
                    int *labels, double *errors)
{
    return orc_label_pairs_ss(npairs, kind, tbeg, tcnt, sbeg, scnt, tbox, sbox, tx, ty, tz, NULL, sx, sy, sz,
                              sq, NULL, kernel, p0, 0, 0.0, tol, labels, errors);
}

/* ---- degree model -------------------------------------------------------- */

#define IN 8
#define H 64
#define OUT 10
#define OW1 0
#define OB1 (OW1 + H * IN)
#define OW2 (OB1 + H)
#define OB2 (OW2 + H * H)
#define OW3 (OB2 + H)
#define OB3 (OW3 + H * H)
#define OW4 (OB3 + H)
#define OB4 (OW4 + OUT * H)
#define NPARAM (OB4 + OUT)

int orc_mlp_num_params(void) { return NPARAM; }

static void layer(const double *W, const double *b, const double *in, int nin, int nout, int relu,
                  double *out)
{
    for (int i = 0; i < nout; i++) {
        double a = b[i];
        for (int j = 0; j < nin; j++) a += W[i * nin + j] * in[j];
        out[i] = relu ? (a > 0.0 ? a : 0.0) : a;
    }
}

void orc_mlp_logits(const double *prm, const double *mean, const double *scale, int n,
                    const float *x, double *logits)
{
    for (int b = 0; b < n; b++) {
        double h0[IN], h1[H], h2[H], h3[H];
        for (int f = 0; f < IN; f++) h0[f] = ((double)x[(size_t)b * IN + f] - mean[f]) * scale[f];
        layer(prm + OW1, prm + OB1, h0, IN, H, 1, h1);
        layer(prm + OW2, prm + OB2, h1, H, H, 1, h2);
        layer(prm + OW3, prm + OB3, h2, H, H, 1, h3);
        layer(prm + OW4, prm + OB4, h3, H, OUT, 0, logits + (size_t)b * OUT);
    }
}

double orc_mlp_loss_grad(const double *prm, const double *mean, const double *scale, int n,
                         const float *x, const int *labels, double *grad)
{
    if (grad) memset(grad, 0, sizeof(double) * NPARAM);
    double loss = 0.0;
    for (int b = 0; b < n; b++) {
        double h0[IN], h1[H], h2[H], h3[H], z[OUT], dz[OUT], d3[H], d2[H], d1[H];
        for (int f = 0; f < IN; f++) h0[f] = ((double)x[(size_t)b * IN + f] - mean[f]) * scale[f];
        layer(prm + OW1, prm + OB1, h0, IN, H, 1, h1);
        layer(prm + OW2, prm + OB2, h1, H, H, 1, h2);
        layer(prm + OW3, prm + OB3, h2, H, H, 1, h3);
        layer(prm + OW4, prm + OB4, h3, H, OUT, 0, z);
        for (int k = 0; k < OUT; k++) {
            double y = labels[b] > DMIN + k ? 1.0 : 0.0;
            loss += (z[k] > 0.0 ? z[k] : 0.0) + log1p(exp(-fabs(z[k]))) - y * z[k];
            dz[k] = (1.0 / (1.0 + exp(-z[k])) - y) / n;
        }
        if (!grad) continue;
        for (int j = 0; j < H; j++) {
            double a = 0.0;
            for (int k = 0; k < OUT; k++) a += prm[OW4 + k * H + j] * dz[k];
            d3[j] = h3[j] > 0.0 ? a : 0.0;
        }
        for (int j = 0; j < H; j++) {
            double a = 0.0;
            for (int k = 0; k < H; k++) a += prm[OW3 + k * H + j] * d3[k];
            d2[j] = h2[j] > 0.0 ? a : 0.0;
        }
        for (int j = 0; j < H; j++) {
            double a = 0.0;
            for (int k = 0; k < H; k++) a += prm[OW2 + k * H + j] * d2[k];
            d1[j] = h1[j] > 0.0 ? a : 0.0;
        }
        for (int k = 0; k < OUT; k++) {
            for (int j = 0; j < H; j++) grad[OW4 + k * H + j] += dz[k] * h3[j];
            grad[OB4 + k] += dz[k];
        }
        for (int i = 0; i < H; i++) {
            for (int j = 0; j < H; j++) {
                grad[OW3 + i * H + j] += d3[i] * h2[j];
                grad[OW2 + i * H + j] += d2[i] * h1[j];
            }
            for (int f = 0; f < IN; f++) grad[OW1 + i * IN + f] += d1[i] * h0[f];
            grad[OB3 + i] += d3[i];
            grad[OB2 + i] += d2[i];
            grad[OB1 + i] += d1[i];
        }
    }
    return loss / n;
}
