// bt_learn.h -- launch interface of the degree-learning kernels (bt_learn.hip).
// Every pointer is a device pointer.
#pragma once

#include <hip/hip_runtime_api.h>

namespace btl {

// ---- pair labels --------------------------------------------------------
// For pair i: targets [tbeg, tbeg+tcnt) of t{x,y,z}, sources [sbeg,
// sbeg+scnt) of s{x,y,z,q}, boxes {xmin,xmax,ymin,ymax,zmin,zmax} of the
// target and source node.  Up to LABEL_NT targets and LABEL_NS sources are
// sampled at stride (index beg + (k*cnt)/n).  The pair error at degree p is
//   sqrt(sum_i (approx_p(x_i) - exact(x_i))^2 / sum_i exact(x_i)^2)
// with exact(x) = sum_s q_s G(x, y_s) and approx_p the kind's degree-p
// approximation (PC: source interpolation, CP: target interpolation, CC:
// both).  label = least p in [2, 12] with error <= tol, 13 if none.
constexpr int LABEL_NT = 32;
constexpr int LABEL_NS = 128;
constexpr int LABEL_DMIN = 2, LABEL_DMAX = 12;
constexpr int LABEL_ND = LABEL_DMAX - LABEL_DMIN + 1;

struct LabelArgs {
    int npairs;
    const int *kind, *tbeg, *tcnt, *sbeg, *scnt;
    const double *tbox, *sbox;  // 6 per pair
    const double *tx, *ty, *tz;
    const double *sx, *sy, *sz, *sq;
    int kernel;                 // 1 Coulomb, 2 Yukawa (p0 = kappa), 3 regularized Coulomb (p0 = epsilon)
    double p0;
    // subtraction (ss = 1): the pair's potential is that of the subtraction
    // kernel, sum_s (q_s - q_t g_w(r)/g(r)) w_s g(r): target charges tq and
    // source weights sw, p1 = 1/kappa^2 (Coulomb) or 1/alpha^2 (regularized
    // Coulomb) for the Gaussian of g_w; Yukawa's g_w is g
    int ss;
    const double *tq, *sw;
    double p1;
    double tol;
    int *labels;
    double *errors;             // LABEL_ND per pair (NaN where not computed)
};
hipError_t launch_label(const LabelArgs &a, hipStream_t st);

// ---- degree model: MLP 8 -> 64 -> 64 -> 64 -> 10 (ReLU), FP32 ----------
// Parameters, one flat array: W1[64][8] b1[64] W2[64][64] b2[64]
// W3[64][64] b3[64] W4[10][64] b4[10].  Output k is the logit of
// "the pair needs a degree above 2+k" (k = 0..9).
constexpr int MLP_IN = 8, MLP_H = 64, MLP_OUT = 10;
constexpr int OFF_W1 = 0;
constexpr int OFF_B1 = OFF_W1 + MLP_H * MLP_IN;
constexpr int OFF_W2 = OFF_B1 + MLP_H;
constexpr int OFF_B2 = OFF_W2 + MLP_H * MLP_H;
constexpr int OFF_W3 = OFF_B2 + MLP_H;
constexpr int OFF_B3 = OFF_W3 + MLP_H * MLP_H;
constexpr int OFF_W4 = OFF_B3 + MLP_H;
constexpr int OFF_B4 = OFF_W4 + MLP_OUT * MLP_H;
constexpr int MLP_PARAMS = OFF_B4 + MLP_OUT;

// Buffers of a batch of b samples (device, feature-major [f][b]): a0 the
// normalised inputs [8][b], a1..a3 the layer outputs [64][b], d1..d3 the
// pre-activation gradients [64][b], dz the logits and then their gradients
// [10][b], loss the per-sample losses [b].
struct TrainBufs {
    float *a0, *a1, *a2, *a3, *d1, *d2, *d3, *dz, *loss;
};

// Inference on x[n][8] raw features, normalised as (x - mean) * scale (tb
// sized for n samples): logits[n][10] (optional) and degree[n] (optional).
// The degree is 3+k for the largest k with sigmoid(logit_k) >= delta (the
// model gives at least probability delta that degree 2+k is not enough),
// else 2.
hipError_t launch_mlp_infer(const float *params, const float *mean, const float *scale, int n,
                            const float *x, float delta, float *logits, int *degree,
                            const TrainBufs &tb, hipStream_t st);

// Training pass over one batch of b samples x[idx[j]] (idx may be null: j):
// forward, loss, backward into tb.  The loss is the batch mean of
// sum_k softplus(z_k) - y_k z_k with y_k = [label > 2+k].
hipError_t launch_mlp_fwdbwd(const float *params, const float *mean, const float *scale,
                             const float *x, const int *labels, const int *idx, int b,
                             const TrainBufs &tb, hipStream_t st);
// Weight gradients from the buffers of launch_mlp_fwdbwd: grad has the
// parameter layout; dW = D^T A and db = sum over the batch, in batch order.
hipError_t launch_mlp_wgrad(const TrainBufs &tb, int b, float *grad, hipStream_t st);
// Adam step t (1-based) on n parameters.
hipError_t launch_adam(float *params, const float *grad, float *m, float *v, int n, float lr,
                       float beta1, float beta2, float eps, int t, hipStream_t st);

}  // namespace btl
