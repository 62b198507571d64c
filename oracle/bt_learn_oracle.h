/*
 * bt_learn_oracle.h -- CPU restatement of the learned-degree parts
 * (pair labels, degree-model loss and gradient).
 *
 * TEST INFRASTRUCTURE ONLY: only tests/ may load it; the product library
 * never links, loads or calls anything here.
 */
#ifndef BT_LEARN_ORACLE_H
#define BT_LEARN_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Labels and per-degree errors (11 per pair, degrees 2..12, NaN beyond the
 * label) of explicitly given pairs; arguments as BaryTree_LabelPairs. */
int orc_label_pairs(int npairs, const int *kind, const int *tbeg, const int *tcnt, const int *sbeg,
                    const int *scnt, const double *tbox, const double *sbox, const double *tx,
                    const double *ty, const double *tz, const double *sx, const double *sy,
                    const double *sz, const double *sq, int kernel, double p0, double tol,
                    int *labels, double *errors);

/* The same with the subtraction kernel when ss = 1: target charges tq,
 * source weights sw, p1 = 1/kappa^2 (Coulomb) or 1/alpha^2 (regularized
 * Coulomb; p0 = epsilon); Yukawa ignores p1. */
int orc_label_pairs_ss(int npairs, const int *kind, const int *tbeg, const int *tcnt, const int *sbeg,
                       const int *scnt, const double *tbox, const double *sbox, const double *tx,
                       const double *ty, const double *tz, const double *tq, const double *sx,
                       const double *sy, const double *sz, const double *sq, const double *sw, int kernel,
                       double p0, int ss, double p1, double tol, int *labels, double *errors);

/* Degree model in FP64: parameter count, logits (10 per sample), and the
 * batch-mean loss with its gradient (grad may be NULL). */
int orc_mlp_num_params(void);
void orc_mlp_logits(const double *params, const double *mean, const double *scale, int n,
                    const float *x, double *logits);
double orc_mlp_loss_grad(const double *params, const double *mean, const double *scale, int n,
                         const float *x, const int *labels, double *grad);

#ifdef __cplusplus
}
#endif

#endif /* BT_LEARN_ORACLE_H */
