/*
 * spx_gnn.h -- the learned control variate of libspx.
 *
 * No counterpart in the reference: a small message-passing graph network
 * over one disorder realisation (couplings J_ij, fields h_i) that predicts
 * its echo curve, trained on curves computed exactly by the propagator and
 * used as a control variate for disorder averaging (the ensemble mean =
 * surrogate mean over many cheap geometries + exact correction on fewer).
 *
 * Network, for a realisation of L spins (node i, hidden width 128), with
 * x_ij = J_ij / s_J, degrees d_i = sum_j |x_ij| and edge weights
 * A_ij = x_ij / sqrt((1 + d_i)(1 + d_j)) (so |A_ij| < 1):
 *   f_i  = (h_i / s_h, d_i, log(1 + sum_j x_ij^2), 1, 0...)     16 inputs
 *   x0   = relu(W0 f + b0)
 *   x_l  = relu(W_l [x_{l-1}, sum_j A_ij x_{l-1,j}] + b_l),       l = 1..3
 *   y    = Wout mean_i x3 + bout                                 P outputs
 * GEMMs run as BF16 MFMA with FP32 accumulation; parameters are kept in
 * FP32 and rounded to BF16 for every product; Adam updates in FP32.
 * Graphs of one batch share L; J must be symmetric. Loss: mean squared
 * error over G x P.
 */
#ifndef SPX_GNN_H
#define SPX_GNN_H

#include <stddef.h>
#include <stdint.h>

#include "spx.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct spx_gnn spx_gnn;

/* outputs = P curve points (P <= 128); seed initialises the weights. */
int spx_gnn_create(spx_ctx* ctx, size_t outputs, uint64_t seed, double scale_h, double scale_j,
                   spx_gnn** out);
int spx_gnn_destroy(spx_gnn* g);
/* Parameter vector (FP32 masters) in the order W0 b0 W1 b1 W2 b2 W3 b3 Wout bout. */
int spx_gnn_num_params(const spx_gnn* g, size_t* n);
int spx_gnn_get_params(const spx_gnn* g, float* host);
int spx_gnn_set_params(spx_gnn* g, const float* host);
/* Forward on `ngraphs` graphs of `nspins`: J [G][L][L], h [G][L] (FP64
 * host), predictions [G][P]. */
int spx_gnn_predict(spx_gnn* g, size_t ngraphs, size_t nspins, const double* J, const double* h,
                    float* pred);
/* Forward on realisations first .. first+count-1 of the dipolar ensemble
 * (the draws of spx_dipolar_graphs), generated on host threads chunk by
 * chunk while the device runs the previous chunk; predictions
 * [count][P]. The surrogate-only side of the control variate. */
int spx_gnn_predict_dipolar(spx_gnn* g, uint64_t seed, size_t first, size_t count, size_t nspins,
                            double scale, double min_separation, int angular, double disorder,
                            float* pred);
/* Loss and gradient (same order as the parameters) for targets [G][P],
 * without updating: the gradient check's entry point. */
int spx_gnn_gradient(spx_gnn* g, size_t ngraphs, size_t nspins, const double* J, const double* h,
                     const float* target, float* loss, float* grad);
/* One Adam step (lr, beta1 0.9, beta2 0.999, eps 1e-8) on one batch. */
int spx_gnn_train_step(spx_gnn* g, size_t ngraphs, size_t nspins, const double* J, const double* h,
                       const float* target, float lr, float* loss);
/* Keep a training set resident on the device: n graphs of `nspins`
 * (J [n][L][L] symmetric, h [n][L], targets [n][P]); replaces any earlier
 * set. */
int spx_gnn_set_training_set(spx_gnn* g, size_t n, size_t nspins, const double* J, const double* h,
                             const float* target);
/* One epoch over the resident set: floor(n / batch) Adam steps, step s on
 * graphs order[s*batch .. (s+1)*batch) (order: a permutation of 0..n-1),
 * gathered on the device; the same updates as spx_gnn_train_step on those
 * batches, with one host synchronisation per epoch. mean_loss: the mean
 * of the steps' losses. */
int spx_gnn_train_epoch(spx_gnn* g, size_t batch, const uint32_t* order, float lr, float* mean_loss);

#ifdef __cplusplus
}
#endif
#endif /* SPX_GNN_H */
