/*
 * C ABI of the learned surrogate in libsbewh_hip.so: an ensemble of
 * multilayer perceptrons (SiLU hidden layers, linear output) trained and
 * evaluated on the MI355X matrix cores with BF16 weights and activations and
 * FP32 accumulation (master weights and Adam moments in FP32).
 *
 * No counterpart in the reference: sbewh has no learned model.  The
 * surrogate maps pulse parameters to per-harmonic log-yields
 * (sbewh_amd/surrogate.py); the solver (sbewh_hip.h) generates its training
 * data and remains the fallback.
 *
 * Parameters of one member, FP32, layer by layer: W_l [d_{l+1}][d_l]
 * (row-major) then b_l [d_{l+1}], l = 0 .. layers-1.  Members train on
 * independent shuffles and initialisations.  Status codes as sbewh_hip.h.
 */
#ifndef SBEWH_SURROGATE_H
#define SBEWH_SURROGATE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sbe_mlp sbe_mlp;

enum sbe_mlp_status { SBE_MLP_OK = 0, SBE_MLP_ERR_ARG = -1, SBE_MLP_ERR_HIP = -2 };

/* dims[0 .. layers]: input width, hidden widths, output width */
int sbe_mlp_create(int members, int layers, const int *dims, unsigned long long seed, sbe_mlp **out);
void sbe_mlp_destroy(sbe_mlp *mlp);
const char *sbe_mlp_last_error(const sbe_mlp *mlp);
long sbe_mlp_param_count(const sbe_mlp *mlp);   /* per member */
int sbe_mlp_get_params(const sbe_mlp *mlp, int member, float *params);
int sbe_mlp_set_params(sbe_mlp *mlp, int member, const float *params);

/* y [members][count][d_out] = every member applied to x [count][d_in] */
int sbe_mlp_forward(sbe_mlp *mlp, int count, const float *x, float *y);

/* mean-squared-error loss of every member on one batch (count <= 4096 rows
 * x, targets t) and its gradient grad [members][P] (no update) */
int sbe_mlp_gradient(sbe_mlp *mlp, int count, const float *x, const float *t, float *loss, float *grad);

/* Adam (beta 0.9 / 0.999, cosine-decayed learning rate from lr) over `epochs`
 * shuffled passes of `batch`-row batches (batch <= 4096) of the data set
 * x [count][d_in], t [count][d_out]; loss_out [epochs][members]: mean batch
 * loss per epoch (optional); seconds: wall time (optional). */
int sbe_mlp_train(sbe_mlp *mlp, int count, const float *x, const float *t, int batch, int epochs, float lr,
                  unsigned long long seed, float *loss_out, double *seconds);

#ifdef __cplusplus
}
#endif

#endif
