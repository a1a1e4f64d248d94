/*
 * CPU oracle for the SBE right-hand side, its Dormand-Prince RK5(4)
 * propagation and the current reduction.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the HIP path is
 * compared against, and the CPU baseline that bench.py times beside it.  The
 * product (libsbewh_hip.so, package sbewh_amd) never links, loads or calls it.
 *
 * It restates, in plain C with OpenMP, the reference's algorithm from its
 * source text (citations per function in sbe_oracle.c).  Parity pins are in
 * tests/test_oracle_*.py: closed forms, invariants, brute force and the
 * reference's own test fixtures.
 */
#ifndef SBE_ORACLE_H
#define SBE_ORACLE_H

#include "../include/sbewh_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_ctx oracle_ctx;

/* method for the Wannier->Bloch phase sum */
enum { ORACLE_PHASE_DIRECT = 0, ORACLE_PHASE_SEPARABLE = 1 };

oracle_ctx *oracle_create(const sbe_model *model, const sbe_grid *grid,
                          const sbe_params *params, const sbe_subpulse *pulses,
                          int num_pulses);
void oracle_destroy(oracle_ctx *c);
const char *oracle_last_error(void);
int oracle_set_pulses(oracle_ctx *c, const sbe_subpulse *pulses, int num_pulses);
void oracle_set_phase_method(oracle_ctx *c, int method);

long oracle_state_size(const oracle_ctx *c);   /* complex elements */
long oracle_num_k(const oracle_ctx *c);
int oracle_padded_mat_size(const oracle_ctx *c);
void oracle_get_nf(const oracle_ctx *c, int nf[3]);
void oracle_k_frac(const oracle_ctx *c, long k, double kf[3]); /* grid part only */

/* pulses */
void oracle_pulse_E(const oracle_ctx *c, double t, double E[3]);
void oracle_pulse_A(const oracle_ctx *c, double t, double A[3]);
double oracle_start_time(const oracle_ctx *c);
double oracle_end_time(const oracle_ctx *c);

/* phase sums */
void oracle_hk_point(const oracle_ctx *c, const double kfrac[3], const double E[3],
                     double *H0, double *HE);
void oracle_hk_full_point(const oracle_ctx *c, const double kfrac[3], double *H0,
                          double *D /*[3][n*n]*/, double *dH /*[3][n*n]*/);
/* H0(k+A), HE(k+A) for every k of the grid, in state k-order:
 * out[k][0] = H0, out[k][1] = HE, each n*n complex */
void oracle_hk_mesh(oracle_ctx *c, const double A[3], const double E[3], double *out);

/* Jacobi eigen-decomposition (cyclic, fixed sweeps).  A (n*n complex,
 * row-major, Hermitian) is overwritten with U, U*A*U^dagger = diag(eig). */
void oracle_jacobi(int n, int sweeps, double *A, double *eig);

/* the RHS functor and friends (reference layout of the state) */
int oracle_rhs(oracle_ctx *c, const double *state, double *dsdt, double t);
int oracle_initial_state(oracle_ctx *c, double mu, double *state);
int oracle_fermi_level(oracle_ctx *c, int occupied_bands, int below_gap, double *mu);
int oracle_observe_j(oracle_ctx *c, const double *state, double j[3]);
/* full observer (regions, occupations, density): per region r
 * weight_sum[r], j[r][3], occ_ham[r][n], occ_wan[r][n], wan[r][n*n] complex
 * (optional), and occ_min[n], occ_max[n] over k */
int oracle_observe_full(oracle_ctx *c, const double *state, const sbe_region *regions, int num_regions,
                        double smearing, double *weight_sum, double *j, double *occ_ham, double *occ_wan,
                        double *occ_min, double *occ_max, double *wan);

/* integrate_const( controlled dopri5, rhs, state, t0, t1 + dt_out/2, dt_out, observer )
 * over the pulse window with output_count outputs.  j_out: [output_count][3],
 * t_out: [output_count].  log receives every step attempt.  If replay is not
 * NULL its accepted entries dictate the step sizes.  Returns the number of
 * accepted steps (>= 0) or a negative error code. */
long oracle_integrate(oracle_ctx *c, double *state, int output_count, double *t_out,
                      double *j_out, sbe_steplog *log, const sbe_steplog *replay,
                      long *rhs_calls);

/* OpenMP threads the oracle's parallel loops use */
int oracle_num_threads(void);

/* wall seconds of `count` dopri5 attempts at fixed (t, dt) (CPU baseline) */
double oracle_bench_attempts(oracle_ctx *c, const double *state, double t, double dt, int count);

#ifdef __cplusplus
}
#endif

#endif
