/*
 * C ABI of libsbewh_hip.so: the MI355X (gfx950) implementation of the SBE
 * right-hand side, its adaptive Dormand-Prince RK5(4) propagation and the
 * current reduction.  Plain pointers and sizes only.
 *
 * State vectors use the reference's CPU layout (code/src/StateContext.h:7-25):
 * complex entries, interleaved (re, im) doubles,
 *   [A_x, A_y, A_z, 5 zero pads][rho_k0 (n*n, row-major, zero padded to P)][rho_k1]...
 * with P = 8 * ceil(n*n / 8) and k in Fourier-block order
 * (code/src/WhFourierTransform.cpp:137-149, 303-312).  Length
 * 8 + P * N_k complex entries (sbe_state_size).  A context created with
 * SBE_FLAG_GPU_LAYOUT uses the reference's GPU layout instead (P = n*n, the
 * same k order; code/src/CuPropagator.cu:50-52,111-114) in every entry point.
 *
 * Every function returns 0 on success and a negative code on error; the
 * message is available from sbe_last_error(ctx) (or sbe_last_error(NULL) for
 * failures of sbe_create).  Nothing throws across the ABI.  A context owns
 * all device memory it uses and one HIP stream; use one context per host
 * thread.  Calls are synchronous unless stated otherwise.
 */
#ifndef SBEWH_HIP_H
#define SBEWH_HIP_H

#include "sbewh_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct sbe_ctx sbe_ctx;

enum sbe_status {
    SBE_OK = 0,
    SBE_ERR_ARG = -1,        /* invalid argument / unsupported configuration */
    SBE_ERR_HIP = -2,        /* HIP runtime or launch failure */
    SBE_ERR_STEP = -3,       /* step-size control failed (500 rejected tries) */
    SBE_ERR_PHYSICS = -4,    /* e.g. no unique band gap for the Fermi level */
    SBE_ERR_REPLAY = -5      /* replay log exhausted / inconsistent */
};

const char *sbe_version(void);
const char *sbe_last_error(const sbe_ctx *ctx);

/* Construction from the real-space model, grid, parameters and pulse.
 * Replaces PropagatorCPU(const StateContext&, WhFourierTransform&)
 * (code/src/PropagatorCPU.h:52) and CuPropagator(cudaStream_t, ...)
 * (code/src/CuPropagator.h:31) together with the WhFourierTransform /
 * CuWhFourierTransformHk setup (code/src/WhFourierTransform.cpp:3-38,
 * code/src/CuWhFourierTransform.cu:3-40).  Applies the reference's Nf rule
 * (code/src/Parameter.cpp:4-56) and dimensional projection
 * (code/src/util/w90util.cpp:417-444).  Supports n <= 16 bands. */
int sbe_create(const sbe_model *model, const sbe_grid *grid, const sbe_params *params,
               const sbe_subpulse *pulses, int num_pulses, sbe_ctx **out);
void sbe_destroy(sbe_ctx *ctx);

/* Replace the pulse (the reference's multi-configuration loop assigns
 * param.pulse per configuration, code/src/main/runSBE.cpp:205-216). */
int sbe_set_pulses(sbe_ctx *ctx, const sbe_subpulse *pulses, int num_pulses);

long sbe_state_size(const sbe_ctx *ctx);   /* complex entries */
long sbe_num_k(const sbe_ctx *ctx);
/* complex stride P between consecutive rho_k of the context's state layout */
int sbe_matrix_stride(const sbe_ctx *ctx);
/* grid[k][3] = integer grid indices (i_0, i_1, i_2), 0 <= i_d < Nk_d, of the
 * state's k-th density matrix (fractional k = i_d / Nk_d + kGlobalShift):
 * the Fourier-block order of code/src/WhFourierTransform.cpp:137-149,303-312
 * (getFineMeshFracShifts / getFFTfracShift), the same in both layouts. */
int sbe_k_index_map(const sbe_ctx *ctx, int *grid);
int sbe_get_nf(const sbe_ctx *ctx, int nf[3]);
/* pulse window [t0, t1] (code/src/pulse/Pulse3D.cpp:42-62) */
int sbe_time_window(const sbe_ctx *ctx, double *t0, double *t1);
/* E(t) of the pulse (Pulse3D::getE, code/src/pulse/Pulse3D.cpp:15-21) */
int sbe_pulse_E(const sbe_ctx *ctx, double t, double E[3]);

/* E(t) and the window [t0, t1] of a set of sub-pulses without a context
 * (host only, no GPU needed): Pulse3D::getE / getStartTime / getEndTime,
 * code/src/pulse/Pulse3D.cpp:15-21,42-62, over the closed forms of
 * code/src/pulse/{Sin2Pulse1D,GaussianPulse1D,PureGauss1D,Sin2RampPulse1D}.cpp. */
int sbe_pulse_field(const sbe_subpulse *pulses, int num_pulses, double t, double E[3]);
int sbe_pulse_window(const sbe_subpulse *pulses, int num_pulses, double *t0, double *t1);

/* Fermi level from all eigenvalues at A = 0 (getFermiLevelCPU / calcFermiLevelBase,
 * code/src/main/runSBE.cpp:97-155).  below_gap != 0 selects occupiedBelowBandGap. */
int sbe_fermi_level(sbe_ctx *ctx, int occupied_bands, int below_gap, double *mu);

/* Initial Fermi-Dirac state, host buffer of sbe_state_size complex entries
 * (PropagatorCPU::calcInitialState, code/src/PropagatorCPU.h:54;
 *  CuPropagator::calcInitialState, code/src/CuPropagator.h:33). */
int sbe_initial_state(sbe_ctx *ctx, double mu, double *state);

/* The RHS functor on host buffers: dsdt = f(state, t), dsdt fully overwritten.
 * Replaces void PropagatorCPU::operator()(const CPUstate&, CPUstate&, double)
 * (code/src/PropagatorCPU.h:55, body :208-256). */
int sbe_rhs(sbe_ctx *ctx, const double *state, double *dsdt, double t);

/* The same on device buffers, enqueued on `stream` (a hipStream_t; NULL = the
 * context's stream) and returning without synchronising.  Replaces
 * CuPropagator::operator()(const GPUstate&, GPUstate&, double)
 * (code/src/CuPropagator.h:34). */
int sbe_rhs_device(sbe_ctx *ctx, const double *d_state, double *d_dsdt, double t, void *stream);

/* Current j(t) of a host state (ObserverCPU::operator() with the default
 * region, code/src/ObserverCPU.h:43, code/src/ObserverCPU.cpp:149-210,293-308). */
int sbe_observe(sbe_ctx *ctx, const double *state, double j[3]);

/* No counterpart in the reference: sbe_observe for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops).
 * d_state: device state in the context's layout; d_j: device buffer of 3
 * doubles that receives j.  Enqueued on `stream` (NULL: the context's
 * stream); returns without synchronising.  Same kernels and summation order
 * as sbe_observe.  The per-k partial currents go to a file-scope device array
 * of 3 * 262144 doubles (6 MiB, held on each GPU the library runs on until the
 * process ends), so grids of at most 262144 k-points are accepted (SBE_ERR_ARG
 * beyond).  Every call on one GPU uses that GPU's array: calls that overlap
 * on one GPU race. */
int sbe_observe_device_noalloc(sbe_ctx *ctx, const double *d_state, double *d_j, void *stream);

/* Whole propagation, device resident:
 * integrate_const(make_controlled<runge_kutta_dopri5>(eps_abs, eps_rel), rhs, state,
 *                 t0, t1 + dt_out/2, dt_out, observer)
 * (code/src/main/runSBE.cpp:170-183, code/src/main/CuRunSBE.cu:19-31) with
 * dt_out = (t1 - t0)/(output_count - 1).  state (host) is advanced in place.
 * t_out[output_count], j_out[output_count][3] receive the observer output
 * (either may be NULL).  log (optional) receives every step attempt.  If
 * replay is non-NULL its accepted entries dictate the step sizes.  Returns
 * the number of accepted steps (>= 0) or an error code; rhs_calls (optional)
 * receives the number of RHS evaluations. */
long sbe_integrate(sbe_ctx *ctx, double *state, int output_count, double *t_out, double *j_out,
                   sbe_steplog *log, const sbe_steplog *replay, long *rhs_calls);
/* 1 if the last sbe_integrate ran on the resident integrator
 * (SBE_FLAG_DEVICE_CONTROLLER), 0 if on the host controller.  No
 * counterpart in the reference (whose controller is Boost.Odeint on the host). */
int sbe_integrate_info(const sbe_ctx *ctx, int *device_controller);

/* ---- full observer ------------------------------------------------------
 * ObserverCPU with k-space regions, occupations and the region-averaged
 * density matrix (code/src/ObserverCPU.h:43, body code/src/ObserverCPU.cpp:
 * 137-260, normalizeExpectationValues :293-308; regions from
 * code/src/ObserverBase.cpp:3-17).  At each recorded output time and for
 * every region sample r (sbe_region, weights w_r(k)):
 *   weight_sum[r] = sum_k w_r(k)
 *   j[r][d]       = sum_k w_r(k) Re[i Tr(D_d [H0, rho_k]) - Tr(dH0/dk_d rho_k)] / weight_sum[r]
 *   occ_ham[r][b] = sum_k w_r(k) o_b(k) / weight_sum[r],  occ_wan[r][b] = sum_k w_r(k) Re rho_k,bb / weight_sum[r]
 *   wan[r]        = sum_k w_r(k) rho_k / weight_sum[r]  (records with index % density_stride == 0)
 *   occ_ham_min[b], occ_ham_max[b] = min / max over k of o_b(k)
 * with H0, D, dH0/dk at k + A, o_b(k) the Hamiltonian occupation
 * <v_b|rho_k|v_b> of the b-th eigenvector of H0(k + A) in ascending energy,
 * smeared over energy, o_b <- sum_q o_q exp(-((e_b - e_q)/w)^2) / sum_q exp(..)
 * when smearing_width w > 0 (the reference's occupationSmearingWidth; w <= 0:
 * unsmeared).  Bands are numbered in ascending energy, the order of the
 * reference's default diagonaliser (LAPACK zheev). */
typedef struct sbe_observer sbe_observer;

/* Collected records, caller-allocated arrays (any may be NULL); count =
 * records, R = regions, n = bands, density records counted separately:
 * t[count], E[count][3] (E(t), zero beyond dim), A[count][3],
 * weight_sum[count][R], j[count][R][3], occ_ham[count][R][n],
 * occ_wan[count][R][n], occ_ham_min[count][n], occ_ham_max[count][n],
 * wan[density_count][R][n][n] (complex, interleaved), density_index[density_count]
 * (the record index of each density record; ExpectationValues_t::tDensIndices). */
typedef struct sbe_observations {
    double *t, *E, *A, *weight_sum, *j, *occ_ham, *occ_wan, *occ_ham_min, *occ_ham_max, *wan;
    int *density_index;
} sbe_observations;

/* regions[num_regions]: the region samples in output order (the reference puts
 * the whole-zone region "expectationValues" first, then the KspaceRegions
 * entries by name, each expanded over its sample grid).  density_stride 0: no
 * density matrices.  capacity: records the observer can hold. */
int sbe_observer_create(sbe_ctx *ctx, const sbe_region *regions, int num_regions, double smearing_width,
                        int density_stride, int capacity, sbe_observer **out);
void sbe_observer_destroy(sbe_observer *obs);
/* forget the records (ObserverCPU::reset, code/src/ObserverCPU.cpp:131-136) */
int sbe_observer_reset(sbe_observer *obs);
/* ObserverCPU::operator()(const CPUstate&, double t) on a host state */
int sbe_observer_record(sbe_observer *obs, const double *state, double t);
/* the same on a device state, enqueued on `stream` (NULL: the context's) */
int sbe_observer_record_device(sbe_observer *obs, const double *d_state, double t, void *stream);
/* records so far (return value) and how many of them hold density matrices */
int sbe_observer_count(const sbe_observer *obs, int *density_count);
/* copy the records out (collectExpectationValues, code/src/ObserverBase.cpp:66-112) */
int sbe_observer_results(sbe_observer *obs, sbe_observations *out);
/* sbe_integrate with the full observer recording every output time (its
 * capacity must cover output_count more records) */
long sbe_integrate_observed(sbe_ctx *ctx, double *state, int output_count, sbe_observer *obs, double *t_out,
                            double *j_out, sbe_steplog *log, const sbe_steplog *replay, long *rhs_calls);

/* Kernel-level entry points (used by tests and the benchmark). */

/* H0(k + A) and H_E(k + A) for every k, out[k][2][n*n] complex (the
 * transformMeshFFT(HkVec...) product, code/src/WhFourierTransform.cpp:259-315). */
int sbe_hk_mesh(sbe_ctx *ctx, const double A[3], const double E[3], double *out);

/* Batched Jacobi on `count` Hermitian n x n matrices H (host, complex):
 * U (count x n x n) with U H U^dagger = diag(eig)  (DiagonalizatorSingle::diag,
 * code/src/util/DiagonalizatorSingle.cpp:99-122; CuDiagonalizator::diag,
 * code/src/CuDiagonalizator.cu:45-51). */
int sbe_jacobi_batch(int n, int sweeps, long count, const double *H, double *U, double *eig);

/* Layout chosen for the fused stage kernel by the built-in tuner (lanes per
 * matrix row; 0 before the first RHS evaluation) and, if the tuner ran, the
 * timed variants (up to 8) with their average launch time in ms.  No
 * reference counterpart: the reference's GPU path has no tunable kernels. */
int sbe_stage_tuning(const sbe_ctx *ctx, int *lanes_per_row, int *count, int *variants, double *ms);

/* Device-resident timing of the adaptive propagation: `warmup`, then `steps`
 * Dormand-Prince step attempts of the controller (accept / reject and
 * step-size adaptation as in sbe_integrate, without truncation onto output
 * times) from the host state at time t with initial step dt; the state
 * advances on the device.  ms: wall time of the timed attempts; rhs: RHS
 * evaluations among them; stage_ms / phase_ms (optional): average duration per
 * RHS of the fused stage kernel and of the phase sum, from HIP events in a
 * separate pass of up to 20 attempts. */
int sbe_bench_steps(sbe_ctx *ctx, const double *state, double t, double dt, int warmup, int steps, double *ms,
                    long *rhs, double *stage_ms, double *phase_ms);

/* High-harmonic spectrum of a current j(t) (host buffers, nt samples of
 * ncomp <= 3 components, uniform t): the reference's post-processing chain
 * fadeout + zero padding + |FFT|^2 omega^(2 power) / N
 * (scripts/util/hhg.py:9-25), masked to omega_min < omega < omega_max
 * (scripts/util/common.py:187-201), j normalised by its global maximum and
 * the per-component spectra summed (scripts/showExpectationValues.py:36-58).
 * Writes up to `capacity` points; *count receives the number of points (if it
 * exceeds capacity nothing is written and SBE_ERR_ARG is returned). */
int sbe_hhg_spectrum(const double *t, const double *j, int nt, int ncomp, double d_omega, double omega_min,
                     double omega_max, double fadeout_rel, int power, double *omega_out, double *S_out,
                     int capacity, int *count);

/* The non-FFT branch of calcSpectrum (scripts/util/common.py:187-201):
 * sbe_direct_omega is directOmega (scripts/util/hhg.py:27-38): the grid
 * omega_min, omega_min + d, ... below omega_max, with omega_max lowered to the
 * Nyquist limit pi/dt and d raised to pi/(t_end - t_0) where needed (count
 * receives the length; nothing written beyond capacity, SBE_ERR_ARG then);
 * sbe_direct_spectrum is directSpectrum (hhg.py:40-42) of one real signal:
 * S(omega) = |omega^power sum_i exp(i omega t_i) j_i|^2; sbe_hhg_spectrum_direct
 * is sbe_hhg_spectrum with the direct sum in place of the FFT (same j
 * normalisation and component sum, no fadeout). */
int sbe_direct_omega(const double *t, int nt, double omega_min, double omega_max, double d_omega, double *omega_out,
                     int capacity, int *count);
int sbe_direct_spectrum(const double *t, int nt, const double *omega, int nw, const double *j, int power,
                        double *S_out);
int sbe_hhg_spectrum_direct(const double *t, const double *j, int nt, int ncomp, double d_omega, double omega_min,
                            double omega_max, int power, double *omega_out, double *S_out, int capacity, int *count);

/* smearSpectrum (scripts/util/hhg.py:45-49): Gaussian filter of full width at
 * half maximum smear_width (in omega units) over the uniform grid omega, zero
 * outside, kernel truncated at 4 standard deviations and normalised (the
 * scipy.ndimage.gaussian_filter1d call); smear_width < 0 copies S.  out may
 * alias S. */
int sbe_smear_spectrum(const double *omega, const double *S, int count, double smear_width, double *out);

/* Per-harmonic yield Y_h = sum of S over omega in [h w0 - width w0, h w0 + width w0),
 * h = 1..hmax (extractProfiles, scripts/ellipticityDependency.py:77-83);
 * omega ascending. */
int sbe_harmonic_yields(const double *omega, const double *S, int count, double omega0, double width, int hmax,
                        double *yields);

/* ---- batched scan: many pulses / grids / T2 over one model -------------
 * Replaces the reference's multi-run loop, which propagates the pulses of a
 * multiRunFile one after the other with one integrate_const each and one
 * shared initial state (code/src/main/runSBE.cpp:205-216; GPU:
 * code/src/main/CuRunSBE.cu:55-71; configurations read by
 * code/src/io/PulseParameter.cpp:201-211).  Here all instances advance
 * together on one stream, each with its own adaptive step size and its own
 * integrate_const output grid; per instance the results are those of
 * sbe_integrate on a context created with the same inputs. */
typedef struct sbe_scan sbe_scan;

/* grid: dim, Nf and k_global_shift shared by all instances (Nk is taken from
 * each instance); params: shared except relaxation_F2 (per instance);
 * instance i uses pulses[pulse_offset .. pulse_offset + num_pulses). */
int sbe_scan_create(const sbe_model *model, const sbe_grid *grid, const sbe_params *params,
                    const sbe_instance *instances, int count, const sbe_subpulse *pulses, int total_pulses,
                    sbe_scan **out);
void sbe_scan_destroy(sbe_scan *scan);
const char *sbe_scan_last_error(const sbe_scan *scan);

/* Propagate every instance from the Fermi-Dirac ground state of its grid
 * (occupied_bands, as sbe_fermi_level; computed once per distinct grid) over
 * its pulse window with output_count outputs.  t_out: [count][output_count],
 * j_out: [count][output_count][3], steps: [count] accepted steps (or a
 * negative status for an instance whose step control failed; the others
 * finish), rhs_calls: [count] (optional), seconds: wall time (optional). */
int sbe_scan_run(sbe_scan *scan, int occupied_bands, int output_count, double *t_out, double *j_out, long *steps,
                 long *rhs_calls, double *seconds);

/* state of instance `index` after sbe_scan_run (sbe_scan_state_size complex entries) */
long sbe_scan_state_size(const sbe_scan *scan, int index);
int sbe_scan_final_state(const sbe_scan *scan, int index, double *state);

/* stage-kernel layout chosen by the tuner on the batched launch (as sbe_stage_tuning) */
int sbe_scan_tuning(const sbe_scan *scan, int *lanes_per_row, int *count, int *variants, double *ms);

#ifdef __cplusplus
}
#endif

#endif
