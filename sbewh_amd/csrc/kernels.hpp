// Host-side launchers of the CDNA4 kernels.  Every launcher enqueues on the
// given stream and returns a hipError_t; no launcher synchronises.
#pragma once
#include <hip/hip_runtime.h>

namespace sbe {

struct KOrderH {
    int Nf[3];
    int Nrep[3];
};

// ---- Wannier -> Bloch phase sum (phase_kernels.hip) ----------------------
struct PhaseGeom {
    int n;          // bands
    int dim;
    int W[3];       // dense R box extents
    int rmin[3];
    int Nk[3];
    double lattice[9];
    double kshift_au[3];
};

// Stage input rho_in = x + sum_{j<nt} c[j] k[j] over the state: the
// Dormand-Prince stage combination, formed by the kernels that consume it.
struct StageInput {
    const double *x;
    const double *k[6];
    double c[6];
    int nt;
};

// Twiddles tw_d[i*W_d + r] = exp(i 2 pi (i/Nk_d + off_d)(r + rmin_d)), off =
// frac(kshift_au + A).  A is Re of the first dim special slots of the stage
// input *in (A = 0 if in == null); use_shift = 0 drops kshift_au.
hipError_t launch_twiddles(const PhaseGeom &g, const StageInput *in, int use_shift, double *tw, hipStream_t s);

// Operator merge of the model box for the RHS, done once per evaluation:
//   out[b][e]      = H(R_b)[e]
//   out[b][nn + e] = (H + sum_{d<dim} E_d r_d)(R_b)[e]      (b < nbox, e < nn)
// (code/src/WhFourierTransform.cpp:63-66).  out = null: no merge.
struct ModelMerge {
    const double *Hb, *Db;
    double E[3];
    long nbox;
    int nn, dim;
    double *out;
};
// twiddles (as launch_twiddles) and the model merge in one launch
hipError_t launch_phase_prep(const PhaseGeom &g, const StageInput *in, int use_shift, double *tw,
                             const ModelMerge &mm, hipStream_t s);

// One contraction stage of the separable phase sum (see phase_kernels.hip).
struct ContractArgs {
    const double *in;    // mode 0: [outer][W][inner][M]
    int ti;              // i-tile height of the tiled kernel
    double *out;         // out mode 0: [outer][Nkd][inner][M]; 1: [k][M] via kmap
    const double *tw;    // [Nkd][W]
    int W, Nkd;
    long outer, inner;
    int M, nn, in_mode, out_mode, dim;
    long nbox;
    const double *Hb, *Db, *dHb;
    double E0, E1, E2;
    const int *kmap;
};
// a stage with its launch shape: kind 0 tiled contraction (grid blocks_x x
// blocks_y), 1 generic contraction, 2 operator merge of the model box (dim 1)
struct ContractStage {
    ContractArgs a;
    int kind;
    long blocks_x;
    int blocks_y;
};

// Full separable phase sum into mesh[k][nops][n*n] (state k order).
//   ops == 2 : {H0, HE = H0 + sum_d E_d D_d}
//   ops == 1 + 2 dim : {H0, D_0.., dH0/dk_0..}
// model arrays are dense over the R box: Hb[box][nn], Db[3][box][nn], dHb[3][box][nn].
struct PhaseBuffers {
    const double *Hb, *Db, *dHb;
    const int *kmap;       // grid index -> k index
    double *tw;            // twiddle scratch (sum_d Nk_d W_d complex)
    double *y1, *y2;       // stage scratch
};
hipError_t launch_phase_sum(const PhaseGeom &g, const PhaseBuffers &b, int ops, double E0, double E1,
                            double E2, double *mesh, hipStream_t s);

// Phase sum over every dimension but the last one L = dim - 1 (the dimension
// along which consecutive k indices run):
//   Y[rest][r_L][ops nn],  rest = grid index over the dimensions < L,
// ops == 2: {H0, H_E}; ops == 1 + 2 dim: {H0, D_0.., dH0/dk_0..}.
// The consuming kernel contracts r_L itself.  For dim == 1, Y is the
// operator-merged model box [W_0][ops nn].
// merged: the model box merged by launch_phase_prep (ops == 2; for dim == 1 it is
// Y itself and no stage runs), or null to read the model operators directly.
hipError_t launch_phase_partial(const PhaseGeom &g, const PhaseBuffers &b, int ops, double E0, double E1,
                                double E2, double *Y, const double *merged, hipStream_t s);
// the stages of launch_phase_partial (at most 2), returns their count; a tiled
// stage gets the largest row tile that leaves >= min_blocks blocks
int phase_partial_plan(const PhaseGeom &g, const PhaseBuffers &b, int ops, double E0, double E1, double E2,
                       double *Y, const double *merged, ContractStage *st, long min_blocks = 2048);
// complex entries of Y and of the stage scratch it needs (y1)
long phase_partial_size(const PhaseGeom &g, int ops);
long phase_partial_scratch(const PhaseGeom &g, int ops);
// twiddle offset (complex entries) of dimension d inside tw
long twiddle_offset(const PhaseGeom &g, int d);

// Last-axis contraction of a partial phase sum for one k-point:
//   X(k)[e] = sum_r twL[kphase[k].y + r] Y[(kphase[k].x + r) M + op nn + e]
// with kphase[k] = (rest(k) W_L, i_L(k) W_L).
struct PhaseLast {
    const double *Y;
    const double *twL;     // [Nk_L][W_L]
    const int *kphase;     // int2 per k
    int WL;
};

// ---- batched forms: several instances (own grid, pulse, buffers) per launch
struct TwiddleEntry {
    PhaseGeom g;
    StageInput in;
    int has_a;
    int use_shift;
    double *tw;
    ModelMerge mm;
};
// grid: ceil(max_count / 256) x m, max_count = max_j (sum_d Nk_d W_d + merged entries)
hipError_t launch_twiddles_batched(const TwiddleEntry *d_entries, int m, int max_count, hipStream_t s);

// ---- the phase sums of the RHS evaluations of one Dormand-Prince attempt,
// formed before its first stage runs: one launch for every twiddle set and
// merged model box, one per contraction pass.  The entries travel as kernel
// arguments (no upload).
constexpr int SET_MAX = 6;
// Special-slot (A) values of stage derivatives that a stage input combines but
// that are not computed when the set is formed: slot e of derivative j of the
// input is v[j][e] where bit j of mask is set (the value the stage kernel will
// store there, -E_e(t_j)), else it is read from memory.
struct AOverride {
    double v[6][3];
    int mask;
};
struct PrepSet {
    TwiddleEntry e[SET_MAX];
    AOverride a[SET_MAX];
    int m, max_count;
};
hipError_t launch_prep_set(const PrepSet &ps, hipStream_t s);
struct ContractSet {
    ContractStage st[SET_MAX];
    long block0[SET_MAX + 1];  // entry j owns blocks [block0[j], block0[j+1])
    int m;
};
hipError_t launch_contract_set(const ContractSet &cs, hipStream_t s);
// entry j owns blocks [block0[j], block0[j] + blocks_x * blocks_y); blocks = total
hipError_t launch_contract_batched(const ContractStage *d_stages, const long *d_block0, int m, long blocks,
                                   hipStream_t s);

// ---- fused RHS stage (stage_kernels.hip) ---------------------------------
struct StageArgs {
    StageInput in;         // stage input state (reference layout), combined on load
    double *xout;          // if set, the combined stage input is also stored here
    double *dxdt;          // output state derivative
    PhaseLast ph;          // {H0, H_E} partial phase sum (ops == 2)
    long nk;
    int n;
    int P;                 // padded matrix stride in the state
    int dim;
    int sweeps;
    double F2;
    double w;              // soothing width
    double E[3];
    int dephase;           // F2 > 0
    double *ucache;        // [k][n*n] eigenvectors of the previous call (or null)
    int warm;              // 1: start the Jacobi from ucache instead of the identity
    // error norm of the attempt fused into its last stage: if err_partial is
    // set, in.k[0..4] are k1, k3, k4, k5, k6 and dxdt is k7; block b stores
    // max over its entries (block 0: and the special slots) of err_entry with
    // coefficients err_e, err_ea, err_er, err_dt into err_partial[b]
    double *err_partial;
    double err_e[6];
    double err_ea, err_er, err_dt;
};
// N: padded band count (2, 4, 8, 16); LR: lanes per matrix row (layout variant)
hipError_t launch_rhs_stage(int N, int LR, const StageArgs &a, hipStream_t s);
// The stages of one Dormand-Prince attempt in one launch, for grids too small
// to fill the GPU (launch gaps dominate there): block b runs st[0 .. m) in
// order on its own k-points.  Stage q reads only what the earlier stages of
// the same block wrote (its k-points' derivatives and eigenvector cache; in
// block 0 also the special slots); phase sums and A slots are formed before
// the launch, so no grid-wide barrier is needed.
struct StageSet {
    StageArgs st[SET_MAX];
    int m;
    // optional: the last stage's per-block error maxima are reduced by the
    // block that finishes last (a ticket counter, zero between launches) and
    // the norm is stored into err_out (mapped host memory)
    unsigned *ticket;
    unsigned long long *err_out;
};
hipError_t launch_rhs_stage_set(int N, int LR, const StageSet &ss, hipStream_t s);
// LR values compiled for N (returns the count written to lr_out)
int stage_variants(int N, int *lr_out, int cap);
// k-points per block (one wavefront) of layout (N, LR)
int stage_k_per_block(int N, int LR);
// entry j owns the logical blocks [block0[j], block0[j+1]) (ceil(nk_j / k_per_block) each)
hipError_t launch_rhs_stage_batched(int N, int LR, const StageArgs *d_args, const long *d_block0, int m, long blocks,
                                    hipStream_t s);

struct InitArgs {
    const double *mesh;    // [k][ops][n*n], H0 = op 0
    int ops;
    double *state;         // receives rho (reference layout), or null
    double *eig;           // receives eigenvalues [k][n], or null
    double *U;             // receives eigenvector rows U [k][n][n], or null
    long nk;
    int n;
    int P;
    int sweeps;
    double mu;
    double kT;
};
hipError_t launch_init_state(int N, int LR, const InitArgs &a, hipStream_t s);

// ---- RK stage algebra (rk_kernels.hip) -----------------------------------
// err_max = max_i |sum_j e[j] k[j]_i| / (ea + er (|x_i| + dt |k0_i|)) (err_entry):
// a single instance's last stage kernel leaves per-block maxima (StageArgs::
// err_partial); this reduces them and stores the bit pattern of the
// non-negative max into *out (device memory or the device alias of mapped
// host memory) with a system-scope store.
hipError_t launch_error_final(const double *partial, long nb, unsigned long long *out, hipStream_t s);

// the same for the instances of a batched last stage: out[j] = max of
// partial[block0[j] .. block0[j+1]) (device arrays; the last range ends at blocks)
hipError_t launch_error_final_batched(const long *d_block0, int m, long blocks, const double *partial, double *out,
                                     hipStream_t s);

// ---- observer (observer_kernels.hip) -------------------------------------
struct ObsArgs {
    const double *x;       // state
    PhaseLast ph;          // observer partial phase sum (ops == 1 + 2 dim)
    long nk;
    int n;
    int P;
    int dim;
    double *partial;       // [nk][pstride], j_k in entries 0..2 of each row
    double *j_out;         // 3 doubles (normalised by nk)
    int pstride;           // row stride of partial (3 for the current alone)
};
hipError_t launch_observe_j(const ObsArgs &a, hipStream_t s);
// entry j owns blocks [block0[j], block0[j] + nk_j) (one block per k-point)
hipError_t launch_observe_j_batched(const ObsArgs *d_args, const long *d_block0, int m, long blocks, hipStream_t s);
// the per-k current alone: partial[k * pstride + d] = j_k,d (d < 3), no reduction
hipError_t launch_observe_jk(const ObsArgs &a, hipStream_t s);

// ---- full observer: occupations and k-space regions ----------------------
// Per k-point Hamiltonian and Wannier occupations (stage_kernels.hip):
// eigen-decomposition of H0(k + A) (op 0 of the observer partial phase sum),
// occupations (U rho U^dagger)_bb in ascending energy, smeared over energy
// with exp(-((e_a - e_b)/smear)^2) normalised over b when smear > 0;
//   X[k Q + 3 + s] = smeared occupation of band s,  X[k Q + 3 + n + s] = Re rho_ss.
struct OccArgs {
    const double *x;       // state
    PhaseLast ph;          // observer partial phase sum (ops == 1 + 2 dim)
    int M;                 // (1 + 2 dim) n n
    long nk;
    int n;
    int P;
    int sweeps;
    double smear;
    double *X;
    int Q;                 // row stride of X: 3 + 2 n
};
hipError_t launch_occupations(int N, int LR, const OccArgs &a, hipStream_t s);

// Region-weighted sums over k (observer_kernels.hip), the reference's
// ObserverCPU::calcWeight / operator() / normalizeExpectationValues
// (code/src/ObserverCPU.cpp:137-147,149-260,293-308).  Quantity q of row k:
// q < Q: X[k Q + q];  q == Q: 1 (the weight sum);  Q < q < Q + 1 + 2 n n: the
// density matrix entries (interleaved doubles of rho_k) when `density`.
struct RegionDev {
    double k[3];
    double inv_sigma2;     // 8 ln 2 / fwhm^2 (0: whole zone)
    int moving;
};
struct RegionArgs {
    const double *x;       // state (A in the special slots, rho_k at 8 + k P)
    const double *X;       // [nk][Q]
    const RegionDev *reg;  // [R]
    int R;
    int Q;
    int n, P;
    int density;
    long nk;
    KOrderH ko;            // k index -> grid indices
    int Nk[3];
    double shift[3];       // fractional k_global_shift
    double lat[9];         // lattice rows a_d (bohr): frac(A)_d = A . a_d / 2 pi
    int dim;
    int chunks;            // k-chunks (grid.y)
    long chunk;            // k-points per chunk
    double *part;          // [R][chunks][Qtot] weighted partial sums
    double *pmin, *pmax;   // [chunks][n] partial min / max over k of X[. + 3 + s]
    double *rec;           // output record: A[3], [R][rstride] normalised (q == Q: raw weight sum), min[n], max[n]
    int rstride;           // >= Qtot: row stride of the record
};
hipError_t launch_region_sums(const RegionArgs &a, hipStream_t s);

// ---- resident integrator (resident_kernels.hip) ---------------------------
// The whole integrate_const of a 1-D single instance in one cooperative
// launch: each block owns the stage kernel's k-points of its index, runs the
// six stages of every attempt on them, and every block evaluates the same
// step-size control from the same error norm (a grid-wide barrier after the
// phase-sum preparation and after the error terms).  Output times as the host
// controller: j(t) at t0 + i dt_out into j_out, the fresh derivative at each
// interval start, steps truncated onto the output grid.
struct SubPulse;
struct ResidentArgs {
    PhaseGeom g;
    const double *Hb, *Db;     // dense model box [W][nn], position operator [3][W][nn]
    long nbox;
    int n, nn, P, dim;
    long nk;
    double *x, *xn;            // state and candidate (swapped on acceptance)
    double *k[8];              // k[1] .. k[7] (k[1], k[7] swapped on acceptance: FSAL)
    double *tw6;               // SET_MAX twiddle sets (each block writes the rows of its k-points)
    double *Yb;                // merged boxes, one SET_MAX group per block: [block][SET_MAX][Y_stride]
    long tw_stride, Y_stride;
    const int *kphase;
    int WL;
    long twL_off;
    double *ucache;
    double *err_partial;       // [blocks]
    const double *Yobs;        // observer operators [W][(1 + 2 dim) nn]
    double *jpartial;          // [nk][3]
    double *j_out, *t_out;     // [output_count][3], [output_count]
    const SubPulse *pulses;
    int npulse;
    double F2, w, eps_abs, eps_rel;
    int sweeps, cold_sweeps, cold_always;
    double t0, t1;
    int output_count;
    long max_attempts;
    long *status;              // [0] accepted steps, [1] RHS evaluations, [2] attempts, [3] code, [4] x / xn parity
};
// status[3]: 0 done, 1 NaN error norm, 2 no step size within 500 tries,
// 3 attempt limit reached.  Returns hipErrorCooperativeLaunchTooLarge when
// the grid cannot be co-resident.
hipError_t launch_resident(int N, int LR, const ResidentArgs &r, hipStream_t s);
int resident_max_blocks(int N, int LR);

}  // namespace sbe
