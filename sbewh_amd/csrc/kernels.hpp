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

// Twiddles tw_d[i*W_d + r] = exp(i 2 pi (i/Nk_d + off_d)(r + rmin_d)), off =
// frac(kshift_au + A).  A is read from Re(a_src[0..dim-1]) if a_src != null,
// otherwise A = 0; use_shift = 0 drops kshift_au.
hipError_t launch_twiddles(const PhaseGeom &g, const double *a_src, int use_shift, double *tw,
                           hipStream_t s);

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

// ---- fused RHS stage (stage_kernels.hip) ---------------------------------
struct StageArgs {
    const double *x;       // stage input state (reference layout)
    double *dxdt;          // output state derivative
    const double *mesh;    // [k][2][n*n]: H0, HE
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
};
// N: padded band count (2, 4, 8, 16); LR: lanes per matrix row (layout variant)
hipError_t launch_rhs_stage(int N, int LR, const StageArgs &a, hipStream_t s);
// LR values compiled for N (returns the count written to lr_out)
int stage_variants(int N, int *lr_out, int cap);

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
// out = x + sum_{j<nt} c[j] * k[j]   over S complex entries
hipError_t launch_combine(long S, const double *x, int nt, const double *const *k, const double *c,
                          double *out, hipStream_t s);
// err_max = max_i |sum_j e[j] k[j]_i| / (ea + er (|x_i| + dt |k0_i|)), written as the bit
// pattern of a non-negative double into *err_bits; partial holds >= 1024 doubles.
hipError_t launch_error_norm(long S, const double *x, const double *const *k, const double *e, double ea,
                             double er, double dt, double *partial, unsigned long long *err_bits, hipStream_t s);

// ---- observer (observer_kernels.hip) -------------------------------------
struct ObsArgs {
    const double *x;       // state
    const double *mesh;    // [k][1+2dim][n*n]: H0, D_d, dH_d
    long nk;
    int n;
    int P;
    int dim;
    double *partial;       // [nblocks][3]
    double *j_out;         // 3 doubles (normalised by nk)
};
hipError_t launch_observe_j(const ObsArgs &a, hipStream_t s);

}  // namespace sbe
