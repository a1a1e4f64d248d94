/*
 * Plain-C data types shared by the HIP library (libsbewh_hip.so) and the CPU
 * oracle.  Every quantity is in Hartree atomic units; complex numbers are
 * interleaved (re, im) doubles.  No torch or HIP types appear here.
 *
 * The structs carry what the reference keeps in Parameter_t and
 * TightBindingParameter_t for the propagation path:
 *   - sbe_model   <- TightBindingParameter_t::realSpaceOps / latticeVectors
 *                    (reference: code/src/io/TightBindingParameter.h,
 *                     code/src/util/w90util.h W90_realSpaceOperators)
 *   - sbe_grid    <- PropagationParameter_t::Nk, kGlobalShift,
 *                    FourierTransformParameter_t::Nf, GeneralParameter dim
 *                    (code/src/Parameter.cpp:4-56)
 *   - sbe_params  <- PropagationParameter_t (relaxationF2, soothingWidth,
 *                    epsAbs, epsRel), DiagonalizationParameter_t
 *                    (maxSweepCount), TightBindingParameter_t::gsTemp
 *                    (code/src/io/PropagationParameter.cpp:65-111)
 *   - sbe_subpulse <- one Pulse1D + polarization + centre inside Pulse3D
 *                    (code/src/pulse/Pulse3D.cpp:8-21,
 *                     code/src/io/PulseParameter.cpp:121-162)
 */
#ifndef SBEWH_TYPES_H
#define SBEWH_TYPES_H

#ifdef __cplusplus
extern "C" {
#endif

/* Real-space tight-binding operators.  H[c] is the n x n Hamiltonian block of
 * cell c (row-major, complex), D[c][d] the position-operator block r_d(R_c).
 * cells[c] holds the integer lattice-vector indices (R_0, R_1, R_2). */
typedef struct sbe_model {
    int num_wann;             /* n: number of Wannier functions (bands) */
    int num_cells;            /* N_R */
    const int *cells;         /* [num_cells][3] */
    const double *H;          /* [num_cells][n][n][2] */
    const double *D;          /* [num_cells][3][n][n][2] */
    double lattice[9];        /* row i = lattice vector a_i (bohr) */
} sbe_model;

/* Simulation k-grid.  Nf is the Fourier block size that fixes the k-point
 * order of the state vector (k index = c + Nf2*(b + Nf1*(a + Nf0*block)),
 * see code/src/WhFourierTransform.cpp:303-312).  Nf[d] == 0 selects the
 * reference's automatic rule (extent of the model, rounded up to a product
 * of primes <= 5). */
typedef struct sbe_grid {
    int dim;                  /* 1, 2 or 3 */
    int Nk[3];
    int Nf[3];
    double k_global_shift[3]; /* fractional coordinates */
} sbe_grid;

/* Propagation / relaxation / eigensolver parameters. */
typedef struct sbe_params {
    double relaxation_F2;     /* 1/T2; <= 0 disables dephasing */
    double soothing_width;    /* w; < 0 selects the (a != b) dephasing mask */
    double gs_temp;           /* k_B T of the initial Fermi-Dirac state */
    double eps_abs;           /* RK45 absolute tolerance */
    double eps_rel;           /* RK45 relative tolerance */
    int sweep_count;          /* Jacobi sweeps per diagonalisation (upper bound) */
    int flags;                /* SBE_FLAG_* */
} sbe_params;

/* sbe_params.flags */
enum {
    /* start every Jacobi diagonalisation of the RHS from the identity (the
     * reference's behaviour) instead of from the previous RHS call's
     * eigenvectors at the same k-point */
    SBE_FLAG_COLD_JACOBI = 1,
    /* state vectors of this context in the reference's GPU layout
     * (CuPropagator, code/src/CuPropagator.cu:50-52,111-114): the 8 special
     * complex slots, then rho_k unpadded (n*n complex each, row-major), k in
     * the same Fourier-block order; length 8 + N_k * n * n complex
     * (2 * that many interleaved doubles, the thrust::device_vector<double>
     * GPUstate).  Without the flag: the CPU layout with each rho_k padded to
     * 8 * ceil(n*n / 8) complex (code/src/StateContext.h:7-25).  Both layouts
     * are identical when n*n is a multiple of 8. */
    SBE_FLAG_GPU_LAYOUT = 2,
    /* sbe_integrate without a step log, replay or observer on a 1-D instance
     * of at most 8 bands whose stage blocks can all be resident: the whole
     * integrate_const in one cooperative launch, the step-size control on the
     * device (sbe_integrate_info reports whether it ran).  Same step rules;
     * the device's pow / cos / sin may round differently from the host's, so
     * the step sequence can differ at that level from the host controller's. */
    SBE_FLAG_DEVICE_CONTROLLER = 4
};

enum sbe_pulse_type {
    SBE_PULSE_SIN2 = 0,       /* code/src/pulse/Sin2Pulse1D.cpp */
    SBE_PULSE_GAUSS = 1,      /* code/src/pulse/GaussianPulse1D.cpp */
    SBE_PULSE_PURE_GAUSS = 2, /* code/src/pulse/PureGauss1D.cpp */
    SBE_PULSE_SIN2_RAMP = 3   /* code/src/pulse/Sin2RampPulse1D.cpp */
};

/* One linearly polarised sub-pulse, already in atomic units:
 * E0 (field amplitude), omega (carrier), fwhm (envelope parameter; for sin2
 * the field vanishes outside |t| > fwhm), t_start_fwhm / t_end_fwhm (window in
 * units of fwhm), center (time shift).  pol need not be normalised. */
typedef struct sbe_subpulse {
    int type;
    int cycles_on;            /* sin2Ramp only */
    double E0;
    double omega;
    double cep;
    double fwhm;
    double t_start_fwhm;
    double t_end_fwhm;
    double center;
    double rising_cycles;     /* sin2Ramp only */
    double pol[3];
} sbe_subpulse;

/* One instance of a batched scan (sbe_scan_create): its own k-grid, pulse
 * (a slice of the scan's sub-pulse array) and dephasing rate. */
typedef struct sbe_instance {
    int Nk[3];
    int pulse_offset;
    int num_pulses;
    double relaxation_F2;     /* 1/T2; <= 0 disables dephasing */
} sbe_instance;

/* One sample point of a k-space region of the observer (ObserverRegion_t,
 * code/src/ObserverBase.h:33-37, built by code/src/ObserverBase.cpp:3-17
 * from the KspaceRegions input, code/src/io/KspaceRegionParameter.cpp:78-164).
 * The weight of state k-point k in the region is
 *   w = prod_{d<3} exp(-c_d^2 * 8 ln2 / fwhm^2),  c_d = p_d - round(p_d),
 *   p = k_frac(k) + k_global_shift + (moving_frame ? 0 : frac(A)) - k,
 * with frac(A)_d = A . a_d / (2 pi) (code/src/ObserverCPU.cpp:137-147,
 * code/src/WhFourierTransform.cpp:46-52,163-174).  fwhm = INFINITY gives
 * w = 1 (the default region "expectationValues"). */
typedef struct sbe_region {
    double k[3];              /* centre, fractional coordinates */
    double fwhm;              /* full width at half maximum, fractional */
    int moving_frame;         /* 1: 'M' (centre follows k), 0: 'F' (fixed in the k + A frame) */
} sbe_region;

/* Record of the adaptive controller: one entry per step attempt. */
typedef struct sbe_steplog {
    int capacity;             /* entries the arrays can hold */
    int count;                /* entries written (may exceed capacity: then
                                 only the first `capacity` are stored) */
    double *t;                /* time at the start of the attempt */
    double *dt;               /* step size tried */
    double *err;              /* max relative error of the attempt */
    int *accepted;            /* 1 accepted, 0 rejected */
} sbe_steplog;

#ifdef __cplusplus
}
#endif

#endif
