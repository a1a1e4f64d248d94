/*
 * CPU oracle (test infrastructure only; see sbe_oracle.h).
 *
 * Plain C99 + OpenMP restatement of the reference's propagation path.  Each
 * function cites the reference file:line it follows.  The reference itself
 * cannot be built in this image (it needs a CMake-generated sbewhConfig.h,
 * Boost.Odeint/Math, FFTW3, MPI), so nothing here calls into it.
 */
#include "sbe_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

#define PI 3.14159265358979323846
#define JACOBI_EPS 1e-15 /* code/src/util/DiagonalizatorSingle.h:24 */

static _Thread_local char g_err[512];
static void set_err(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}
const char *oracle_last_error(void) { return g_err; }

struct oracle_ctx {
    int n, ncell, dim;
    int *cells;          /* projected cells [ncell][3] */
    cplx *H;             /* [ncell][n*n] */
    cplx *D;             /* [ncell][3][n*n] */
    double lat[9];
    int Nk[3], Nf[3], Nrep[3];
    long nk;
    int P;               /* padded matrix size (StateIndexMapper) */
    int nspecial;        /* special index count (8) */
    long state_size;
    double kshift_au[3];
    sbe_params prm;
    sbe_subpulse *pulses;
    int npulses;
    double pol[8][3];
    int phase_method;
    /* separable phase-sum box */
    int rmin[3], W[3];
    /* per-k mesh buffer [nk][2][n*n] */
    cplx *mesh;
};

/* ------------------------------------------------------------------------ */
/* geometry: code/src/StateContext.cpp:3-9 (kFracTo_au),
 *           code/src/WhFourierTransform.cpp:46-52 (toFracK)                  */
static void cross3(const double *a, const double *b, double *r)
{
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}
static double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

static void kfrac_to_au(const double *lat, const double *kf, double *out)
{
    double c12[3], c20[3], c01[3];
    cross3(lat + 3, lat + 6, c12);
    cross3(lat + 6, lat + 0, c20);
    cross3(lat + 0, lat + 3, c01);
    double pre = 2 * PI / dot3(lat, c12);
    for (int i = 0; i < 3; i++)
        out[i] = pre * (kf[0] * c12[i] + kf[1] * c20[i] + kf[2] * c01[i]);
}
static void to_frac_k(const double *lat, const double *k, double *out)
{
    for (int i = 0; i < 3; i++)
        out[i] = dot3(k, lat + 3 * i) / (2 * PI);
}

/* ------------------------------------------------------------------------ */
/* pulses: code/src/pulse/Pulse3D.cpp:8-62, Sin2Pulse1D.cpp:3-34,
 *         GaussianPulse1D.cpp:3-36, PureGauss1D.cpp:3-35,
 *         Sin2RampPulse1D.cpp:3-74                                           */
static double sub_E(const sbe_subpulse *p, double t)
{
    double w = p->omega, F = p->fwhm, E0 = p->E0;
    switch (p->type) {
    case SBE_PULSE_SIN2: {
        if (t < -F || t > F) return 0;
        double beta = PI / 2 / F;
        double cp = w * t + p->cep, ec = cos(beta * t);
        return E0 * (cos(cp) * ec * ec - 2 * beta / w * sin(cp) * ec * sin(beta * t));
    }
    case SBE_PULSE_GAUSS: {
        double beta = 4 * log(2.0) / (F * F);
        double cp = w * t + p->cep;
        return E0 * (-2 * beta / w * t * sin(cp) + cos(cp)) * exp(-beta * t * t);
    }
    case SBE_PULSE_PURE_GAUSS: {
        double beta = 4 * log(2.0) / (F * F);
        return E0 * exp(-beta * t * t);
    }
    case SBE_PULSE_SIN2_RAMP: {
        double rc = p->rising_cycles;
        int co = p->cycles_on;
        double beta = w / 4 / rc;
        double rising = 2 * PI / w * rc, on = PI / w * co, half = rising + PI / w * co;
        if (t < -half) return 0;
        if (t < -on) {
            t += on;
            double cp = w * t, ec = cos(beta * t);
            return E0 * (cos(cp) * ec * ec - 2 * beta / w * sin(cp) * ec * sin(beta * t));
        }
        if (t < on) return ((co % 2 == 0) ? 1 : -1) * E0 * cos(w * t);
        if (t < half) {
            t -= on;
            double cp = w * t, ec = cos(beta * t);
            return E0 * (cos(cp) * ec * ec - 2 * beta / w * sin(cp) * ec * sin(beta * t));
        }
        return 0;
    }
    }
    return 0;
}

static double sub_start(const sbe_subpulse *p)
{
    if (p->type == SBE_PULSE_SIN2_RAMP)
        return -2 * PI / p->omega * (p->rising_cycles + 0.5 * p->cycles_on);
    return p->t_start_fwhm * p->fwhm;
}
static double sub_end(const sbe_subpulse *p)
{
    if (p->type == SBE_PULSE_SIN2_RAMP)
        return 2 * PI / p->omega * (p->rising_cycles + 0.5 * p->cycles_on);
    return p->t_end_fwhm * p->fwhm;
}

static double sub_A(const sbe_subpulse *p, double t)
{
    double w = p->omega, F = p->fwhm, E0 = p->E0;
    switch (p->type) {
    case SBE_PULSE_SIN2: {
        if (t < -F || t > F) return 0;
        double beta = PI / 2 / F, ec = cos(beta * t);
        return -E0 / w * ec * ec * sin(w * t + p->cep);
    }
    case SBE_PULSE_GAUSS: {
        double beta = 4 * log(2.0) / (F * F);
        return -E0 / w * exp(-beta * t * t) * sin(w * t + p->cep);
    }
    case SBE_PULSE_PURE_GAUSS: {
        double beta = 4 * log(2.0) / (F * F);
        double apre = E0 * sqrt(PI / beta) / 2;
        double erf_start = apre * erf(sub_start(p)); /* as written in PureGauss1D.cpp:9 */
        return -(apre * erf(t * sqrt(beta)) - erf_start);
    }
    case SBE_PULSE_SIN2_RAMP: {
        double rc = p->rising_cycles;
        int co = p->cycles_on;
        double beta = w / 4 / rc;
        double rising = 2 * PI / w * rc, on = PI / w * co, half = rising + PI / w * co;
        if (t < -half) return 0;
        if (t < -on) {
            t += on;
            double ec = cos(beta * t);
            return -E0 / w * ec * ec * sin(w * t);
        }
        if (t < on) return ((co % 2 == 0) ? 1 : -1) * -E0 / w * sin(w * t);
        if (t < half) {
            t -= on;
            double ec = cos(beta * t);
            return -E0 / w * ec * ec * sin(w * t);
        }
        return 0;
    }
    }
    return 0;
}

void oracle_pulse_E(const oracle_ctx *c, double t, double E[3])
{
    E[0] = E[1] = E[2] = 0;
    for (int p = 0; p < c->npulses; p++) {
        double e = sub_E(&c->pulses[p], t - c->pulses[p].center);
        for (int d = 0; d < 3; d++) E[d] += e * c->pol[p][d];
    }
}
void oracle_pulse_A(const oracle_ctx *c, double t, double A[3])
{
    A[0] = A[1] = A[2] = 0;
    for (int p = 0; p < c->npulses; p++) {
        double a = sub_A(&c->pulses[p], t - c->pulses[p].center);
        for (int d = 0; d < 3; d++) A[d] += a * c->pol[p][d];
    }
}
double oracle_start_time(const oracle_ctx *c)
{
    double s = INFINITY;
    for (int p = 0; p < c->npulses; p++) {
        double t = sub_start(&c->pulses[p]) + c->pulses[p].center;
        if (t < s) s = t;
    }
    return s;
}
double oracle_end_time(const oracle_ctx *c)
{
    double s = -INFINITY;
    for (int p = 0; p < c->npulses; p++) {
        double t = sub_end(&c->pulses[p]) + c->pulses[p].center;
        if (t > s) s = t;
    }
    return s;
}

int oracle_set_pulses(oracle_ctx *c, const sbe_subpulse *pulses, int np)
{
    if (np < 1 || np > 8) {
        set_err("number of sub-pulses must be in [1, 8] (got %d)", np);
        return -1;
    }
    free(c->pulses);
    c->pulses = malloc(sizeof(sbe_subpulse) * np);
    memcpy(c->pulses, pulses, sizeof(sbe_subpulse) * np);
    c->npulses = np;
    for (int p = 0; p < np; p++) {
        /* GeomVector::normalize (code/src/pulse/Pulse3D.cpp:10-11) */
        double nrm = sqrt(dot3(pulses[p].pol, pulses[p].pol));
        for (int d = 0; d < 3; d++) c->pol[p][d] = pulses[p].pol[d] / nrm;
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* grid rules: code/src/Parameter.cpp:4-56, code/src/io/TightBindingParameter.cpp:254-268 */
static int smooth5(int N)
{
    int n = N - 1, r;
    do {
        r = ++n;
        int pr[3] = {2, 3, 5};
        for (int i = 0; i < 3; i++)
            while (r % pr[i] == 0) r /= pr[i];
    } while (r != 1);
    return n;
}

oracle_ctx *oracle_create(const sbe_model *m, const sbe_grid *g, const sbe_params *prm,
                          const sbe_subpulse *pulses, int np)
{
    if (m->num_wann < 1 || m->num_cells < 1 || g->dim < 1 || g->dim > 3) {
        set_err("invalid model/grid");
        return NULL;
    }
    oracle_ctx *c = calloc(1, sizeof *c);
    int n = c->n = m->num_wann;
    c->dim = g->dim;
    memcpy(c->lat, m->lattice, sizeof c->lat);
    c->prm = *prm;
    c->phase_method = ORACLE_PHASE_SEPARABLE;
    /* projection onto dim dimensions: code/src/util/w90util.cpp:417-444 */
    c->cells = calloc((size_t)m->num_cells * 3, sizeof(int));
    c->H = calloc((size_t)m->num_cells * n * n, sizeof(cplx));
    c->D = calloc((size_t)m->num_cells * 3 * n * n, sizeof(cplx));
    int nc = 0;
    for (int i = 0; i < m->num_cells; i++) {
        int r[3];
        for (int d = 0; d < 3; d++) r[d] = d < g->dim ? m->cells[3 * i + d] : 0;
        int found = -1;
        for (int j = 0; j < nc; j++)
            if (c->cells[3 * j] == r[0] && c->cells[3 * j + 1] == r[1] && c->cells[3 * j + 2] == r[2]) {
                found = j;
                break;
            }
        if (found < 0) {
            found = nc++;
            memcpy(c->cells + 3 * found, r, sizeof r);
        }
        for (int e = 0; e < n * n; e++) {
            c->H[(size_t)found * n * n + e] += m->H[2 * ((size_t)i * n * n + e)] + I * m->H[2 * ((size_t)i * n * n + e) + 1];
            for (int d = 0; d < 3; d++) {
                size_t src = ((size_t)i * 3 + d) * n * n + e;
                c->D[((size_t)found * 3 + d) * n * n + e] += m->D[2 * src] + I * m->D[2 * src + 1];
            }
        }
    }
    c->ncell = nc;
    int hasOrigin = 0;
    for (int j = 0; j < nc; j++)
        if (!c->cells[3 * j] && !c->cells[3 * j + 1] && !c->cells[3 * j + 2]) hasOrigin = 1;
    if (!hasOrigin) {
        set_err("projected hamiltonian does not contain R=(0,0,0)");
        oracle_destroy(c);
        return NULL;
    }
    int mn[3] = {0, 0, 0}, mx[3] = {0, 0, 0};
    for (int j = 0; j < nc; j++)
        for (int d = 0; d < 3; d++) {
            if (c->cells[3 * j + d] < mn[d]) mn[d] = c->cells[3 * j + d];
            if (c->cells[3 * j + d] > mx[d]) mx[d] = c->cells[3 * j + d];
        }
    for (int d = 0; d < 3; d++) {
        c->rmin[d] = mn[d];
        c->W[d] = mx[d] - mn[d] + 1;
    }
    /* Nf rule */
    int Nf[3];
    for (int d = 0; d < 3; d++) Nf[d] = d < g->dim ? (g->Nf[d] > 0 ? g->Nf[d] : 1) : 1;
    for (int d = 0; d < g->dim; d++)
        if (Nf[d] < c->W[d]) Nf[d] = c->W[d];
    int maxNf = Nf[0] > Nf[1] ? Nf[0] : Nf[1];
    if (Nf[2] > maxNf) maxNf = Nf[2];
    if (maxNf > 5)
        for (int d = 0; d < g->dim; d++) Nf[d] = smooth5(Nf[d]);
    for (int d = 0; d < 3; d++) {
        c->Nk[d] = d < g->dim ? g->Nk[d] : 1;
        if (c->Nk[d] < 1 || c->Nk[d] % Nf[d] != 0) {
            set_err("In k-dimension %d: Nk (%d) is not a multiple of Nf (%d)", d + 1, c->Nk[d], Nf[d]);
            oracle_destroy(c);
            return NULL;
        }
        c->Nf[d] = Nf[d];
        c->Nrep[d] = c->Nk[d] / Nf[d];
    }
    c->nk = (long)c->Nk[0] * c->Nk[1] * c->Nk[2];
    /* StateIndexMapper: code/src/StateContext.h:7-25 */
    c->P = 8 * ((n * n + 7) / 8);
    c->nspecial = 8;
    c->state_size = c->nspecial + (long)c->P * c->nk;
    kfrac_to_au(c->lat, g->k_global_shift, c->kshift_au);
    if (oracle_set_pulses(c, pulses, np)) {
        oracle_destroy(c);
        return NULL;
    }
    c->mesh = malloc(sizeof(cplx) * 2 * n * n * (size_t)c->nk);
    return c;
}

void oracle_destroy(oracle_ctx *c)
{
    if (!c) return;
    free(c->cells);
    free(c->H);
    free(c->D);
    free(c->pulses);
    free(c->mesh);
    free(c);
}

void oracle_set_phase_method(oracle_ctx *c, int method) { c->phase_method = method; }
long oracle_state_size(const oracle_ctx *c) { return c->state_size; }
long oracle_num_k(const oracle_ctx *c) { return c->nk; }
int oracle_padded_mat_size(const oracle_ctx *c) { return c->P; }
void oracle_get_nf(const oracle_ctx *c, int nf[3])
{
    for (int d = 0; d < 3; d++) nf[d] = c->Nf[d];
}

/* k index -> global grid indices, following the FFT block order
 * (code/src/WhFourierTransform.cpp:137-149 for the block shifts,
 *  :303-312 for the in-block index) */
static void k_index(const oracle_ctx *c, long k, int i[3])
{
    long fft = (long)c->Nf[0] * c->Nf[1] * c->Nf[2];
    long blk = k / fft, f = k % fft;
    int a = (int)(f / ((long)c->Nf[1] * c->Nf[2])), b = (int)((f / c->Nf[2]) % c->Nf[1]), cc = (int)(f % c->Nf[2]);
    int ar = (int)(blk / ((long)c->Nrep[1] * c->Nrep[2])), br = (int)((blk / c->Nrep[2]) % c->Nrep[1]),
        cr = (int)(blk % c->Nrep[2]);
    i[0] = a * c->Nrep[0] + ar;
    i[1] = b * c->Nrep[1] + br;
    i[2] = cc * c->Nrep[2] + cr;
}

void oracle_k_frac(const oracle_ctx *c, long k, double kf[3])
{
    int i[3];
    k_index(c, k, i);
    for (int d = 0; d < 3; d++) kf[d] = (double)i[d] / c->Nk[d];
}

/* ------------------------------------------------------------------------ */
/* phase sums: direct form code/src/WhFourierTransform.cpp:54-83 */
void oracle_hk_point(const oracle_ctx *c, const double kf[3], const double E[3], double *H0o, double *HEo)
{
    int n = c->n;
    cplx *H0 = (cplx *)H0o, *HE = (cplx *)HEo;
    for (int e = 0; e < n * n; e++) H0[e] = HE[e] = 0;
    for (int ci = 0; ci < c->ncell; ci++) {
        const int *r = c->cells + 3 * ci;
        double ph = 2 * PI * (kf[0] * r[0] + kf[1] * r[1] + kf[2] * r[2]);
        cplx eikr = cos(ph) + I * sin(ph);
        for (int e = 0; e < n * n; e++) {
            cplx h = c->H[(size_t)ci * n * n + e];
            H0[e] += h * eikr;
            cplx rhe = h;
            for (int d = 0; d < c->dim; d++) rhe += c->D[((size_t)ci * 3 + d) * n * n + e] * E[d];
            HE[e] += rhe * eikr;
        }
    }
}

/* full operators for the observer: H0, D_d and dH0/dk_d = sum_R i R_cart,d H(R) e^{..}
 * (code/src/WhFourierTransform.cpp:17-26, :71-83, :317-368) */
void oracle_hk_full_point(const oracle_ctx *c, const double kf[3], double *H0o, double *Do, double *dHo)
{
    int n = c->n;
    cplx *H0 = (cplx *)H0o, *D = (cplx *)Do, *dH = (cplx *)dHo;
    memset(H0, 0, sizeof(cplx) * n * n);
    memset(D, 0, sizeof(cplx) * 3 * n * n);
    memset(dH, 0, sizeof(cplx) * 3 * n * n);
    for (int ci = 0; ci < c->ncell; ci++) {
        const int *r = c->cells + 3 * ci;
        double pos[3];
        for (int d = 0; d < 3; d++) pos[d] = r[0] * c->lat[d] + r[1] * c->lat[3 + d] + r[2] * c->lat[6 + d];
        double ph = 2 * PI * (kf[0] * r[0] + kf[1] * r[1] + kf[2] * r[2]);
        cplx eikr = cos(ph) + I * sin(ph);
        for (int e = 0; e < n * n; e++) {
            cplx h = c->H[(size_t)ci * n * n + e];
            H0[e] += h * eikr;
            for (int d = 0; d < c->dim; d++) {
                D[d * n * n + e] += c->D[((size_t)ci * 3 + d) * n * n + e] * eikr;
                dH[d * n * n + e] += I * pos[d] * h * eikr;
            }
        }
    }
}

/* Separable form of the same sum on the dense R box: the phase factorises as
 * prod_d exp(i 2 pi (i_d/Nk_d + off_d) R_d); contract one axis at a time.
 * Mathematically identical to the reference's pre-phase + FFT + post-phase
 * (code/src/WhFourierTransform.cpp:259-315). */
static void mesh_separable(oracle_ctx *c, const double off[3], const double E[3])
{
    int n = c->n, nn = n * n, M = 2 * nn;
    int W0 = c->W[0], W1 = c->W[1], W2 = c->W[2];
    int N0 = c->Nk[0], N1 = c->Nk[1], N2 = c->Nk[2];
    cplx *tw[3];
    int Wd[3] = {W0, W1, W2}, Nd[3] = {N0, N1, N2};
    for (int d = 0; d < 3; d++) {
        tw[d] = malloc(sizeof(cplx) * Nd[d] * Wd[d]);
        for (int i = 0; i < Nd[d]; i++)
            for (int r = 0; r < Wd[d]; r++) {
                double ph = 2 * PI * (((double)i / Nd[d] + off[d]) * (r + c->rmin[d]));
                tw[d][i * Wd[d] + r] = cos(ph) + I * sin(ph);
            }
    }
    /* X[R0][R1][R2][m], m = op*nn + e */
    size_t szX = (size_t)W0 * W1 * W2 * M;
    cplx *X = calloc(szX, sizeof(cplx));
    for (int ci = 0; ci < c->ncell; ci++) {
        const int *r = c->cells + 3 * ci;
        size_t base = ((((size_t)(r[0] - c->rmin[0]) * W1) + (r[1] - c->rmin[1])) * W2 + (r[2] - c->rmin[2])) * M;
        for (int e = 0; e < nn; e++) {
            cplx h = c->H[(size_t)ci * nn + e];
            cplx he = h;
            for (int d = 0; d < c->dim; d++) he += c->D[((size_t)ci * 3 + d) * nn + e] * E[d];
            X[base + e] = h;
            X[base + nn + e] = he;
        }
    }
    /* stage over R2 -> i2 */
    cplx *Y2 = calloc((size_t)W0 * W1 * N2 * M, sizeof(cplx));
#pragma omp parallel for collapse(2) schedule(static)
    for (int a = 0; a < W0 * W1; a++)
        for (int i2 = 0; i2 < N2; i2++) {
            cplx *o = Y2 + ((size_t)a * N2 + i2) * M;
            for (int r2 = 0; r2 < W2; r2++) {
                cplx t = tw[2][i2 * W2 + r2];
                const cplx *x = X + ((size_t)a * W2 + r2) * M;
                for (int m = 0; m < M; m++) o[m] += t * x[m];
            }
        }
    free(X);
    /* stage over R1 -> i1 */
    cplx *Y1 = calloc((size_t)W0 * N1 * N2 * M, sizeof(cplx));
#pragma omp parallel for collapse(2) schedule(static)
    for (int a = 0; a < W0; a++)
        for (int i1 = 0; i1 < N1; i1++) {
            for (int r1 = 0; r1 < W1; r1++) {
                cplx t = tw[1][i1 * W1 + r1];
                for (int i2 = 0; i2 < N2; i2++) {
                    cplx *o = Y1 + (((size_t)a * N1 + i1) * N2 + i2) * M;
                    const cplx *x = Y2 + (((size_t)a * W1 + r1) * N2 + i2) * M;
                    for (int m = 0; m < M; m++) o[m] += t * x[m];
                }
            }
        }
    free(Y2);
    /* stage over R0 -> i0, scattered into state k-order */
    long fft = (long)c->Nf[0] * c->Nf[1] * c->Nf[2];
    (void)fft;
#pragma omp parallel for schedule(static)
    for (long k = 0; k < c->nk; k++) {
        int ii[3];
        k_index(c, k, ii);
        cplx *o = c->mesh + (size_t)k * M;
        for (int m = 0; m < M; m++) o[m] = 0;
        for (int r0 = 0; r0 < W0; r0++) {
            cplx t = tw[0][ii[0] * W0 + r0];
            const cplx *x = Y1 + (((size_t)r0 * N1 + ii[1]) * N2 + ii[2]) * M;
            for (int m = 0; m < M; m++) o[m] += t * x[m];
        }
    }
    free(Y1);
    for (int d = 0; d < 3; d++) free(tw[d]);
}

static void mesh_compute(oracle_ctx *c, const double A[3], const double E[3], int with_shift)
{
    double kOff[3], off[3];
    for (int d = 0; d < 3; d++) kOff[d] = (with_shift ? c->kshift_au[d] : 0) + A[d];
    to_frac_k(c->lat, kOff, off);
    int nn = c->n * c->n;
    if (c->phase_method == ORACLE_PHASE_SEPARABLE) {
        mesh_separable(c, off, E);
        return;
    }
#pragma omp parallel for schedule(static)
    for (long k = 0; k < c->nk; k++) {
        double kf[3];
        int ii[3];
        k_index(c, k, ii);
        for (int d = 0; d < 3; d++) kf[d] = (double)ii[d] / c->Nk[d] + off[d];
        oracle_hk_point(c, kf, E, (double *)(c->mesh + (size_t)k * 2 * nn), (double *)(c->mesh + (size_t)k * 2 * nn + nn));
    }
}

void oracle_hk_mesh(oracle_ctx *c, const double A[3], const double E[3], double *out)
{
    mesh_compute(c, A, E, 1);
    memcpy(out, c->mesh, sizeof(cplx) * 2 * c->n * c->n * (size_t)c->nk);
}

/* ------------------------------------------------------------------------ */
/* cyclic complex Jacobi: code/src/util/DiagonalizatorSingle.cpp:57-122 */
static void cs_half(double y, double x, double *c, double *s)
{
    double t = y / (fabs(x) + sqrt(x * x + y * y));
    double cc = 1 / sqrt(1 + t * t);
    double ss = t * cc;
    if (x >= 0) {
        *c = cc;
        *s = ss;
    } else {
        *c = fabs(ss);
        *s = copysign(cc, y);
    }
}

static void jacobi_rotate(int n, int p, int q, cplx *A, cplx *U)
{
    cplx Apq = A[p * n + q];
    double c1, s1, c2, s2;
    cs_half(-creal(Apq), cimag(Apq), &c1, &s1);
    cs_half(2 * cabs(Apq), creal(A[p * (n + 1)]) - creal(A[q * (n + 1)]), &c2, &s2);
    cplx Rpp = (-s1 * s2 - 1) + I * (-c1 * s2);
    cplx Rpq = (-c1 * c2) + I * (-s1 * c2);
    cplx Rqp = -conj(Rpq);
    cplx Rqq = conj(Rpp);
    for (int i = 0; i < n; i++) {
        cplx dUp = Rpp * U[p * n + i] + Rpq * U[q * n + i];
        cplx dUq = Rqp * U[p * n + i] + Rqq * U[q * n + i];
        U[p * n + i] += dUp;
        U[q * n + i] += dUq;
    }
    for (int i = 0; i < n; i++) {
        cplx dApq = Rpp * A[p * n + i] + Rpq * A[q * n + i];
        cplx dAqp = Rqp * A[p * n + i] + Rqq * A[q * n + i];
        A[p * n + i] += dApq;
        A[q * n + i] += dAqp;
    }
    for (int i = 0; i < n; i++) {
        cplx dApq = Rqq * A[i * n + p] - Rqp * A[i * n + q];
        cplx dAqp = -Rpq * A[i * n + p] + Rpp * A[i * n + q];
        A[i * n + p] += dApq;
        A[i * n + q] += dAqp;
    }
    A[p * n + q] = 0;
    A[q * n + p] = 0;
}

static void jacobi_work(int n, int sweeps, cplx *A, double *eig, cplx *U)
{
    for (int a = 0; a < n; a++)
        for (int b = 0; b < n; b++) U[a * n + b] = (a == b);
    for (int sw = 0; sw < sweeps; sw++) {
        double sum = 0;
        for (int i = 0; i < n; i++)
            for (int j = i; j < n; j++) sum += fabs(creal(A[i * n + j])) + fabs(cimag(A[i * n + j]));
        sum /= 2;
        for (int a = 0; a < n; a++)
            for (int b = a + 1; b < n; b++)
                if (cabs(A[a * n + b]) > JACOBI_EPS * sum) jacobi_rotate(n, a, b, A, U);
    }
    for (int i = 0; i < n; i++) eig[i] = creal(A[i * (n + 1)]);
    for (int i = 0; i < n * n; i++) A[i] = U[i];
}

void oracle_jacobi(int n, int sweeps, double *A, double *eig)
{
    cplx *U = malloc(sizeof(cplx) * n * n);
    jacobi_work(n, sweeps, (cplx *)A, eig, U);
    free(U);
}

/* ------------------------------------------------------------------------ */
/* RHS: code/src/PropagatorCPU.cpp:73-87 (dephaseRhoH), :109-119 (manualCoherent),
 * :121-155 (addManualDephasing), :208-256 (operator())                       */
int oracle_rhs(oracle_ctx *c, const double *state_d, double *dsdt_d, double t)
{
    const cplx *state = (const cplx *)state_d;
    cplx *dsdt = (cplx *)dsdt_d;
    int n = c->n, nn = n * n;
    memset(dsdt, 0, sizeof(cplx) * c->state_size);
    double E[3], A[3] = {0, 0, 0};
    oracle_pulse_E(c, t, E);
    for (int d = 0; d < c->dim; d++) {
        A[d] = creal(state[d]);
        dsdt[d] = -E[d];
    }
    mesh_compute(c, A, E, 1);
    double F2 = c->prm.relaxation_F2, w = c->prm.soothing_width;
#pragma omp parallel
    {
        cplx *U = malloc(sizeof(cplx) * nn), *Uw = malloc(sizeof(cplx) * nn);
        cplx *T = malloc(sizeof(cplx) * nn), *rH = malloc(sizeof(cplx) * nn);
        double *en = malloc(sizeof(double) * n);
#pragma omp for schedule(static)
        for (long k = 0; k < c->nk; k++) {
            const cplx *rho = state + c->nspecial + (size_t)c->P * k;
            cplx *ds = dsdt + c->nspecial + (size_t)c->P * k;
            const cplx *HE = c->mesh + (size_t)k * 2 * nn + nn;
            for (int a = 0; a < n; a++)
                for (int b = 0; b < n; b++) {
                    cplx s = 0;
                    for (int i = 0; i < n; i++) s += HE[a * n + i] * rho[i * n + b] - HE[i * n + b] * rho[a * n + i];
                    ds[a * n + b] = -I * s;
                }
            if (F2 > 0) {
                memcpy(U, c->mesh + (size_t)k * 2 * nn, sizeof(cplx) * nn);
                jacobi_work(n, c->prm.sweep_count, U, en, Uw);
                for (int a = 0; a < n; a++)
                    for (int b = 0; b < n; b++) {
                        cplx s = 0;
                        for (int i = 0; i < n; i++) s += U[a * n + i] * rho[i * n + b];
                        T[a * n + b] = s;
                    }
                for (int a = 0; a < n; a++)
                    for (int b = 0; b < n; b++) {
                        cplx s = 0;
                        for (int i = 0; i < n; i++) s += T[a * n + i] * conj(U[b * n + i]);
                        rH[a * n + b] = s;
                    }
                for (int a = 0; a < n; a++)
                    for (int b = 0; b < n; b++) {
                        if (w < 0) {
                            rH[a * n + b] = -(double)(a != b) * F2 * rH[a * n + b];
                        } else {
                            double dE = (en[a] - en[b]) / w;
                            rH[a * n + b] = -(1 - exp(-dE * dE)) * F2 * rH[a * n + b];
                        }
                    }
                for (int a = 0; a < n; a++)
                    for (int b = 0; b < n; b++) {
                        cplx s = 0;
                        for (int i = 0; i < n; i++) s += conj(U[i * n + a]) * rH[i * n + b];
                        T[a * n + b] = s;
                    }
                for (int a = 0; a < n; a++)
                    for (int b = 0; b < n; b++) {
                        cplx s = 0;
                        for (int i = 0; i < n; i++) s += T[a * n + i] * U[i * n + b];
                        ds[a * n + b] += s;
                    }
            }
        }
        free(U);
        free(Uw);
        free(T);
        free(rH);
        free(en);
    }
    return 0;
}

/* initial state: code/src/PropagatorCPU.cpp:31-70 */
int oracle_initial_state(oracle_ctx *c, double mu, double *state_d)
{
    cplx *st = (cplx *)state_d;
    int n = c->n, nn = n * n;
    memset(st, 0, sizeof(cplx) * c->state_size);
    double A[3] = {0, 0, 0}, E[3] = {0, 0, 0};
    mesh_compute(c, A, E, 1);
    double kT = c->prm.gs_temp;
#pragma omp parallel
    {
        cplx *Uw = malloc(sizeof(cplx) * nn), *U = malloc(sizeof(cplx) * nn);
        double *e = malloc(sizeof(double) * n);
#pragma omp for schedule(static)
        for (long k = 0; k < c->nk; k++) {
            cplx *rho = st + c->nspecial + (size_t)c->P * k;
            memcpy(U, c->mesh + (size_t)k * 2 * nn, sizeof(cplx) * nn);
            jacobi_work(n, c->prm.sweep_count, U, e, Uw);
            for (int i = 0; i < n; i++) e[i] = 1 / (1 + exp((e[i] - mu) / kT));
            for (int a = 0; a < n; a++)
                for (int b = 0; b < n; b++) {
                    cplx s = 0;
                    for (int i = 0; i < n; i++) s += U[i * n + b] * e[i] * conj(U[i * n + a]);
                    rho[a * n + b] = s;
                }
        }
        free(Uw);
        free(U);
        free(e);
    }
    return 0;
}

static int cmp_double(const void *a, const void *b)
{
    double x = *(const double *)a, y = *(const double *)b;
    return (x > y) - (x < y);
}

/* Fermi level: code/src/main/runSBE.cpp:97-155 (no kGlobalShift, A = 0) */
int oracle_fermi_level(oracle_ctx *c, int occupied_bands, int below_gap, double *mu)
{
    int n = c->n, nn = n * n;
    double A[3] = {0, 0, 0}, E[3] = {0, 0, 0};
    mesh_compute(c, A, E, 0);
    double *ev = malloc(sizeof(double) * n * (size_t)c->nk);
#pragma omp parallel
    {
        cplx *Uw = malloc(sizeof(cplx) * nn), *U = malloc(sizeof(cplx) * nn);
#pragma omp for schedule(static)
        for (long k = 0; k < c->nk; k++) {
            memcpy(U, c->mesh + (size_t)k * 2 * nn, sizeof(cplx) * nn);
            jacobi_work(n, c->prm.sweep_count, U, ev + (size_t)k * n, Uw);
        }
        free(Uw);
        free(U);
    }
    double kT = c->prm.gs_temp;
    int rc = 0;
    if (below_gap) {
        double *bmin = malloc(sizeof(double) * n), *bmax = malloc(sizeof(double) * n), *tmp = malloc(sizeof(double) * n);
        for (int i = 0; i < n; i++) {
            bmin[i] = INFINITY;
            bmax[i] = -INFINITY;
        }
        for (long k = 0; k < c->nk; k++) {
            memcpy(tmp, ev + (size_t)k * n, sizeof(double) * n);
            qsort(tmp, n, sizeof(double), cmp_double);
            for (int i = 0; i < n; i++) {
                if (tmp[i] < bmin[i]) bmin[i] = tmp[i];
                if (tmp[i] > bmax[i]) bmax[i] = tmp[i];
            }
        }
        int gaps = 0;
        double f = 0;
        for (int i = 0; i + 1 < n; i++)
            if (bmax[i] + 3 * kT < bmin[i + 1]) {
                gaps++;
                f = 0.5 * (bmax[i] + bmin[i + 1]);
            }
        if (gaps != 1) {
            set_err("Detected %d != 1 band gaps", gaps);
            rc = -1;
        }
        *mu = f;
        free(bmin);
        free(bmax);
        free(tmp);
    } else {
        if (occupied_bands <= 0 || occupied_bands >= n) {
            set_err("occupiedBands must be in (0, %d)", n);
            free(ev);
            return -1;
        }
        size_t tot = (size_t)n * c->nk;
        qsort(ev, tot, sizeof(double), cmp_double);
        size_t hi = (size_t)occupied_bands * c->nk - 1;
        double homo = ev[hi], lumo = ev[hi + 1];
        *mu = 0.5 * (homo + lumo);
        if (homo + 3 * kT > lumo) {
            set_err("Fermi level calculation for metals is not implemented");
            rc = -1;
        }
    }
    free(ev);
    return rc;
}

/* current: code/src/ObserverCPU.cpp:149-210 (default region, w = 1) and
 * normalizeExpectationValues :293-308 */
int oracle_observe_j(oracle_ctx *c, const double *state_d, double jout[3])
{
    const cplx *state = (const cplx *)state_d;
    int n = c->n, nn = n * n, dim = c->dim;
    double A[3], kOff[3], off[3];
    for (int d = 0; d < 3; d++) {
        A[d] = creal(state[d]);
        kOff[d] = c->kshift_au[d] + A[d];
    }
    to_frac_k(c->lat, kOff, off);
    double js[3] = {0, 0, 0};
    /* per-thread partial sums combined in thread order: deterministic for a
     * fixed thread count (static schedule) */
    int nth = 1;
#ifdef _OPENMP
    nth = omp_get_max_threads();
#endif
    double *part = calloc((size_t)nth * 3, sizeof(double));
#pragma omp parallel
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        cplx *H0 = malloc(sizeof(cplx) * nn), *D = malloc(sizeof(cplx) * 3 * nn), *dH = malloc(sizeof(cplx) * 3 * nn);
        cplx *T = malloc(sizeof(cplx) * nn);
        double loc[3] = {0, 0, 0};
#pragma omp for schedule(static)
        for (long k = 0; k < c->nk; k++) {
            int ii[3];
            double kf[3];
            k_index(c, k, ii);
            for (int d = 0; d < 3; d++) kf[d] = (double)ii[d] / c->Nk[d] + off[d];
            oracle_hk_full_point(c, kf, (double *)H0, (double *)D, (double *)dH);
            const cplx *rho = state + c->nspecial + (size_t)c->P * k;
            for (int a = 0; a < n; a++)
                for (int b = 0; b < n; b++) {
                    cplx s = 0;
                    for (int i = 0; i < n; i++) s += H0[a * n + i] * rho[i * n + b] - H0[i * n + b] * rho[a * n + i];
                    T[a * n + b] = s;
                }
            for (int d = 0; d < dim; d++) {
                cplx jk = 0;
                for (int a = 0; a < n; a++)
                    for (int b = 0; b < n; b++)
                        jk += I * D[d * nn + a * n + b] * T[b * n + a] - dH[d * nn + a * n + b] * rho[b * n + a];
                loc[d] += creal(jk);
            }
        }
        for (int d = 0; d < 3; d++) part[3 * tid + d] = loc[d];
        free(H0);
        free(D);
        free(dH);
        free(T);
    }
    for (int th = 0; th < nth; th++)
        for (int d = 0; d < 3; d++) js[d] += part[3 * th + d];
    free(part);
    for (int d = 0; d < 3; d++) jout[d] = js[d] / (double)c->nk;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Full observer at one output time: ObserverCPU::operator()
 * (code/src/ObserverCPU.cpp:137-260) and normalizeExpectationValues (:293-308),
 * regions from code/src/ObserverBase.cpp:3-17.  Per k: current j_k (as
 * oracle_observe_j), eigen-decomposition of H0(k + A) (cyclic Jacobi, sweep
 * count of the parameters), Hamiltonian occupations <v_b|rho|v_b> smeared over
 * energy with weights exp(-((e_a - e_b)/w)^2) normalised over b (w <= 0: no
 * smearing), Wannier occupations Re rho_bb, and the density matrix; each
 * summed with the region weight (calcWeight, :137-147) and divided by the
 * weight sum; min / max over k of the smeared occupations.  Bands are
 * numbered in ascending energy (the order LAPACK zheev, the reference's
 * default diagonaliser, returns; the Jacobi's own order is not sorted). */
static double region_weight(const sbe_region *r, const double kgrid[3], const double off[3], const double negA[3])
{
    double inv = isinf(r->fwhm) ? 0.0 : 8 * log(2.0) / (r->fwhm * r->fwhm);
    double w = 1;
    for (int d = 0; d < 3; d++) {
        double pos = kgrid[d] + off[d];
        if (r->moving_frame) pos += negA[d];
        pos -= r->k[d];
        double cp = pos - round(pos);
        w *= exp(-cp * cp * inv);
    }
    return w;
}

int oracle_observe_full(oracle_ctx *c, const double *state_d, const sbe_region *reg, int R, double smearing,
                        double *weight_sum, double *j_out, double *occ_ham, double *occ_wan, double *occ_min,
                        double *occ_max, double *wan)
{
    const cplx *state = (const cplx *)state_d;
    const int n = c->n, nn = n * n, dim = c->dim;
    double A[3], kOff[3], off[3], mA[3], negA[3];
    for (int d = 0; d < 3; d++) {
        A[d] = creal(state[d]);
        kOff[d] = c->kshift_au[d] + A[d];
        mA[d] = -A[d];
    }
    to_frac_k(c->lat, kOff, off);
    to_frac_k(c->lat, mA, negA);
    /* per k-point quantities first (in parallel), then the region sums in k order */
    const int Q = 3 + 2 * n;
    double *X = calloc((size_t)c->nk * Q, sizeof(double));
    double *W = malloc(sizeof(double) * (size_t)c->nk * (R > 0 ? R : 1));
#pragma omp parallel
    {
        cplx *H0 = malloc(sizeof(cplx) * nn), *D = malloc(sizeof(cplx) * 3 * nn), *dH = malloc(sizeof(cplx) * 3 * nn);
        cplx *T = malloc(sizeof(cplx) * nn), *U = malloc(sizeof(cplx) * nn);
        double *eig = malloc(sizeof(double) * n), *occ = malloc(sizeof(double) * n);
        int *ord = malloc(sizeof(int) * n);
#pragma omp for schedule(static)
        for (long k = 0; k < c->nk; k++) {
            int ii[3];
            double kf[3], kg[3];
            k_index(c, k, ii);
            for (int d = 0; d < 3; d++) {
                kg[d] = (double)ii[d] / c->Nk[d];
                kf[d] = kg[d] + off[d];
            }
            for (int r = 0; r < R; r++) W[(size_t)k * R + r] = region_weight(reg + r, kg, off, negA);
            oracle_hk_full_point(c, kf, (double *)H0, (double *)D, (double *)dH);
            const cplx *rho = state + c->nspecial + (size_t)c->P * k;
            double *x = X + (size_t)k * Q;
            for (int a = 0; a < n; a++)
                for (int b = 0; b < n; b++) {
                    cplx s = 0;
                    for (int i = 0; i < n; i++) s += H0[a * n + i] * rho[i * n + b] - H0[i * n + b] * rho[a * n + i];
                    T[a * n + b] = s;
                }
            for (int d = 0; d < dim; d++) {
                cplx jk = 0;
                for (int a = 0; a < n; a++)
                    for (int b = 0; b < n; b++)
                        jk += I * D[d * nn + a * n + b] * T[b * n + a] - dH[d * nn + a * n + b] * rho[b * n + a];
                x[d] = creal(jk);
            }
            /* eigen-decomposition of H0(k + A): rows of U are conjugated eigenvectors */
            jacobi_work(n, c->prm.sweep_count, H0, eig, U);
            /* ascending energy (stable for ties) */
            for (int i = 0; i < n; i++) ord[i] = i;
            for (int i = 1; i < n; i++)
                for (int m = i; m > 0 && eig[ord[m - 1]] > eig[ord[m]]; m--) {
                    int tmp = ord[m];
                    ord[m] = ord[m - 1];
                    ord[m - 1] = tmp;
                }
            /* raw occupations (U rho U^dagger)_bb, code/src/ObserverCPU.cpp:228-233 */
            for (int s = 0; s < n; s++) {
                int bi = ord[s];
                cplx sum = 0;
                for (int a = 0; a < n; a++)
                    for (int b = 0; b < n; b++) sum += conj(U[bi * n + b]) * rho[a * n + b] * U[bi * n + a];
                occ[s] = creal(sum);
            }
            for (int s = 0; s < n; s++) {
                double v;
                if (smearing > 0) {
                    double wsum = 0, acc = 0;
                    for (int q = 0; q < n; q++) {
                        double dE = (eig[ord[s]] - eig[ord[q]]) / smearing;
                        double w = exp(-dE * dE);
                        wsum += w;
                        acc += occ[q] * w;
                    }
                    v = acc / wsum;
                } else {
                    v = occ[s];
                }
                x[3 + s] = v;
                x[3 + n + s] = creal(rho[s * (n + 1)]);
            }
        }
        free(H0);
        free(D);
        free(dH);
        free(T);
        free(U);
        free(eig);
        free(occ);
        free(ord);
    }
    for (int s = 0; s < n; s++) {
        occ_min[s] = INFINITY;
        occ_max[s] = -INFINITY;
    }
    for (long k = 0; k < c->nk; k++)
        for (int s = 0; s < n; s++) {
            double v = X[(size_t)k * Q + 3 + s];
            if (v < occ_min[s]) occ_min[s] = v;
            if (v > occ_max[s]) occ_max[s] = v;
        }
    for (int r = 0; r < R; r++) {
        double ws = 0, acc[3 + 2 * 16] = {0};
        cplx *wr = wan ? (cplx *)wan + (size_t)r * nn : NULL;
        if (wr)
            for (int e = 0; e < nn; e++) wr[e] = 0;
        for (long k = 0; k < c->nk; k++) {
            double w = W[(size_t)k * R + r];
            ws += w;
            for (int q = 0; q < Q; q++) acc[q] += X[(size_t)k * Q + q] * w;
            if (wr) {
                const cplx *rho = state + c->nspecial + (size_t)c->P * k;
                for (int e = 0; e < nn; e++) wr[e] += rho[e] * w;
            }
        }
        weight_sum[r] = ws;
        for (int d = 0; d < 3; d++) j_out[3 * r + d] = acc[d] / ws;
        for (int s = 0; s < n; s++) {
            occ_ham[(size_t)r * n + s] = acc[3 + s] / ws;
            occ_wan[(size_t)r * n + s] = acc[3 + n + s] / ws;
        }
        if (wr)
            for (int e = 0; e < nn; e++) wr[e] /= ws;
    }
    free(X);
    free(W);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Dormand-Prince 5(4) with Boost.Odeint's controller semantics, restated from
 * the library's published algorithm (Boost.Odeint >= 1.53; exact version not
 * pinned by the reference, code/CMakeLists.txt:28):
 *   runge_kutta_dopri5::do_step_impl        (tableau, FSAL, error estimate)
 *   default_error_checker::error            (max_i |err_i| /
 *                                            (eps_abs + eps_rel (|x_i| + dt |dxdt_i|)))
 *   default_step_adjuster                   (decrease: max(0.9 err^(-1/3), 0.2);
 *                                            increase if err < 0.5:
 *                                            0.9 max(err, 5^-5)^(-1/5))
 *   integrate_adaptive / integrate_const    (steps truncated onto output times,
 *                                            dt carried across intervals, a fresh
 *                                            stepper copy - hence a fresh FSAL
 *                                            derivative - per output interval)
 * Call site: code/src/main/runSBE.cpp:170-183.                               */

typedef struct {
    oracle_ctx *c;
    long S;
    cplx *k[7];
    cplx *xt, *xn, *dxn, *err;
    long rhs;
} dp_work;

static const double A2 = 1.0 / 5, A3 = 3.0 / 10, A4 = 4.0 / 5, A5 = 8.0 / 9;
static const double B21 = 1.0 / 5;
static const double B31 = 3.0 / 40, B32 = 9.0 / 40;
static const double B41 = 44.0 / 45, B42 = -56.0 / 15, B43 = 32.0 / 9;
static const double B51 = 19372.0 / 6561, B52 = -25360.0 / 2187, B53 = 64448.0 / 6561, B54 = -212.0 / 729;
static const double B61 = 9017.0 / 3168, B62 = -355.0 / 33, B63 = 46732.0 / 5247, B64 = 49.0 / 176,
                    B65 = -5103.0 / 18656;
static const double C1 = 35.0 / 384, C3 = 500.0 / 1113, C4 = 125.0 / 192, C5 = -2187.0 / 6784, C6 = 11.0 / 84;

static void dp_rhs(dp_work *w, const cplx *x, cplx *dx, double t)
{
    oracle_rhs(w->c, (const double *)x, (double *)dx, t);
    w->rhs++;
}

/* one dopri5 attempt from (x, dxdt, t) with step dt -> xn, dxn; returns the
 * error norm */
static double dp_attempt(dp_work *w, const cplx *x, const cplx *dxdt, double t, double dt)
{
    long S = w->S;
    cplx *k2 = w->k[1], *k3 = w->k[2], *k4 = w->k[3], *k5 = w->k[4], *k6 = w->k[5];
    cplx *xt = w->xt;
    double b21 = dt * B21;
    for (long i = 0; i < S; i++) xt[i] = 1.0 * x[i] + b21 * dxdt[i];
    dp_rhs(w, xt, k2, t + dt * A2);
    double b31 = dt * B31, b32 = dt * B32;
    for (long i = 0; i < S; i++) xt[i] = 1.0 * x[i] + b31 * dxdt[i] + b32 * k2[i];
    dp_rhs(w, xt, k3, t + dt * A3);
    double b41 = dt * B41, b42 = dt * B42, b43 = dt * B43;
    for (long i = 0; i < S; i++) xt[i] = 1.0 * x[i] + b41 * dxdt[i] + b42 * k2[i] + b43 * k3[i];
    dp_rhs(w, xt, k4, t + dt * A4);
    double b51 = dt * B51, b52 = dt * B52, b53 = dt * B53, b54 = dt * B54;
    for (long i = 0; i < S; i++) xt[i] = 1.0 * x[i] + b51 * dxdt[i] + b52 * k2[i] + b53 * k3[i] + b54 * k4[i];
    dp_rhs(w, xt, k5, t + dt * A5);
    double b61 = dt * B61, b62 = dt * B62, b63 = dt * B63, b64 = dt * B64, b65 = dt * B65;
    for (long i = 0; i < S; i++)
        xt[i] = 1.0 * x[i] + b61 * dxdt[i] + b62 * k2[i] + b63 * k3[i] + b64 * k4[i] + b65 * k5[i];
    dp_rhs(w, xt, k6, t + dt);
    double c1 = dt * C1, c3 = dt * C3, c4 = dt * C4, c5 = dt * C5, c6 = dt * C6;
    for (long i = 0; i < S; i++)
        w->xn[i] = 1.0 * x[i] + c1 * dxdt[i] + c3 * k3[i] + c4 * k4[i] + c5 * k5[i] + c6 * k6[i];
    dp_rhs(w, w->xn, w->dxn, t + dt);
    /* error estimate (dc_i = c_i - chat_i) */
    const double dc1 = C1 - 5179.0 / 57600, dc3 = C3 - 7571.0 / 16695, dc4 = C4 - 393.0 / 640,
                 dc5 = C5 - -92097.0 / 339200, dc6 = C6 - 187.0 / 2100, dc7 = -1.0 / 40;
    double e1 = dt * dc1, e3 = dt * dc3, e4 = dt * dc4, e5 = dt * dc5, e6 = dt * dc6, e7 = dt * dc7;
    double ea = w->c->prm.eps_abs, er = w->c->prm.eps_rel;
    double mx = 0;
    for (long i = 0; i < S; i++) {
        cplx er_i = e1 * dxdt[i] + e3 * k3[i] + e4 * k4[i] + e5 * k5[i] + e6 * k6[i] + e7 * w->dxn[i];
        double v = cabs(er_i) / (ea + er * (1.0 * cabs(x[i]) + (1.0 * dt) * cabs(dxdt[i])));
        if (v > mx) mx = v;
    }
    return mx;
}

static int less_with_sign(double t1, double t2) { return t2 - t1 > 2.220446049250313e-16; }
static int less_eq_with_sign(double t1, double t2) { return t1 - t2 <= 2.220446049250313e-16; }

static void log_push(sbe_steplog *log, double t, double dt, double err, int acc)
{
    if (!log) return;
    if (log->count < log->capacity) {
        log->t[log->count] = t;
        log->dt[log->count] = dt;
        log->err[log->count] = err;
        log->accepted[log->count] = acc;
    }
    log->count++;
}

/* the controller's time arithmetic is kept free of FMA contraction */
__attribute__((optimize("fp-contract=off")))
long oracle_integrate(oracle_ctx *c, double *state_d, int output_count, double *t_out, double *j_out,
                      sbe_steplog *log, const sbe_steplog *replay, long *rhs_calls)
{
    if (output_count < 2) {
        set_err("output_count must be >= 2");
        return -1;
    }
    cplx *x = (cplx *)state_d;
    dp_work w;
    w.c = c;
    w.S = c->state_size;
    w.rhs = 0;
    for (int i = 0; i < 7; i++) w.k[i] = malloc(sizeof(cplx) * w.S);
    w.xt = malloc(sizeof(cplx) * w.S);
    w.xn = malloc(sizeof(cplx) * w.S);
    w.dxn = malloc(sizeof(cplx) * w.S);
    cplx *dxdt = w.k[0];
    double t0 = oracle_start_time(c), t1 = oracle_end_time(c);
    double dt_out = (t1 - t0) / (output_count - 1);
    double t_end = t1 + dt_out / 2;
    double time = t0, dt = dt_out;
    long steps = 0, step = 0;
    int out = 0, rp = 0;
    long rc = 0;
    if (log) log->count = 0;
    while (less_eq_with_sign(time + dt_out, t_end)) {
        if (out < output_count) {
            if (t_out) t_out[out] = time;
            if (j_out) oracle_observe_j(c, (double *)x, j_out + 3 * out);
        }
        out++;
        /* integrate_adaptive with a fresh controlled stepper */
        double start = time, end = time + dt_out;
        int first = 1, fails = 0;
        if (replay) {
            /* replay: the logged accepted attempts starting in this output
             * interval, with the logged (t, dt) */
            dp_rhs(&w, x, dxdt, start);
            double tol = 1e-9 * (fabs(end) > 1 ? fabs(end) : 1);
            for (;;) {
                while (rp < replay->count && rp < replay->capacity && !replay->accepted[rp]) rp++;
                if (rp >= replay->count || rp >= replay->capacity) break;
                if (replay->t[rp] >= end - tol) break;
                start = replay->t[rp];
                dt = replay->dt[rp++];
                double err = dp_attempt(&w, x, dxdt, start, dt);
                log_push(log, start, dt, err, 1);
                memcpy(x, w.xn, sizeof(cplx) * w.S);
                memcpy(dxdt, w.dxn, sizeof(cplx) * w.S);
                steps++;
            }
            if (steps == 0) {
                set_err("replay log does not cover the interval ending at t=%g", end);
                rc = -3;
                goto done;
            }
        }
        while (!replay && less_with_sign(start, end)) {
            if (less_with_sign(end, start + dt)) dt = end - start;
            if (first) {
                dp_rhs(&w, x, dxdt, start);
                first = 0;
            }
            for (;;) {
                double err = dp_attempt(&w, x, dxdt, start, dt);
                if (err > 1.0) {
                    log_push(log, start, dt, err, 0);
                    double f = 0.9 * pow(err, -1.0 / 3.0);
                    dt *= f > 0.2 ? f : 0.2;
                    if (++fails >= 500) {
                        set_err("Max number of iterations exceeded (500). A new step size was not found.");
                        rc = -2;
                        goto done;
                    }
                    continue;
                }
                log_push(log, start, dt, err, 1);
                start += dt;
                if (err < 0.5) {
                    double e = err > pow(5.0, -5.0) ? err : pow(5.0, -5.0);
                    dt *= 0.9 * pow(e, -1.0 / 5.0);
                }
                break;
            }
            fails = 0;
            /* accept: x <- xn, dxdt <- dxn */
            memcpy(x, w.xn, sizeof(cplx) * w.S);
            memcpy(dxdt, w.dxn, sizeof(cplx) * w.S);
            steps++;
        }
        step++;
        time = t0 + (double)step * dt_out;
    }
    if (out < output_count) {
        if (t_out) t_out[out] = time;
        if (j_out) oracle_observe_j(c, (double *)x, j_out + 3 * out);
    }
    out++;
    rc = steps;
done:
    if (rhs_calls) *rhs_calls = w.rhs;
    for (int i = 0; i < 7; i++) free(w.k[i]);
    free(w.xt);
    free(w.xn);
    free(w.dxn);
    return rc;
}

/* Time `count` Dormand-Prince attempts at fixed (t, dt) from `state` (each
 * attempt: 6 RHS evaluations plus stage algebra and error norm; the state is
 * not advanced).  Used as the CPU baseline of the benchmark. Returns the
 * elapsed wall time in seconds. */
int oracle_num_threads(void) { return omp_get_max_threads(); }

double oracle_bench_attempts(oracle_ctx *c, const double *state_d, double t, double dt, int count)
{
    dp_work w;
    w.c = c;
    w.S = c->state_size;
    w.rhs = 0;
    for (int i = 0; i < 7; i++) w.k[i] = malloc(sizeof(cplx) * w.S);
    w.xt = malloc(sizeof(cplx) * w.S);
    w.xn = malloc(sizeof(cplx) * w.S);
    w.dxn = malloc(sizeof(cplx) * w.S);
    const cplx *x = (const cplx *)state_d;
    dp_rhs(&w, x, w.k[0], t);
    double t0 = omp_get_wtime();
    for (int i = 0; i < count; i++) dp_attempt(&w, x, w.k[0], t, dt);
    double el = omp_get_wtime() - t0;
    for (int i = 0; i < 7; i++) free(w.k[i]);
    free(w.xt);
    free(w.xn);
    free(w.dxn);
    return el;
}
