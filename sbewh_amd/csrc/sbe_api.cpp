// Host side of libsbewh_hip.so: context set-up, pulses, the device-resident
// Dormand-Prince controller and the C ABI declared in include/sbewh_hip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/sbewh_hip.h"
#include "kernels.hpp"

using namespace sbe;

namespace {

thread_local std::string g_create_err;

struct HipFail {
    hipError_t e;
    const char *what;
};

#define HIPCHK(x)                                          \
    do {                                                   \
        hipError_t e_ = (x);                               \
        if (e_ != hipSuccess) throw HipFail{e_, #x};       \
    } while (0)

struct ArgFail {
    int code;
    std::string msg;
};

[[noreturn]] void fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    throw ArgFail{code, buf};
}

// ---- pulses: code/src/pulse/{Pulse3D,Sin2Pulse1D,GaussianPulse1D,PureGauss1D,Sin2RampPulse1D}.cpp
struct SubPulse {
    sbe_subpulse p;
    double pol[3];
};

double sin2_env_field(double E0, double w, double cep, double beta, double t)
{
    double cp = w * t + cep, ec = std::cos(beta * t);
    return E0 * (std::cos(cp) * ec * ec - 2 * beta / w * std::sin(cp) * ec * std::sin(beta * t));
}

double sub_E(const sbe_subpulse &p, double t)
{
    const double w = p.omega, F = p.fwhm, E0 = p.E0;
    switch (p.type) {
    case SBE_PULSE_SIN2:
        if (t < -F || t > F) return 0;
        return sin2_env_field(E0, w, p.cep, M_PI / 2 / F, t);
    case SBE_PULSE_GAUSS: {
        double beta = 4 * std::log(2.0) / (F * F), cp = w * t + p.cep;
        return E0 * (-2 * beta / w * t * std::sin(cp) + std::cos(cp)) * std::exp(-beta * t * t);
    }
    case SBE_PULSE_PURE_GAUSS: {
        double beta = 4 * std::log(2.0) / (F * F);
        return E0 * std::exp(-beta * t * t);
    }
    case SBE_PULSE_SIN2_RAMP: {
        double rc = p.rising_cycles, beta = w / 4 / rc;
        int co = p.cycles_on;
        double rising = 2 * M_PI / w * rc, on = M_PI / w * co, half = rising + M_PI / w * co;
        if (t < -half) return 0;
        if (t < -on) return sin2_env_field(E0, w, 0.0, beta, t + on);
        if (t < on) return ((co % 2 == 0) ? 1 : -1) * E0 * std::cos(w * t);
        if (t < half) return sin2_env_field(E0, w, 0.0, beta, t - on);
        return 0;
    }
    }
    return 0;
}

double sub_start(const sbe_subpulse &p)
{
    if (p.type == SBE_PULSE_SIN2_RAMP) return -2 * M_PI / p.omega * (p.rising_cycles + 0.5 * p.cycles_on);
    return p.t_start_fwhm * p.fwhm;
}
double sub_end(const sbe_subpulse &p)
{
    if (p.type == SBE_PULSE_SIN2_RAMP) return 2 * M_PI / p.omega * (p.rising_cycles + 0.5 * p.cycles_on);
    return p.t_end_fwhm * p.fwhm;
}

int smooth5(int N)
{
    int n = N - 1, r;
    do {
        r = ++n;
        for (int p : {2, 3, 5})
            while (r % p == 0) r /= p;
    } while (r != 1);
    return n;
}

// Dormand-Prince 5(4) tableau (runge_kutta_dopri5)
const double A2 = 1.0 / 5, A3 = 3.0 / 10, A4 = 4.0 / 5, A5 = 8.0 / 9;
const double B21 = 1.0 / 5;
const double B31 = 3.0 / 40, B32 = 9.0 / 40;
const double B41 = 44.0 / 45, B42 = -56.0 / 15, B43 = 32.0 / 9;
const double B51 = 19372.0 / 6561, B52 = -25360.0 / 2187, B53 = 64448.0 / 6561, B54 = -212.0 / 729;
const double B61 = 9017.0 / 3168, B62 = -355.0 / 33, B63 = 46732.0 / 5247, B64 = 49.0 / 176,
             B65 = -5103.0 / 18656;
const double C1 = 35.0 / 384, C3 = 500.0 / 1113, C4 = 125.0 / 192, C5 = -2187.0 / 6784, C6 = 11.0 / 84;
const double DC1 = C1 - 5179.0 / 57600, DC3 = C3 - 7571.0 / 16695, DC4 = C4 - 393.0 / 640,
             DC5 = C5 - -92097.0 / 339200, DC6 = C6 - 187.0 / 2100, DC7 = -1.0 / 40;

}  // namespace

struct sbe_ctx {
    std::string err;
    int n = 0, N = 0, dim = 0, P = 0, nn = 0;
    long nk = 0, S = 0;
    sbe_params prm{};
    std::vector<SubPulse> pulses;
    PhaseGeom geom{};
    KOrderH korder{};
    int obs_ops = 0;
    hipStream_t stream = nullptr;
    int device = 0;
    // device buffers
    double *Hb = nullptr, *Db = nullptr, *dHb = nullptr;
    int *kmap = nullptr;
    double *tw = nullptr, *y1 = nullptr, *y2 = nullptr;
    double *mesh = nullptr, *mesh_obs = nullptr;
    double *st[10] = {};  // x, k1..k7, xt, xn
    double *ucache = nullptr;  // [k][n*n] eigenvectors of the last RHS (warm starts)
    bool ucache_valid = false;
    unsigned long long *err_bits = nullptr;
    double *err_partial = nullptr;
    unsigned long long *err_host = nullptr;  // pinned
    double *partial = nullptr;
    double *jbuf = nullptr;
    long jbuf_cap = 0;
    hipEvent_t ev[4] = {};
    long rhs_calls = 0;
    // optional timing of the phase-sum and stage kernels: events recorded
    // around each launch, evaluated after the timed region
    bool timing = false;
    std::vector<hipEvent_t> evpool;
    size_t evused = 0;
    hipEvent_t next_event()
    {
        if (evused == evpool.size()) {
            hipEvent_t e;
            HIPCHK(hipEventCreate(&e));
            evpool.push_back(e);
        }
        return evpool[evused++];
    }

    double *x() { return st[0]; }

    // stage-kernel layout (lanes per row), chosen by tune_stage()
    int lr = 0;
    double tune_ms[8] = {};
    int tune_lr[8] = {};
    int tune_count = 0;

    static int default_lr(int N) { return N <= 4 ? 1 : 2; }

    // Time every compiled layout variant of the stage kernel on the current
    // inputs (3 timed launches after one warm-up each) and keep the fastest.
    // SBEWH_STAGE_LR in the environment forces a variant.
    void tune_stage(const StageArgs &a, hipStream_t s)
    {
        int lrs[8];
        int nv = stage_variants(N, lrs, 8);
        if (const char *env = std::getenv("SBEWH_STAGE_LR")) {
            int want = std::atoi(env);
            for (int v = 0; v < nv; v++)
                if (lrs[v] == want) lr = want;
            if (lr) return;
        }
        if (nv == 1) {
            lr = lrs[0];
            return;
        }
        float best = 1e30f;
        tune_count = nv;
        for (int v = 0; v < nv; v++) {
            HIPCHK(launch_rhs_stage(N, lrs[v], a, s));
            HIPCHK(hipEventRecord(ev[0], s));
            for (int it = 0; it < 3; it++) HIPCHK(launch_rhs_stage(N, lrs[v], a, s));
            HIPCHK(hipEventRecord(ev[1], s));
            HIPCHK(hipEventSynchronize(ev[1]));
            float ms = 0;
            HIPCHK(hipEventElapsedTime(&ms, ev[0], ev[1]));
            tune_lr[v] = lrs[v];
            tune_ms[v] = ms / 3;
            if (ms < best) {
                best = ms;
                lr = lrs[v];
            }
        }
    }

    void E_at(double t, double E[3]) const
    {
        E[0] = E[1] = E[2] = 0;
        for (const auto &s : pulses) {
            double e = sub_E(s.p, t - s.p.center);
            for (int d = 0; d < 3; d++) E[d] += e * s.pol[d];
        }
    }
    double t_start() const
    {
        double r = INFINITY;
        for (const auto &s : pulses) r = std::min(r, sub_start(s.p) + s.p.center);
        return r;
    }
    double t_end() const
    {
        double r = -INFINITY;
        for (const auto &s : pulses) r = std::max(r, sub_end(s.p) + s.p.center);
        return r;
    }

    PhaseBuffers pbufs() const { return PhaseBuffers{Hb, Db, dHb, kmap, tw, y1, y2}; }

    // dsdt = f(x_in, t) on device
    void rhs_dev(const double *x_in, double *dxdt, double t, hipStream_t s)
    {
        double E[3];
        E_at(t, E);
        if (timing) HIPCHK(hipEventRecord(next_event(), s));
        HIPCHK(launch_twiddles(geom, x_in, 1, tw, s));
        HIPCHK(launch_phase_sum(geom, pbufs(), 2, E[0], E[1], E[2], mesh, s));
        if (timing) HIPCHK(hipEventRecord(next_event(), s));
        StageArgs a{};
        a.x = x_in;
        a.dxdt = dxdt;
        a.mesh = mesh;
        a.nk = nk;
        a.n = n;
        a.P = P;
        a.dim = dim;
        a.sweeps = prm.sweep_count;
        a.F2 = prm.relaxation_F2;
        a.w = prm.soothing_width;
        for (int d = 0; d < 3; d++) a.E[d] = E[d];
        a.dephase = prm.relaxation_F2 > 0;
        a.ucache = ucache;
        a.warm = 0;
        if (!lr) tune_stage(a, s);
        a.warm = (ucache_valid && !(prm.flags & SBE_FLAG_COLD_JACOBI)) ? 1 : 0;
        HIPCHK(launch_rhs_stage(N, lr, a, s));
        if (a.dephase) ucache_valid = true;
        if (timing) HIPCHK(hipEventRecord(next_event(), s));
        rhs_calls++;
    }

    void observe_dev(const double *xs, double *jdst, hipStream_t s)
    {
        HIPCHK(launch_twiddles(geom, xs, 1, tw, s));
        HIPCHK(launch_phase_sum(geom, pbufs(), obs_ops, 0, 0, 0, mesh_obs, s));
        ObsArgs o{};
        o.x = xs;
        o.mesh = mesh_obs;
        o.nk = nk;
        o.n = n;
        o.P = P;
        o.dim = dim;
        o.partial = partial;
        o.j_out = jdst;
        HIPCHK(launch_observe_j(o, s));
    }

    // one Dormand-Prince attempt from (x, k1, t, dt) -> xn, k7; returns the error norm
    double attempt(double t, double dt)
    {
        double *X = st[0], *k1 = st[1], *k2 = st[2], *k3 = st[3], *k4 = st[4], *k5 = st[5], *k6 = st[6],
               *k7 = st[7], *xt = st[8], *xn = st[9];
        hipStream_t s = stream;
        {
            const double *kk[1] = {k1};
            double c[1] = {dt * B21};
            HIPCHK(launch_combine(S, X, 1, kk, c, xt, s));
            rhs_dev(xt, k2, t + dt * A2, s);
        }
        {
            const double *kk[2] = {k1, k2};
            double c[2] = {dt * B31, dt * B32};
            HIPCHK(launch_combine(S, X, 2, kk, c, xt, s));
            rhs_dev(xt, k3, t + dt * A3, s);
        }
        {
            const double *kk[3] = {k1, k2, k3};
            double c[3] = {dt * B41, dt * B42, dt * B43};
            HIPCHK(launch_combine(S, X, 3, kk, c, xt, s));
            rhs_dev(xt, k4, t + dt * A4, s);
        }
        {
            const double *kk[4] = {k1, k2, k3, k4};
            double c[4] = {dt * B51, dt * B52, dt * B53, dt * B54};
            HIPCHK(launch_combine(S, X, 4, kk, c, xt, s));
            rhs_dev(xt, k5, t + dt * A5, s);
        }
        {
            const double *kk[5] = {k1, k2, k3, k4, k5};
            double c[5] = {dt * B61, dt * B62, dt * B63, dt * B64, dt * B65};
            HIPCHK(launch_combine(S, X, 5, kk, c, xt, s));
            rhs_dev(xt, k6, t + dt, s);
        }
        {
            const double *kk[5] = {k1, k3, k4, k5, k6};
            double c[5] = {dt * C1, dt * C3, dt * C4, dt * C5, dt * C6};
            HIPCHK(launch_combine(S, X, 5, kk, c, xn, s));
            rhs_dev(xn, k7, t + dt, s);
        }
        const double *ek[6] = {k1, k3, k4, k5, k6, k7};
        double e[6] = {dt * DC1, dt * DC3, dt * DC4, dt * DC5, dt * DC6, dt * DC7};
        HIPCHK(launch_error_norm(S, X, ek, e, prm.eps_abs, prm.eps_rel, dt, err_partial, err_bits, s));
        HIPCHK(hipMemcpyAsync(err_host, err_bits, sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        double v;
        std::memcpy(&v, err_host, sizeof v);
        return v;
    }

    void accept()
    {
        std::swap(st[0], st[9]);  // x <- xn
        std::swap(st[1], st[7]);  // k1 <- k7 (FSAL)
    }

    ~sbe_ctx()
    {
        double *ptrs[] = {Hb, Db, dHb, tw, y1, y2, mesh, mesh_obs, partial, jbuf, ucache};
        for (double *p : ptrs)
            if (p) (void)hipFree(p);
        for (double *p : st)
            if (p) (void)hipFree(p);
        if (kmap) (void)hipFree(kmap);
        if (err_bits) (void)hipFree(err_bits);
        if (err_partial) (void)hipFree(err_partial);
        if (err_host) (void)hipHostFree(err_host);
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
        for (auto &e : evpool) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

template <class F>
int guarded(sbe_ctx *c, F &&f)
{
    try {
        f();
        return SBE_OK;
    } catch (const HipFail &h) {
        char buf[512];
        snprintf(buf, sizeof buf, "HIP error %d (%s) in %s", (int)h.e, hipGetErrorString(h.e), h.what);
        if (c) c->err = buf;
        else g_create_err = buf;
        return SBE_ERR_HIP;
    } catch (const ArgFail &a) {
        if (c) c->err = a.msg;
        else g_create_err = a.msg;
        return a.code;
    } catch (const std::exception &e) {
        if (c) c->err = e.what();
        else g_create_err = e.what();
        return SBE_ERR_ARG;
    }
}

void set_pulses(sbe_ctx *c, const sbe_subpulse *p, int np)
{
    if (!p || np < 1 || np > 8) fail(SBE_ERR_ARG, "number of sub-pulses must be in [1, 8] (got %d)", np);
    c->pulses.clear();
    for (int i = 0; i < np; i++) {
        SubPulse s;
        s.p = p[i];
        double nrm = std::sqrt(p[i].pol[0] * p[i].pol[0] + p[i].pol[1] * p[i].pol[1] + p[i].pol[2] * p[i].pol[2]);
        if (!(nrm > 0)) fail(SBE_ERR_ARG, "sub-pulse %d has a zero polarisation", i);
        for (int d = 0; d < 3; d++) s.pol[d] = p[i].pol[d] / nrm;
        if (p[i].type < 0 || p[i].type > 3) fail(SBE_ERR_ARG, "unknown pulse type %d", p[i].type);
        if (p[i].type == SBE_PULSE_SIN2_RAMP && !(p[i].rising_cycles > 0))
            fail(SBE_ERR_ARG, "sin2Ramp needs rising_cycles > 0");
        if (p[i].type != SBE_PULSE_PURE_GAUSS && !(p[i].omega > 0)) fail(SBE_ERR_ARG, "omega must be > 0");
        if (p[i].type != SBE_PULSE_SIN2_RAMP && !(p[i].fwhm > 0)) fail(SBE_ERR_ARG, "fwhm must be > 0");
        c->pulses.push_back(s);
    }
}

template <class T>
T *dalloc(size_t count)
{
    void *p = nullptr;
    if (count == 0) count = 1;
    HIPCHK(hipMalloc(&p, count * sizeof(T)));
    return static_cast<T *>(p);
}

void build(sbe_ctx *c, const sbe_model *m, const sbe_grid *g, const sbe_params *prm, const sbe_subpulse *pul,
           int np)
{
    if (!m || !g || !prm) fail(SBE_ERR_ARG, "null argument");
    if (m->num_wann < 1 || m->num_wann > 16) fail(SBE_ERR_ARG, "num_wann must be in [1, 16] (got %d)", m->num_wann);
    if (m->num_cells < 1 || !m->cells || !m->H || !m->D) fail(SBE_ERR_ARG, "empty model");
    if (g->dim < 1 || g->dim > 3) fail(SBE_ERR_ARG, "dim must be 1, 2 or 3");
    if (prm->sweep_count < 0 || prm->sweep_count > 50) fail(SBE_ERR_ARG, "sweep_count must be in [0, 50]");
    const int n = m->num_wann, nn = n * n;
    c->n = n;
    c->nn = nn;
    c->N = n <= 2 ? 2 : n <= 4 ? 4 : n <= 8 ? 8 : 16;
    c->dim = g->dim;
    c->prm = *prm;
    set_pulses(c, pul, np);
    // projection onto dim dimensions (code/src/util/w90util.cpp:417-444)
    std::map<std::tuple<int, int, int>, int> idx;
    std::vector<std::array<int, 3>> cells;
    std::vector<std::complex<double>> H, D;
    for (int i = 0; i < m->num_cells; i++) {
        std::array<int, 3> r;
        for (int d = 0; d < 3; d++) r[d] = d < g->dim ? m->cells[3 * i + d] : 0;
        auto key = std::make_tuple(r[0], r[1], r[2]);
        auto it = idx.find(key);
        int ci;
        if (it == idx.end()) {
            ci = (int)cells.size();
            idx[key] = ci;
            cells.push_back(r);
            H.resize(H.size() + nn);
            D.resize(D.size() + 3 * nn);
        } else {
            ci = it->second;
        }
        for (int e = 0; e < nn; e++) {
            H[(size_t)ci * nn + e] += std::complex<double>(m->H[2 * ((size_t)i * nn + e)], m->H[2 * ((size_t)i * nn + e) + 1]);
            for (int d = 0; d < 3; d++) {
                size_t s = ((size_t)i * 3 + d) * nn + e;
                D[((size_t)ci * 3 + d) * nn + e] += std::complex<double>(m->D[2 * s], m->D[2 * s + 1]);
            }
        }
    }
    if (!idx.count(std::make_tuple(0, 0, 0))) fail(SBE_ERR_ARG, "projected hamiltonian does not contain R=(0,0,0)");
    int mn[3] = {0, 0, 0}, mx[3] = {0, 0, 0};
    for (auto &r : cells)
        for (int d = 0; d < 3; d++) {
            mn[d] = std::min(mn[d], r[d]);
            mx[d] = std::max(mx[d], r[d]);
        }
    PhaseGeom &G = c->geom;
    G.n = n;
    G.dim = g->dim;
    for (int d = 0; d < 3; d++) {
        G.W[d] = mx[d] - mn[d] + 1;
        G.rmin[d] = mn[d];
    }
    std::memcpy(G.lattice, m->lattice, sizeof G.lattice);
    // Nf rule (code/src/Parameter.cpp:4-56)
    int Nf[3];
    for (int d = 0; d < 3; d++) Nf[d] = d < g->dim ? (g->Nf[d] > 0 ? g->Nf[d] : 1) : 1;
    for (int d = 0; d < g->dim; d++) Nf[d] = std::max(Nf[d], G.W[d]);
    if (std::max({Nf[0], Nf[1], Nf[2]}) > 5)
        for (int d = 0; d < g->dim; d++) Nf[d] = smooth5(Nf[d]);
    for (int d = 0; d < 3; d++) {
        int Nk = d < g->dim ? g->Nk[d] : 1;
        if (Nk < 1 || Nk % Nf[d] != 0) fail(SBE_ERR_ARG, "In k-dimension %d: Nk (%d) is not a multiple of Nf (%d)", d + 1, Nk, Nf[d]);
        G.Nk[d] = Nk;
        c->korder.Nf[d] = Nf[d];
        c->korder.Nrep[d] = Nk / Nf[d];
    }
    c->nk = (long)G.Nk[0] * G.Nk[1] * G.Nk[2];
    c->P = 8 * ((nn + 7) / 8);
    c->S = 8 + (long)c->P * c->nk;
    // kGlobalShift -> a.u. (code/src/StateContext.cpp:3-9)
    {
        const double *L = m->lattice;
        double c12[3] = {L[4] * L[8] - L[5] * L[7], L[5] * L[6] - L[3] * L[8], L[3] * L[7] - L[4] * L[6]};
        double c20[3] = {L[7] * L[2] - L[8] * L[1], L[8] * L[0] - L[6] * L[2], L[6] * L[1] - L[7] * L[0]};
        double c01[3] = {L[1] * L[5] - L[2] * L[4], L[2] * L[3] - L[0] * L[5], L[0] * L[4] - L[1] * L[3]};
        double vol = L[0] * c12[0] + L[1] * c12[1] + L[2] * c12[2];
        if (!(std::fabs(vol) > 0)) fail(SBE_ERR_ARG, "degenerate lattice vectors");
        double pre = 2 * M_PI / vol;
        const double *kf = g->k_global_shift;
        for (int i = 0; i < 3; i++) G.kshift_au[i] = pre * (kf[0] * c12[i] + kf[1] * c20[i] + kf[2] * c01[i]);
    }
    c->obs_ops = 1 + 2 * g->dim;
    // dense R box arrays
    long nbox = (long)G.W[0] * G.W[1] * G.W[2];
    std::vector<std::complex<double>> Hb(nbox * nn), Db(3 * nbox * nn), dHb(3 * nbox * nn);
    for (size_t ci = 0; ci < cells.size(); ci++) {
        auto &r = cells[ci];
        long b = ((long)(r[0] - mn[0]) * G.W[1] + (r[1] - mn[1])) * G.W[2] + (r[2] - mn[2]);
        double pos[3];
        for (int d = 0; d < 3; d++) pos[d] = r[0] * m->lattice[d] + r[1] * m->lattice[3 + d] + r[2] * m->lattice[6 + d];
        for (int e = 0; e < nn; e++) {
            std::complex<double> h = H[ci * nn + e];
            Hb[b * nn + e] = h;
            for (int d = 0; d < 3; d++) {
                Db[((long)d * nbox + b) * nn + e] = D[(ci * 3 + d) * nn + e];
                dHb[((long)d * nbox + b) * nn + e] = std::complex<double>(0, 1) * pos[d] * h;
            }
        }
    }
    // grid index -> k index map
    std::vector<int> kmap(c->nk);
    {
        const auto &o = c->korder;
        long fft = (long)o.Nf[0] * o.Nf[1] * o.Nf[2];
        for (long k = 0; k < c->nk; k++) {
            long blk = k / fft, f = k % fft;
            int a = (int)(f / (o.Nf[1] * o.Nf[2])), bb = (int)((f / o.Nf[2]) % o.Nf[1]), cc = (int)(f % o.Nf[2]);
            int ar = (int)(blk / (o.Nrep[1] * o.Nrep[2])), br = (int)((blk / o.Nrep[2]) % o.Nrep[1]),
                cr = (int)(blk % o.Nrep[2]);
            long i0 = a * o.Nrep[0] + ar, i1 = bb * o.Nrep[1] + br, i2 = cc * o.Nrep[2] + cr;
            kmap[(i0 * G.Nk[1] + i1) * G.Nk[2] + i2] = (int)k;
        }
    }
    if (c->nk > 0x7fffffffL) fail(SBE_ERR_ARG, "too many k-points");
    // device memory
    HIPCHK(hipGetDevice(&c->device));
    HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    for (auto &e : c->ev) HIPCHK(hipEventCreate(&e));
    c->Hb = dalloc<double>(2 * Hb.size());
    c->Db = dalloc<double>(2 * Db.size());
    c->dHb = dalloc<double>(2 * dHb.size());
    HIPCHK(hipMemcpy(c->Hb, Hb.data(), Hb.size() * 16, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->Db, Db.data(), Db.size() * 16, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(c->dHb, dHb.data(), dHb.size() * 16, hipMemcpyHostToDevice));
    c->kmap = dalloc<int>(c->nk);
    HIPCHK(hipMemcpy(c->kmap, kmap.data(), kmap.size() * sizeof(int), hipMemcpyHostToDevice));
    long ntw = (long)G.Nk[0] * G.W[0] + (long)G.Nk[1] * G.W[1] + (long)G.Nk[2] * G.W[2];
    c->tw = dalloc<double>(2 * ntw);
    long M = (long)c->obs_ops * nn;
    long ysz = std::max((long)G.W[0] * G.W[1] * G.Nk[2], (long)G.W[0] * G.Nk[1] * G.Nk[2]) * M;
    c->y1 = dalloc<double>(2 * ysz);
    c->y2 = dalloc<double>(2 * ysz);
    c->mesh = dalloc<double>(2 * c->nk * 2 * nn);
    c->mesh_obs = dalloc<double>(2 * c->nk * M);
    for (auto &p : c->st) {
        p = dalloc<double>(2 * c->S);
        HIPCHK(hipMemset(p, 0, 16 * c->S));
    }
    c->ucache = dalloc<double>(2 * c->nk * nn);
    c->err_bits = dalloc<unsigned long long>(1);
    c->err_partial = dalloc<double>(1024);
    HIPCHK(hipHostMalloc((void **)&c->err_host, sizeof(unsigned long long), hipHostMallocDefault));
    c->partial = dalloc<double>(3 * c->nk);
}

bool less_with_sign(double t1, double t2) { return t2 - t1 > 2.220446049250313e-16; }
bool less_eq_with_sign(double t1, double t2) { return t1 - t2 <= 2.220446049250313e-16; }

void log_push(sbe_steplog *log, double t, double dt, double err, int acc)
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

}  // namespace

extern "C" {

const char *sbe_version(void) { return "sbewh-mi355x 0.1 (gfx950, FP64)"; }

const char *sbe_last_error(const sbe_ctx *ctx) { return ctx ? ctx->err.c_str() : g_create_err.c_str(); }

int sbe_create(const sbe_model *model, const sbe_grid *grid, const sbe_params *params, const sbe_subpulse *pulses,
               int num_pulses, sbe_ctx **out)
{
    if (!out) return SBE_ERR_ARG;
    *out = nullptr;
    sbe_ctx *c = new sbe_ctx();
    int rc = guarded(nullptr, [&] { build(c, model, grid, params, pulses, num_pulses); });
    if (rc) {
        delete c;
        return rc;
    }
    *out = c;
    return SBE_OK;
}

void sbe_destroy(sbe_ctx *ctx) { delete ctx; }

int sbe_set_pulses(sbe_ctx *ctx, const sbe_subpulse *pulses, int num_pulses)
{
    if (!ctx) return SBE_ERR_ARG;
    return guarded(ctx, [&] { set_pulses(ctx, pulses, num_pulses); });
}

long sbe_state_size(const sbe_ctx *ctx) { return ctx ? ctx->S : SBE_ERR_ARG; }
long sbe_num_k(const sbe_ctx *ctx) { return ctx ? ctx->nk : SBE_ERR_ARG; }

int sbe_get_nf(const sbe_ctx *ctx, int nf[3])
{
    if (!ctx || !nf) return SBE_ERR_ARG;
    for (int d = 0; d < 3; d++) nf[d] = ctx->korder.Nf[d];
    return SBE_OK;
}

int sbe_time_window(const sbe_ctx *ctx, double *t0, double *t1)
{
    if (!ctx || !t0 || !t1) return SBE_ERR_ARG;
    *t0 = ctx->t_start();
    *t1 = ctx->t_end();
    return SBE_OK;
}

int sbe_pulse_E(const sbe_ctx *ctx, double t, double E[3])
{
    if (!ctx || !E) return SBE_ERR_ARG;
    ctx->E_at(t, E);
    return SBE_OK;
}

int sbe_fermi_level(sbe_ctx *ctx, int occupied_bands, int below_gap, double *mu)
{
    if (!ctx || !mu) return SBE_ERR_ARG;
    return guarded(ctx, [&] {
        sbe_ctx &c = *ctx;
        const int n = c.n;
        if (!below_gap && (occupied_bands <= 0 || occupied_bands >= n))
            fail(SBE_ERR_ARG, "occupiedBands must be in (0, %d)", n);
        // all eigenvalues of H0(k) at A = 0 without kGlobalShift (code/src/main/runSBE.cpp:130-155)
        HIPCHK(launch_twiddles(c.geom, nullptr, 0, c.tw, c.stream));
        HIPCHK(launch_phase_sum(c.geom, c.pbufs(), 2, 0, 0, 0, c.mesh, c.stream));
        double *eig = c.st[8];  // scratch: n*nk doubles fit into a state buffer
        InitArgs a{};
        a.mesh = c.mesh;
        a.ops = 2;
        a.eig = eig;
        a.nk = c.nk;
        a.n = n;
        a.P = c.P;
        a.sweeps = c.prm.sweep_count;
        HIPCHK(launch_init_state(c.N, c.lr ? c.lr : sbe_ctx::default_lr(c.N), a, c.stream));
        std::vector<double> ev((size_t)n * c.nk);
        HIPCHK(hipMemcpyAsync(ev.data(), eig, ev.size() * sizeof(double), hipMemcpyDeviceToHost, c.stream));
        HIPCHK(hipStreamSynchronize(c.stream));
        const double kT = c.prm.gs_temp;
        if (below_gap) {
            std::vector<double> bmin(n, INFINITY), bmax(n, -INFINITY), tmp(n);
            for (long k = 0; k < c.nk; k++) {
                std::copy(ev.begin() + k * n, ev.begin() + (k + 1) * n, tmp.begin());
                std::sort(tmp.begin(), tmp.end());
                for (int i = 0; i < n; i++) {
                    bmin[i] = std::min(bmin[i], tmp[i]);
                    bmax[i] = std::max(bmax[i], tmp[i]);
                }
            }
            int gaps = 0;
            double f = 0;
            for (int i = 0; i + 1 < n; i++)
                if (bmax[i] + 3 * kT < bmin[i + 1]) {
                    gaps++;
                    f = 0.5 * (bmax[i] + bmin[i + 1]);
                }
            *mu = f;
            if (gaps != 1) fail(SBE_ERR_PHYSICS, "Detected %d != 1 band gaps", gaps);
        } else {
            std::sort(ev.begin(), ev.end());
            size_t hi = (size_t)occupied_bands * c.nk - 1;
            double homo = ev[hi], lumo = ev[hi + 1];
            *mu = 0.5 * (homo + lumo);
            if (homo + 3 * kT > lumo) fail(SBE_ERR_PHYSICS, "Fermi level calculation for metals is not implemented");
        }
    });
}

static void initial_state_dev(sbe_ctx &c, double mu, double *dst)
{
    HIPCHK(launch_twiddles(c.geom, nullptr, 1, c.tw, c.stream));
    HIPCHK(launch_phase_sum(c.geom, c.pbufs(), 2, 0, 0, 0, c.mesh, c.stream));
    InitArgs a{};
    a.mesh = c.mesh;
    a.ops = 2;
    a.state = dst;
    a.nk = c.nk;
    a.n = c.n;
    a.P = c.P;
    a.sweeps = c.prm.sweep_count;
    a.mu = mu;
    a.kT = c.prm.gs_temp;
    HIPCHK(launch_init_state(c.N, c.lr ? c.lr : sbe_ctx::default_lr(c.N), a, c.stream));
}

int sbe_initial_state(sbe_ctx *ctx, double mu, double *state)
{
    if (!ctx || !state) return SBE_ERR_ARG;
    return guarded(ctx, [&] {
        sbe_ctx &c = *ctx;
        initial_state_dev(c, mu, c.st[8]);
        HIPCHK(hipMemcpyAsync(state, c.st[8], 16 * c.S, hipMemcpyDeviceToHost, c.stream));
        HIPCHK(hipStreamSynchronize(c.stream));
    });
}

int sbe_rhs(sbe_ctx *ctx, const double *state, double *dsdt, double t)
{
    if (!ctx || !state || !dsdt) return SBE_ERR_ARG;
    return guarded(ctx, [&] {
        sbe_ctx &c = *ctx;
        HIPCHK(hipMemcpyAsync(c.st[8], state, 16 * c.S, hipMemcpyHostToDevice, c.stream));
        c.rhs_dev(c.st[8], c.st[9], t, c.stream);
        HIPCHK(hipMemcpyAsync(dsdt, c.st[9], 16 * c.S, hipMemcpyDeviceToHost, c.stream));
        HIPCHK(hipStreamSynchronize(c.stream));
    });
}

int sbe_rhs_device(sbe_ctx *ctx, const double *d_state, double *d_dsdt, double t, void *stream)
{
    if (!ctx || !d_state || !d_dsdt) return SBE_ERR_ARG;
    return guarded(ctx, [&] {
        hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
        ctx->rhs_dev(d_state, d_dsdt, t, s);
    });
}

int sbe_observe(sbe_ctx *ctx, const double *state, double j[3])
{
    if (!ctx || !state || !j) return SBE_ERR_ARG;
    return guarded(ctx, [&] {
        sbe_ctx &c = *ctx;
        if (c.jbuf_cap < 1) {
            if (c.jbuf) HIPCHK(hipFree(c.jbuf));
            c.jbuf = dalloc<double>(3);
            c.jbuf_cap = 1;
        }
        HIPCHK(hipMemcpyAsync(c.st[8], state, 16 * c.S, hipMemcpyHostToDevice, c.stream));
        c.observe_dev(c.st[8], c.jbuf, c.stream);
        HIPCHK(hipMemcpyAsync(j, c.jbuf, 3 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
        HIPCHK(hipStreamSynchronize(c.stream));
    });
}

long sbe_integrate(sbe_ctx *ctx, double *state, int output_count, double *t_out, double *j_out, sbe_steplog *log,
                   const sbe_steplog *replay, long *rhs_calls)
{
    if (!ctx || !state) return SBE_ERR_ARG;
    long steps = 0;
    int rc = guarded(ctx, [&] {
        sbe_ctx &c = *ctx;
        if (output_count < 2) fail(SBE_ERR_ARG, "output_count must be >= 2");
        if (c.jbuf_cap < output_count) {
            if (c.jbuf) HIPCHK(hipFree(c.jbuf));
            c.jbuf = dalloc<double>(3 * (size_t)output_count);
            c.jbuf_cap = output_count;
        }
        c.rhs_calls = 0;
        if (log) log->count = 0;
        HIPCHK(hipMemcpyAsync(c.x(), state, 16 * c.S, hipMemcpyHostToDevice, c.stream));
        const double t0 = c.t_start(), t1 = c.t_end();
        const double dt_out = (t1 - t0) / (output_count - 1);
        const double t_end = t1 + dt_out / 2;
        double time = t0, dt = dt_out;
        long step = 0;
        int out = 0, rp = 0;
        std::vector<double> tv;
        while (less_eq_with_sign(time + dt_out, t_end)) {
            if (out < output_count) {
                tv.push_back(time);
                c.observe_dev(c.x(), c.jbuf + 3 * out, c.stream);
            }
            out++;
            double start = time, end = time + dt_out;
            bool first = true;
            int fails = 0;
            if (replay) {
                // replay: the logged accepted attempts whose start lies in this
                // output interval, with the logged (t, dt)
                c.rhs_dev(c.x(), c.st[1], start, c.stream);
                const double tol = 1e-9 * std::max(1.0, std::fabs(end));
                for (;;) {
                    while (rp < replay->count && rp < replay->capacity && !replay->accepted[rp]) rp++;
                    if (rp >= replay->count || rp >= replay->capacity) break;
                    if (replay->t[rp] >= end - tol) break;
                    start = replay->t[rp];
                    dt = replay->dt[rp++];
                    double err = c.attempt(start, dt);
                    log_push(log, start, dt, err, 1);
                    c.accept();
                    steps++;
                }
                if (steps == 0 || (rp >= replay->count && less_with_sign(start + dt, end - tol)))
                    fail(SBE_ERR_REPLAY, "replay log does not cover the interval ending at t=%g", end);
            }
            while (!replay && less_with_sign(start, end)) {
                if (less_with_sign(end, start + dt)) dt = end - start;
                if (first) {
                    c.rhs_dev(c.x(), c.st[1], start, c.stream);
                    first = false;
                }
                for (;;) {
                    double err = c.attempt(start, dt);
                    if (!(err <= 1.0)) {
                        log_push(log, start, dt, err, 0);
                        if (std::isnan(err)) fail(SBE_ERR_STEP, "error norm is NaN at t=%g", start);
                        double f = 0.9 * std::pow(err, -1.0 / 3.0);
                        dt *= f > 0.2 ? f : 0.2;
                        if (++fails >= 500)
                            fail(SBE_ERR_STEP, "Max number of iterations exceeded (500). A new step size was not found.");
                        continue;
                    }
                    log_push(log, start, dt, err, 1);
                    start += dt;
                    if (err < 0.5) {
                        double e = std::max(std::pow(5.0, -5.0), err);
                        dt *= 0.9 * std::pow(e, -1.0 / 5.0);
                    }
                    break;
                }
                fails = 0;
                c.accept();
                steps++;
            }
            step++;
            time = t0 + (double)step * dt_out;
        }
        if (out < output_count) {
            tv.push_back(time);
            c.observe_dev(c.x(), c.jbuf + 3 * out, c.stream);
        }
        out++;
        HIPCHK(hipMemcpyAsync(state, c.x(), 16 * c.S, hipMemcpyDeviceToHost, c.stream));
        std::vector<double> jh(3 * (size_t)output_count, 0.0);
        int nobs = std::min(out, output_count);
        HIPCHK(hipMemcpyAsync(jh.data(), c.jbuf, 3 * sizeof(double) * nobs, hipMemcpyDeviceToHost, c.stream));
        HIPCHK(hipStreamSynchronize(c.stream));
        if (j_out) std::memcpy(j_out, jh.data(), 3 * sizeof(double) * output_count);
        if (t_out)
            for (int i = 0; i < output_count; i++) t_out[i] = i < (int)tv.size() ? tv[i] : 0.0;
    });
    if (rhs_calls) *rhs_calls = ctx->rhs_calls;
    return rc ? rc : steps;
}

int sbe_hk_mesh(sbe_ctx *ctx, const double A[3], const double E[3], double *out)
{
    if (!ctx || !A || !E || !out) return SBE_ERR_ARG;
    return guarded(ctx, [&] {
        sbe_ctx &c = *ctx;
        // A enters through the state's special slots
        std::vector<double> head(16, 0.0);
        for (int d = 0; d < 3; d++) head[2 * d] = A[d];
        HIPCHK(hipMemcpyAsync(c.st[8], head.data(), 16 * sizeof(double), hipMemcpyHostToDevice, c.stream));
        HIPCHK(launch_twiddles(c.geom, c.st[8], 1, c.tw, c.stream));
        HIPCHK(launch_phase_sum(c.geom, c.pbufs(), 2, E[0], E[1], E[2], c.mesh, c.stream));
        HIPCHK(hipMemcpyAsync(out, c.mesh, 16 * c.nk * 2 * c.nn, hipMemcpyDeviceToHost, c.stream));
        HIPCHK(hipStreamSynchronize(c.stream));
    });
}

int sbe_jacobi_batch(int n, int sweeps, long count, const double *H, double *U, double *eig)
{
    if (n < 1 || n > 16 || count < 1 || !H || sweeps < 0) {
        g_create_err = "sbe_jacobi_batch: invalid argument";
        return SBE_ERR_ARG;
    }
    double *dH = nullptr, *dU = nullptr, *de = nullptr;
    int rc = guarded(nullptr, [&] {
        dH = dalloc<double>(2 * (size_t)count * n * n);
        dU = dalloc<double>(2 * (size_t)count * n * n);
        de = dalloc<double>((size_t)count * n);
        HIPCHK(hipMemcpy(dH, H, 16 * (size_t)count * n * n, hipMemcpyHostToDevice));
        InitArgs a{};
        a.mesh = dH;
        a.ops = 1;
        a.eig = de;
        a.U = dU;
        a.nk = count;
        a.n = n;
        a.P = 8 * ((n * n + 7) / 8);
        a.sweeps = sweeps;
        int N = n <= 2 ? 2 : n <= 4 ? 4 : n <= 8 ? 8 : 16;
        HIPCHK(launch_init_state(N, sbe_ctx::default_lr(N), a, nullptr));
        HIPCHK(hipDeviceSynchronize());
        if (U) HIPCHK(hipMemcpy(U, dU, 16 * (size_t)count * n * n, hipMemcpyDeviceToHost));
        if (eig) HIPCHK(hipMemcpy(eig, de, 8 * (size_t)count * n, hipMemcpyDeviceToHost));
    });
    if (dH) (void)hipFree(dH);
    if (dU) (void)hipFree(dU);
    if (de) (void)hipFree(de);
    return rc;
}

int sbe_stage_tuning(const sbe_ctx *ctx, int *lanes_per_row, int *count, int *variants, double *ms)
{
    if (!ctx || !lanes_per_row) return SBE_ERR_ARG;
    *lanes_per_row = ctx->lr;
    if (count) *count = ctx->tune_count;
    for (int v = 0; v < ctx->tune_count; v++) {
        if (variants) variants[v] = ctx->tune_lr[v];
        if (ms) ms[v] = ctx->tune_ms[v];
    }
    return SBE_OK;
}

int sbe_bench_steps(sbe_ctx *ctx, const double *state, double t, double dt, int warmup, int steps, double *ms,
                    long *rhs, double *stage_ms, double *phase_ms)
{
    if (!ctx || !state || steps < 1 || warmup < 0 || !(dt > 0)) return SBE_ERR_ARG;
    return guarded(ctx, [&] {
        sbe_ctx &c = *ctx;
        HIPCHK(hipMemcpyAsync(c.x(), state, 16 * c.S, hipMemcpyHostToDevice, c.stream));
        c.rhs_dev(c.x(), c.st[1], t, c.stream);
        // the controller of integrate_adaptive without output truncation
        auto run = [&](int count) {
            for (int i = 0; i < count; i++) {
                double err = c.attempt(t, dt);
                if (!(err <= 1.0)) {
                    if (std::isnan(err)) fail(SBE_ERR_STEP, "error norm is NaN at t=%g", t);
                    dt *= std::max(0.9 * std::pow(err, -1.0 / 3.0), 0.2);
                    continue;
                }
                c.accept();
                t += dt;
                if (err < 0.5) dt *= 0.9 * std::pow(std::max(std::pow(5.0, -5.0), err), -1.0 / 5.0);
            }
        };
        run(warmup);
        HIPCHK(hipStreamSynchronize(c.stream));
        long r0 = c.rhs_calls;
        HIPCHK(hipEventRecord(c.ev[2], c.stream));
        run(steps);
        HIPCHK(hipEventRecord(c.ev[3], c.stream));
        HIPCHK(hipEventSynchronize(c.ev[3]));
        float tot = 0;
        HIPCHK(hipEventElapsedTime(&tot, c.ev[2], c.ev[3]));
        if (ms) *ms = tot;
        if (rhs) *rhs = c.rhs_calls - r0;
        // optional second pass: per-kernel durations from events around each launch
        if (stage_ms || phase_ms) {
            c.timing = true;
            c.evused = 0;
            run(std::min(steps, 20));
            HIPCHK(hipStreamSynchronize(c.stream));
            c.timing = false;
            double sm = 0, pm = 0;
            long nr = 0;
            for (size_t e = 0; e + 2 < c.evused; e += 3) {
                float a = 0, b = 0;
                HIPCHK(hipEventElapsedTime(&a, c.evpool[e], c.evpool[e + 1]));
                HIPCHK(hipEventElapsedTime(&b, c.evpool[e + 1], c.evpool[e + 2]));
                pm += a;
                sm += b;
                nr++;
            }
            if (stage_ms) *stage_ms = nr ? sm / nr : 0;
            if (phase_ms) *phase_ms = nr ? pm / nr : 0;
        }
    });
}

}  // extern "C"
