// Host-side context of libsbewh_hip.so shared by the single-instance C ABI
// (sbe_api.cpp) and the batched scan (scan_api.cpp): error handling, pulses,
// the Dormand-Prince tableau, the per-instance device context with its
// single-instance controller pieces, and the context set-up.
#pragma once
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
#include "dopri5.hpp"
#include "pulse.hpp"

using namespace sbe;

namespace sbe_host {

inline thread_local std::string g_create_err;

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

[[noreturn]] inline void fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    throw ArgFail{code, buf};
}

inline int smooth5(int N)
{
    int n = N - 1, r;
    do {
        r = ++n;
        for (int p : {2, 3, 5})
            while (r % p == 0) r /= p;
    } while (r != 1);
    return n;
}

}  // namespace sbe_host

using namespace sbe_host;

struct sbe_ctx {
    std::string err;
    int n = 0, N = 0, dim = 0, P = 0, nn = 0;
    long nk = 0, S = 0;
    sbe_params prm{};
    std::vector<SubPulse> pulses;
    PhaseGeom geom{};
    KOrderH korder{};
    std::vector<int> kgrid;    // [k][3] grid indices of state k (Fourier-block order)
    int obs_ops = 0;
    hipStream_t stream = nullptr;
    int device = 0;
    // device buffers
    double *Hb = nullptr, *Db = nullptr, *dHb = nullptr;
    int *kmap = nullptr;
    double *tw = nullptr, *y1 = nullptr, *y2 = nullptr;
    double *mesh = nullptr;    // full phase sum {H0, H_E} (initial state, Fermi level, hk_mesh)
    double *Y = nullptr;       // partial phase sum of the RHS operators (all axes but the last)
    double *hm = nullptr;      // model box with the operator merge of the current RHS (dim >= 2)
    double *Yobs = nullptr;    // partial phase sum of the observer operators
    int *kphase = nullptr;     // per k: (rest(k) W_L, i_L(k) W_L)
    int WL = 1;
    long twL_off = 0;
    double *st[10] = {};  // x, k1..k7, scratch, xn
    double *ucache = nullptr;  // [k][n*n] eigenvectors of the last RHS (warm starts)
    // per-stage phase data of one attempt (twiddles, merged model box, partial
    // phase sum, contraction scratch), SET_MAX copies each
    double *tw6 = nullptr, *hm6 = nullptr, *Y6 = nullptr, *y16 = nullptr;
    long tw_stride = 0, hm_stride = 0, Y_stride = 0, y1_stride = 0;
    long fuse_below_blocks = 512;  // stage grids smaller than this run an attempt's stages in one launch
    long set_min_blocks = 512;  // block target of a set's contraction launch (best on the 150 x 150 2-D model)
    bool ucache_valid = false;
    bool resident_used = false;  // the last integration ran on the resident integrator
    double *err_partial = nullptr;
    unsigned *err_ticket = nullptr;              // ticket of the fused attempt's last block (zero between launches)
    unsigned long long *err_host = nullptr;      // pinned, mapped, fine-grained
    unsigned long long *err_host_dev = nullptr;  // its device address
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

    // Jacobi sweep bound of one diagonalisation.  A call that starts from the
    // previous call's eigenvectors at the same k (warm) runs at most the
    // configured sweep count; one that starts from the identity runs at most
    // max(sweep count, COLD_SWEEPS), so that the first RHS of a run (and the
    // ground state) begins the warm chain from a converged decomposition (16
    // bands need 6 sweeps of the parallel ordering from the identity).  The
    // iteration ends after the first sweep without a rotation, so a bound
    // above what convergence needs changes nothing.  SBE_FLAG_COLD_JACOBI
    // keeps the reference's fixed count on every call.
    static constexpr int COLD_SWEEPS = 12;
    int sweeps_for(int warm) const
    {
        if (warm || (prm.flags & SBE_FLAG_COLD_JACOBI)) return prm.sweep_count;
        return std::max(prm.sweep_count, COLD_SWEEPS);
    }

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

    void E_at(double t, double E[3]) const { pulse_E(pulses.data(), (int)pulses.size(), t, E); }
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
    PhaseLast last(const double *Yp) const { return PhaseLast{Yp, tw + 2 * twL_off, kphase, WL}; }
    // the merged model box of the RHS (for dim == 1 it is the partial phase sum itself)
    double *merged() const { return dim == 1 ? Y : hm; }
    ModelMerge merge_of(const double E[3]) const
    {
        ModelMerge mm{};
        mm.Hb = Hb;
        mm.Db = Db;
        for (int d = 0; d < 3; d++) mm.E[d] = d < dim ? E[d] : 0.0;
        mm.nbox = (long)geom.W[0] * geom.W[1] * geom.W[2];
        mm.nn = nn;
        mm.dim = dim;
        mm.out = merged();
        return mm;
    }

    // dxdt = f(rho_in, t) on device with rho_in = in.x + sum_j in.c[j] in.k[j];
    // xout (optional) receives rho_in
    void rhs_dev(const StageInput &in, double *dxdt, double t, hipStream_t s, double *xout = nullptr)
    {
        double E[3];
        E_at(t, E);
        if (timing) HIPCHK(hipEventRecord(next_event(), s));
        HIPCHK(launch_phase_prep(geom, &in, 1, tw, merge_of(E), s));
        HIPCHK(launch_phase_partial(geom, pbufs(), 2, E[0], E[1], E[2], Y, merged(), s));
        if (timing) HIPCHK(hipEventRecord(next_event(), s));
        stage_dev(in, dxdt, E, last(Y), s, xout);
    }

    // the stage kernel alone, on a partial phase sum formed beforehand
    void stage_dev(const StageInput &in, double *dxdt, const double E[3], const PhaseLast &ph, hipStream_t s,
                   double *xout, const double *err_e = nullptr, double dt = 0)
    {
        StageArgs a = stage_args(in, dxdt, E, ph, xout, err_e, dt);
        if (!lr) tune_stage(a, s);
        a.warm = (ucache_valid && !(prm.flags & SBE_FLAG_COLD_JACOBI)) ? 1 : 0;
        a.sweeps = sweeps_for(a.warm);
        HIPCHK(launch_rhs_stage(N, lr, a, s));
        if (a.dephase) ucache_valid = true;
        if (timing) HIPCHK(hipEventRecord(next_event(), s));
        rhs_calls++;
    }

    StageArgs stage_args(const StageInput &in, double *dxdt, const double E[3], const PhaseLast &ph, double *xout,
                         const double *err_e, double dt) const
    {
        StageArgs a{};
        if (err_e) {
            // the attempt's error norm, per block, fused into this (last) stage
            a.err_partial = err_partial;
            for (int j = 0; j < 6; j++) a.err_e[j] = err_e[j];
            a.err_ea = prm.eps_abs;
            a.err_er = prm.eps_rel;
            a.err_dt = dt;
        }
        a.in = in;
        a.xout = xout;
        a.dxdt = dxdt;
        a.ph = ph;
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
        return a;
    }

    // The phase sums of the m stage inputs of one attempt (fields E[q]), all
    // formed before the first stage runs: they depend on rho only through the
    // special slots A, and the A slot of a derivative that an earlier stage of
    // this attempt computes (outs[r]) is -E[r], known here; derivatives that
    // are no stage's output (k1) are read from memory.
    void phase_set(const StageInput *ins, const double (*E)[3], const double *const *outs, int m, hipStream_t s)
    {
        // the prep launch goes out before the contraction is planned: the GPU
        // is idle from the error norm's arrival until it starts
        PrepSet ps{};
        ps.m = m;
        for (int q = 0; q < m; q++) {
            TwiddleEntry &e = ps.e[q];
            e.g = geom;
            e.in = ins[q];
            e.has_a = 1;
            e.use_shift = 1;
            e.tw = tw6 + q * tw_stride;
            e.mm = merge_of(E[q]);
            e.mm.out = dim == 1 ? Y6 + q * Y_stride : hm6 + q * hm_stride;
            AOverride &ao = ps.a[q];
            for (int j = 0; j < ins[q].nt; j++)
                for (int r = 0; r < q; r++)
                    if (ins[q].k[j] == outs[r]) {
                        ao.mask |= 1 << j;
                        for (int d = 0; d < 3; d++) ao.v[j][d] = d < dim ? -E[r][d] : 0.0;
                    }
            ps.max_count = std::max(ps.max_count, (int)(twiddle_offset(geom, 3) + (e.mm.out ? e.mm.nbox * 2 * nn : 0)));
        }
        HIPCHK(launch_prep_set(ps, s));
        ContractSet cs[2]{};
        int ns = 0;
        for (int q = 0; q < m; q++) {
            PhaseBuffers pb{Hb, Db, dHb, kmap, tw6 + q * tw_stride, y16 + q * y1_stride, nullptr};
            ContractStage st[2];
            // the set's stages share one launch: larger row tiles reuse the inputs
            ns = phase_partial_plan(geom, pb, 2, E[q][0], E[q][1], E[q][2], Y6 + q * Y_stride, ps.e[q].mm.out, st,
                                    set_min_blocks / m);
            for (int d = 0; d < ns; d++) {
                cs[d].st[q] = st[d];
                cs[d].block0[q + 1] = cs[d].block0[q] + st[d].blocks_x * st[d].blocks_y;
                cs[d].m = m;
            }
        }
        for (int d = 0; d < ns; d++) HIPCHK(launch_contract_set(cs[d], s));
    }
    PhaseLast last_set(int q) const
    {
        return PhaseLast{Y6 + q * Y_stride, tw6 + q * tw_stride + 2 * twL_off, kphase, WL};
    }

    void rhs_dev(const double *x_in, double *dxdt, double t, hipStream_t s)
    {
        StageInput in{};
        in.x = x_in;
        rhs_dev(in, dxdt, t, s);
    }

    void observe_dev(const double *xs, double *jdst, hipStream_t s)
    {
        StageInput in{};
        in.x = xs;
        HIPCHK(launch_twiddles(geom, &in, 1, tw, s));
        HIPCHK(launch_phase_partial(geom, pbufs(), obs_ops, 0, 0, 0, Yobs, nullptr, s));
        ObsArgs o{};
        o.x = xs;
        o.ph = last(Yobs);
        o.nk = nk;
        o.n = n;
        o.P = P;
        o.dim = dim;
        o.partial = partial;
        o.j_out = jdst;
        o.pstride = 3;
        HIPCHK(launch_observe_j(o, s));
    }

    // one Dormand-Prince attempt from (x, k1, t, dt) -> xn, k7; returns the error norm.
    // Stage inputs x + dt sum_j b_sj k_j are formed inside the stage kernel.
    double attempt(double t, double dt)
    {
        double *X = st[0], *k1 = st[1], *k2 = st[2], *k3 = st[3], *k4 = st[4], *k5 = st[5], *k6 = st[6],
               *k7 = st[7], *xn = st[9];
        hipStream_t s = stream;
        auto input = [&](std::initializer_list<const double *> ks, std::initializer_list<double> cs) {
            StageInput in{};
            in.x = X;
            int j = 0;
            for (const double *k : ks) in.k[j++] = k;
            in.nt = j;
            j = 0;
            for (double c : cs) in.c[j++] = dt * c;
            return in;
        };
        const StageInput ins[6] = {input({k1}, {B21}),
                                   input({k1, k2}, {B31, B32}),
                                   input({k1, k2, k3}, {B41, B42, B43}),
                                   input({k1, k2, k3, k4}, {B51, B52, B53, B54}),
                                   input({k1, k2, k3, k4, k5}, {B61, B62, B63, B64, B65}),
                                   input({k1, k3, k4, k5, k6}, {C1, C3, C4, C5, C6})};
        double *const outs[6] = {k2, k3, k4, k5, k6, k7};
        const double ts[6] = {t + dt * A2, t + dt * A3, t + dt * A4, t + dt * A5, t + dt, t + dt};
        double E[6][3];
        for (int q = 0; q < 6; q++) E_at(ts[q], E[q]);
        // error coefficients of k1, k3, k4, k5, k6, k7
        const double err_e[6] = {dt * DC1, dt * DC3, dt * DC4, dt * DC5, dt * DC6, dt * DC7};
        const long blocks = lr ? (nk + stage_k_per_block(N, lr) - 1) / stage_k_per_block(N, lr) : 0;
        const bool fused = lr && !timing && blocks < fuse_below_blocks;
        // the last kernel of the attempt overwrites this with the norm
        *reinterpret_cast<volatile unsigned long long *>(err_host) = ERR_PENDING;
        // the six phase sums first, then the six stage kernels
        if (timing) HIPCHK(hipEventRecord(next_event(), s));
        phase_set(ins, E, outs, 6, s);
        if (timing) HIPCHK(hipEventRecord(next_event(), s));
        if (fused) {
            // a grid this small is launch-bound: the six stages in one launch,
            // whose last block also reduces the error norm
            StageSet ss{};
            ss.m = 6;
            ss.ticket = err_ticket;
            ss.err_out = err_host_dev;
            for (int q = 0; q < 6; q++) {
                ss.st[q] = stage_args(ins[q], outs[q], E[q], last_set(q), q == 5 ? xn : nullptr,
                                      q == 5 ? err_e : nullptr, dt);
                ss.st[q].warm = (ucache_valid && !(prm.flags & SBE_FLAG_COLD_JACOBI)) ? 1 : 0;
                ss.st[q].sweeps = sweeps_for(ss.st[q].warm);
                if (ss.st[q].dephase) ucache_valid = true;
            }
            HIPCHK(launch_rhs_stage_set(N, lr, ss, s));
            rhs_calls += 6;
        } else {
            for (int q = 0; q < 6; q++) {
                if (timing && q > 0) {
                    HIPCHK(hipEventRecord(next_event(), s));
                    HIPCHK(hipEventRecord(next_event(), s));
                }
                stage_dev(ins[q], outs[q], E[q], last_set(q), s, q == 5 ? xn : nullptr, q == 5 ? err_e : nullptr,
                          dt);
            }
        }
        // the last stage kernel left one error-norm partial per block
        // (the stage layout, and with it the block count, is known only after
        // the first stage launch when the tuner has not run yet)
        if (!fused)
            HIPCHK(launch_error_final(err_partial, (nk + stage_k_per_block(N, lr) - 1) / stage_k_per_block(N, lr),
                                      err_host_dev, s));
        return wait_error(s);
    }

    // The error norm's final kernel stores the norm into mapped host memory;
    // poll it instead of waiting for the stream to drain (the next attempt's
    // launches then queue behind the retiring kernel).  A stream that has
    // completed, or failed, without the value arriving is an error.
    static constexpr unsigned long long ERR_PENDING = ~0ull;  // a negative NaN: never a max of |.| terms
    double wait_error(hipStream_t s)
    {
        const volatile unsigned long long *p = err_host;
        unsigned long long b;
        for (unsigned it = 1;; it++) {
            b = *p;
            if (b != ERR_PENDING) break;
            if ((it & 4095) == 0) {
                const hipError_t q = hipStreamQuery(s);
                if (q == hipErrorNotReady) continue;
                HIPCHK(q);
                b = *p;
                if (b == ERR_PENDING) {
                    // a fused launch that ended without its last block leaves
                    // the ticket counter non-zero: clear it for the next attempt
                    if (err_ticket) (void)hipMemsetAsync(err_ticket, 0, sizeof(unsigned), s);
                    throw HipFail{hipErrorUnknown, "error norm not delivered"};
                }
                break;
            }
        }
        double v;
        std::memcpy(&v, &b, sizeof v);
        return v;
    }

    void accept()
    {
        std::swap(st[0], st[9]);  // x <- xn
        std::swap(st[1], st[7]);  // k1 <- k7 (FSAL)
    }

    ~sbe_ctx()
    {
        double *ptrs[] = {Hb, Db, dHb, tw, y1, y2, mesh, Y, hm, Yobs, partial, jbuf, ucache, tw6, hm6, Y6, y16};
        for (double *p : ptrs)
            if (p) (void)hipFree(p);
        for (double *p : st)
            if (p) (void)hipFree(p);
        if (kmap) (void)hipFree(kmap);
        if (kphase) (void)hipFree(kphase);
        if (err_partial) (void)hipFree(err_partial);
        if (err_ticket) (void)hipFree(err_ticket);
        if (err_host) (void)hipHostFree(err_host);
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
        for (auto &e : evpool) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace sbe_host {

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

inline void set_pulses(sbe_ctx *c, const sbe_subpulse *p, int np)
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

// member_of_scan: the context is one instance of a batched scan, which owns
// the stream and the host-side read-back buffers
inline void build(sbe_ctx *c, const sbe_model *m, const sbe_grid *g, const sbe_params *prm, const sbe_subpulse *pul,
                  int np, bool member_of_scan = false)
{
    if (!m || !g || !prm) fail(SBE_ERR_ARG, "null argument");
    if (m->num_wann < 1 || m->num_wann > 16) fail(SBE_ERR_ARG, "num_wann must be in [1, 16] (got %d)", m->num_wann);
    if (m->num_cells < 1 || !m->cells || !m->H || !m->D) fail(SBE_ERR_ARG, "empty model");
    if (g->dim < 1 || g->dim > 3) fail(SBE_ERR_ARG, "dim must be 1, 2 or 3");
    if (prm->sweep_count < 0 || prm->sweep_count > 50) fail(SBE_ERR_ARG, "sweep_count must be in [0, 50]");
    if (prm->flags & ~(SBE_FLAG_COLD_JACOBI | SBE_FLAG_GPU_LAYOUT | SBE_FLAG_DEVICE_CONTROLLER))
        fail(SBE_ERR_ARG, "unknown flags 0x%x", prm->flags);
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
    // reference CPU layout: rho_k padded to a multiple of 8 complex; GPU layout: unpadded
    c->P = (prm->flags & SBE_FLAG_GPU_LAYOUT) ? nn : 8 * ((nn + 7) / 8);
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
            c->kgrid.push_back((int)i0);
            c->kgrid.push_back((int)i1);
            c->kgrid.push_back((int)i2);
        }
    }
    if (c->nk > 0x7fffffffL) fail(SBE_ERR_ARG, "too many k-points");
    // device memory
    HIPCHK(hipGetDevice(&c->device));
    if (!member_of_scan) {
        HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        for (auto &e : c->ev) HIPCHK(hipEventCreate(&e));
    }
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
    ysz = std::max(ysz, phase_partial_scratch(G, c->obs_ops));
    c->y1 = dalloc<double>(2 * ysz);
    c->y2 = dalloc<double>(2 * ysz);
    c->mesh = dalloc<double>(2 * c->nk * 2 * nn);
    c->Y = dalloc<double>(2 * phase_partial_size(G, 2));
    if (g->dim > 1) c->hm = dalloc<double>(2 * nbox * 2 * nn);
    c->Yobs = dalloc<double>(2 * phase_partial_size(G, c->obs_ops));
    {
        // per k: Y row offset of rest(k) and twiddle offset of i_L(k), L = dim - 1
        const int L = g->dim - 1;
        c->WL = G.W[L];
        c->twL_off = twiddle_offset(G, L);
        const long rowlen = G.W[L];
        if ((long)c->nk / G.Nk[L] * rowlen > 0x7fffffffL) fail(SBE_ERR_ARG, "partial phase sum too large");
        std::vector<int> kph(2 * c->nk);
        for (long gi = 0; gi < c->nk; gi++) {
            long k = kmap[gi];
            kph[2 * k] = (int)((gi / G.Nk[L]) * rowlen);
            kph[2 * k + 1] = (int)((gi % G.Nk[L]) * G.W[L]);
        }
        c->kphase = dalloc<int>(2 * c->nk);
        HIPCHK(hipMemcpy(c->kphase, kph.data(), kph.size() * sizeof(int), hipMemcpyHostToDevice));
    }
    for (auto &p : c->st) {
        p = dalloc<double>(2 * c->S);
        HIPCHK(hipMemset(p, 0, 16 * c->S));
    }
    c->ucache = dalloc<double>(2 * c->nk * nn);
    if (!member_of_scan) {
        c->tw_stride = 2 * ntw;
        c->hm_stride = g->dim > 1 ? 2 * nbox * 2 * nn : 0;
        c->Y_stride = 2 * phase_partial_size(G, 2);
        c->y1_stride = 2 * phase_partial_scratch(G, 2);
        c->tw6 = dalloc<double>(SET_MAX * c->tw_stride);
        if (g->dim > 1) c->hm6 = dalloc<double>(SET_MAX * c->hm_stride);
        c->Y6 = dalloc<double>(SET_MAX * c->Y_stride);
        c->y16 = dalloc<double>(SET_MAX * c->y1_stride);
    }
    c->err_partial = dalloc<double>(std::max(1024L, c->nk));  // one per stage block (>= 1 k-point each)
    c->err_ticket = dalloc<unsigned>(1);
    HIPCHK(hipMemset(c->err_ticket, 0, sizeof(unsigned)));
    // SBEWH_FUSE_BELOW overrides the stage-grid size under which an attempt's
    // stages share one launch (0: never, as for large grids)
    if (const char *env = std::getenv("SBEWH_FUSE_BELOW")) c->fuse_below_blocks = std::atol(env);
    if (!member_of_scan) {
        HIPCHK(hipHostMalloc((void **)&c->err_host, sizeof(unsigned long long),
                             hipHostMallocMapped | hipHostMallocCoherent));
        HIPCHK(hipHostGetDevicePointer((void **)&c->err_host_dev, c->err_host, 0));
    }
    c->partial = dalloc<double>(3 * c->nk);
}

// Fermi-Dirac ground state rho_k = f(H0(k)) of the context into dst (device)
void initial_state_dev(sbe_ctx &c, double mu, double *dst, hipStream_t s);


inline void log_push(sbe_steplog *log, double t, double dt, double err, int acc)
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

}  // namespace sbe_host
