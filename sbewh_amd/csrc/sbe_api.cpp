// Single-instance C ABI of libsbewh_hip.so (include/sbewh_hip.h): context
// life cycle, pulses, ground state, RHS, observer and the device-resident
// Dormand-Prince controller.
#include "observer.hpp"

using namespace sbe_host;

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
int sbe_matrix_stride(const sbe_ctx *ctx) { return ctx ? ctx->P : SBE_ERR_ARG; }

int sbe_k_index_map(const sbe_ctx *ctx, int *grid)
{
    if (!ctx || !grid) return SBE_ERR_ARG;
    std::memcpy(grid, ctx->kgrid.data(), ctx->kgrid.size() * sizeof(int));
    return SBE_OK;
}

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

int sbe_pulse_field(const sbe_subpulse *pulses, int num_pulses, double t, double E[3])
{
    if (!E) return SBE_ERR_ARG;
    sbe_ctx tmp;
    int rc = guarded(nullptr, [&] { set_pulses(&tmp, pulses, num_pulses); });
    if (rc) return rc;
    tmp.E_at(t, E);
    return SBE_OK;
}

int sbe_pulse_window(const sbe_subpulse *pulses, int num_pulses, double *t0, double *t1)
{
    if (!t0 || !t1) return SBE_ERR_ARG;
    sbe_ctx tmp;
    int rc = guarded(nullptr, [&] { set_pulses(&tmp, pulses, num_pulses); });
    if (rc) return rc;
    *t0 = tmp.t_start();
    *t1 = tmp.t_end();
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
        a.sweeps = c.sweeps_for(0);
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

}  // extern "C"

void sbe_host::initial_state_dev(sbe_ctx &c, double mu, double *dst, hipStream_t s)
{
    HIPCHK(launch_twiddles(c.geom, nullptr, 1, c.tw, s));
    HIPCHK(launch_phase_sum(c.geom, c.pbufs(), 2, 0, 0, 0, c.mesh, s));
    InitArgs a{};
    a.mesh = c.mesh;
    a.ops = 2;
    a.state = dst;
    a.nk = c.nk;
    a.n = c.n;
    a.P = c.P;
    a.sweeps = c.sweeps_for(0);
    a.mu = mu;
    a.kT = c.prm.gs_temp;
    HIPCHK(launch_init_state(c.N, c.lr ? c.lr : sbe_ctx::default_lr(c.N), a, s));
}

extern "C" {

int sbe_initial_state(sbe_ctx *ctx, double mu, double *state)
{
    if (!ctx || !state) return SBE_ERR_ARG;
    return guarded(ctx, [&] {
        sbe_ctx &c = *ctx;
        initial_state_dev(c, mu, c.st[8], c.stream);
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
    return integrate_impl(ctx, state, output_count, t_out, j_out, log, replay, rhs_calls, nullptr);
}

int sbe_integrate_info(const sbe_ctx *ctx, int *device_controller)
{
    if (!ctx || !device_controller) return SBE_ERR_ARG;
    *device_controller = ctx->resident_used ? 1 : 0;
    return SBE_OK;
}

}  // extern "C"

namespace {

// integrate_const on the resident integrator (one cooperative launch, the
// step-size control on the device); false when the instance does not qualify
// (not 1-D, more than 8 bands, or a grid whose blocks cannot all be resident)
bool resident_integrate(sbe_ctx &c, int output_count, std::vector<double> &tv, long &steps)
{
    // (1-D, at most 8 bands, an R extent of at most 32 cells)
    if (c.dim != 1 || c.N > 8 || c.geom.W[0] > 32 || c.twL_off != 0) return false;
    if (!c.lr) {
        // the layout tuner runs on a first stage (not part of the integration)
        const long calls = c.rhs_calls;
        c.rhs_dev(c.x(), c.st[1], c.t_start(), c.stream);
        c.rhs_calls = calls;
    }
    const long kpb = stage_k_per_block(c.N, c.lr), blocks = (c.nk + kpb - 1) / kpb;
    if (blocks > resident_max_blocks(c.N, c.lr)) return false;
    // the observer operators do not depend on E: formed once
    HIPCHK(launch_phase_partial(c.geom, c.pbufs(), c.obs_ops, 0, 0, 0, c.Yobs, nullptr, c.stream));
    struct Buf {
        void *p = nullptr;
        ~Buf()
        {
            if (p) (void)hipFree(p);
        }
    } pulses, tdev, status, boxes;
    HIPCHK(hipMalloc(&pulses.p, sizeof(SubPulse) * std::max<size_t>(1, c.pulses.size())));
    HIPCHK(hipMemcpyAsync(pulses.p, c.pulses.data(), sizeof(SubPulse) * c.pulses.size(), hipMemcpyHostToDevice,
                          c.stream));
    HIPCHK(hipMalloc(&tdev.p, sizeof(double) * output_count));
    HIPCHK(hipMalloc(&status.p, sizeof(long) * 8));
    HIPCHK(hipMalloc(&boxes.p, sizeof(double) * blocks * SET_MAX * c.Y_stride));
    ResidentArgs r{};
    r.g = c.geom;
    r.Hb = c.Hb;
    r.Db = c.Db;
    r.nbox = (long)c.geom.W[0] * c.geom.W[1] * c.geom.W[2];
    r.n = c.n;
    r.nn = c.nn;
    r.P = c.P;
    r.dim = c.dim;
    r.nk = c.nk;
    r.x = c.st[0];
    r.xn = c.st[9];
    for (int i = 1; i <= 7; i++) r.k[i] = c.st[i];
    r.tw6 = c.tw6;
    r.Yb = static_cast<double *>(boxes.p);
    r.tw_stride = c.tw_stride;
    r.Y_stride = c.Y_stride;
    r.kphase = c.kphase;
    r.WL = c.WL;
    r.twL_off = c.twL_off;
    r.ucache = c.ucache;
    r.err_partial = c.err_partial;
    r.Yobs = c.Yobs;
    r.jpartial = c.partial;
    r.j_out = c.jbuf;
    r.t_out = static_cast<double *>(tdev.p);
    r.pulses = static_cast<const SubPulse *>(pulses.p);
    r.npulse = (int)c.pulses.size();
    r.F2 = c.prm.relaxation_F2;
    r.w = c.prm.soothing_width;
    r.eps_abs = c.prm.eps_abs;
    r.eps_rel = c.prm.eps_rel;
    r.sweeps = c.sweeps_for(1);
    r.cold_sweeps = c.sweeps_for(0);
    r.cold_always = (c.prm.flags & SBE_FLAG_COLD_JACOBI) ? 1 : 0;
    r.t0 = c.t_start();
    r.t1 = c.t_end();
    r.output_count = output_count;
    r.max_attempts = 100000000L;
    r.status = static_cast<long *>(status.p);
    const hipError_t e = launch_resident(c.N, c.lr, r, c.stream);
    if (e == hipErrorCooperativeLaunchTooLarge) {
        (void)hipGetLastError();
        return false;
    }
    HIPCHK(e);
    long st[8] = {};
    tv.assign(output_count, 0.0);
    HIPCHK(hipMemcpyAsync(st, status.p, sizeof st, hipMemcpyDeviceToHost, c.stream));
    HIPCHK(hipMemcpyAsync(tv.data(), tdev.p, sizeof(double) * output_count, hipMemcpyDeviceToHost, c.stream));
    HIPCHK(hipStreamSynchronize(c.stream));
    if (st[4]) {
        std::swap(c.st[0], c.st[9]);
        std::swap(c.st[1], c.st[7]);
    }
    c.rhs_calls += st[1];
    if (c.prm.relaxation_F2 > 0) c.ucache_valid = true;
    if (st[3] == 1) fail(SBE_ERR_STEP, "error norm is NaN (resident integrator, attempt %ld)", st[2]);
    if (st[3] == 2) fail(SBE_ERR_STEP, "Max number of iterations exceeded (500). A new step size was not found.");
    if (st[3] == 3) fail(SBE_ERR_STEP, "resident integrator: attempt limit %ld reached", r.max_attempts);
    steps = st[0];
    c.resident_used = true;
    return true;
}

}  // namespace

long sbe_host::integrate_impl(sbe_ctx *ctx, double *state, int output_count, double *t_out, double *j_out,
                              sbe_steplog *log, const sbe_steplog *replay, long *rhs_calls, sbe_observer *obs)
{
    if (!ctx || !state) return SBE_ERR_ARG;
    long steps = 0;
    int rc = guarded(ctx, [&] {
        sbe_ctx &c = *ctx;
        if (output_count < 2) fail(SBE_ERR_ARG, "output_count must be >= 2");
        if (obs && obs->count + output_count > obs->capacity)
            fail(SBE_ERR_ARG, "observer capacity %d cannot hold %d more records", obs->capacity, output_count);
        if (c.jbuf_cap < output_count) {
            if (c.jbuf) HIPCHK(hipFree(c.jbuf));
            c.jbuf = dalloc<double>(3 * (size_t)output_count);
            c.jbuf_cap = output_count;
        }
        c.rhs_calls = 0;
        // every integration starts its Jacobi chain from the identity, so that
        // its results do not depend on what the context computed before
        c.ucache_valid = false;
        // the fused attempt's ticket counter starts from zero
        HIPCHK(hipMemsetAsync(c.err_ticket, 0, sizeof(unsigned), c.stream));
        if (log) log->count = 0;
        HIPCHK(hipMemcpyAsync(c.x(), state, 16 * c.S, hipMemcpyHostToDevice, c.stream));
        c.resident_used = false;
        std::vector<double> tres;
        if ((c.prm.flags & SBE_FLAG_DEVICE_CONTROLLER) && !replay && !log && !obs &&
            resident_integrate(c, output_count, tres, steps)) {
            HIPCHK(hipMemcpyAsync(state, c.x(), 16 * c.S, hipMemcpyDeviceToHost, c.stream));
            std::vector<double> jh(3 * (size_t)output_count, 0.0);
            HIPCHK(hipMemcpyAsync(jh.data(), c.jbuf, 3 * sizeof(double) * output_count, hipMemcpyDeviceToHost,
                                  c.stream));
            HIPCHK(hipStreamSynchronize(c.stream));
            if (j_out) std::memcpy(j_out, jh.data(), 3 * sizeof(double) * output_count);
            if (t_out) std::memcpy(t_out, tres.data(), sizeof(double) * output_count);
            return;
        }
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
                if (obs) obs->record_dev(c.x(), time, c.stream);
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
            if (obs) obs->record_dev(c.x(), time, c.stream);
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

extern "C" {

int sbe_hk_mesh(sbe_ctx *ctx, const double A[3], const double E[3], double *out)
{
    if (!ctx || !A || !E || !out) return SBE_ERR_ARG;
    return guarded(ctx, [&] {
        sbe_ctx &c = *ctx;
        // A enters through the state's special slots
        std::vector<double> head(16, 0.0);
        for (int d = 0; d < 3; d++) head[2 * d] = A[d];
        HIPCHK(hipMemcpyAsync(c.st[8], head.data(), 16 * sizeof(double), hipMemcpyHostToDevice, c.stream));
        StageInput in{};
        in.x = c.st[8];
        HIPCHK(launch_twiddles(c.geom, &in, 1, c.tw, c.stream));
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
        HIPCHK(hipMemsetAsync(c.err_ticket, 0, sizeof(unsigned), c.stream));
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
