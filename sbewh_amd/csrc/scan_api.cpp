// Batched scan of many pulses over one model (include/sbewh_hip.h, sbe_scan_*).
//
// The reference propagates the pulses of a multi-run strictly one after the
// other, each with its own integrate_const (code/src/main/runSBE.cpp:205-216,
// GPU: code/src/main/CuRunSBE.cu:55-71), reusing one initial state.  Here every
// instance (own k-grid, pulse and T2) keeps its own integrate_const state
// machine on the host, and all instances that are in flight advance together:
// one launch per kernel kind covers every instance (twiddles, partial phase
// sum, fused stage kernel, error norm, observer), each kernel finding its
// instance from a per-launch table of block offsets.  The host reads back one
// error norm per instance per attempt.  Per instance, the arithmetic and the
// step sequence are those of the single-instance path (same kernel bodies,
// same controller), so a scan reproduces independent runs.
#include <array>
#include <chrono>

#include "context.hpp"

using namespace sbe_host;

namespace {

struct Ctl {
    double t0 = 0, t1 = 0, dt_out = 0, t_end = 0;
    double time = 0;  // output-grid time of the current interval start
    long step = 0;    // output interval index
    int out = 0;      // outputs observed
    bool open = false, first = false, done = false;
    double start = 0, end = 0, dt = 0;
    int fails = 0;
    long steps = 0;
    int status = 0;
};

// a batched RHS request: dxdt = f(in, t) for instance c (xout: combined input)
struct RhsReq {
    sbe_ctx *c;
    StageInput in;
    double *dxdt;
    double t;
    double *xout;
    // last stage of an attempt: the error terms' coefficients of k1, k3..k7
    bool err = false;
    double err_e[6] = {};
    double err_dt = 0;
};

}  // namespace

struct sbe_scan {
    std::string err;
    std::vector<sbe_ctx *> inst;
    std::vector<Ctl> ctl;
    int n = 0, N = 0, dim = 0, ns = 0, ns_obs = 0;  // contraction stages of the RHS / observer sums
    hipStream_t stream = nullptr;
    // stage layout shared by all instances (tuned on the first batched RHS)
    int lr = 0;
    double tune_ms[8] = {};
    int tune_lr[8] = {};
    int tune_count = 0;
    // table arena: pinned host staging mirrored by device memory
    char *h_arena = nullptr, *d_arena = nullptr;
    size_t arena_cap = 0, arena_used = 0, arena_flushed = 0;
    double *d_partial = nullptr;  // error-norm maxima per stage block of the last stage
    long partial_cap = 0;
    long *stage_b0 = nullptr;     // block offsets of the last batched stage launch (device, arena)
    long stage_blocks = 0;
    double *d_err = nullptr, *h_err = nullptr;
    double *d_jout = nullptr;    // [instance][output][3]
    long jout_cap = 0;
    int output_count = 0;
    hipEvent_t ev[2] = {};

    ~sbe_scan()
    {
        for (sbe_ctx *c : inst) {
            c->stream = nullptr;  // borrowed
            delete c;
        }
        if (h_arena) (void)hipHostFree(h_arena);
        if (d_arena) (void)hipFree(d_arena);
        if (d_partial) (void)hipFree(d_partial);
        if (d_err) (void)hipFree(d_err);
        if (h_err) (void)hipHostFree(h_err);
        if (d_jout) (void)hipFree(d_jout);
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
        if (stream) (void)hipStreamDestroy(stream);
    }

    // ---- table arena: add() copies into pinned staging and returns the
    // device address; flush() uploads everything added since the last flush.
    template <class T>
    T *add(const T *src, size_t count)
    {
        size_t off = (arena_used + 255) & ~(size_t)255;
        size_t bytes = sizeof(T) * count;
        if (off + bytes > arena_cap) fail(SBE_ERR_ARG, "scan table arena exhausted (%zu bytes)", arena_cap);
        std::memcpy(h_arena + off, src, bytes);
        arena_used = off + bytes;
        return reinterpret_cast<T *>(d_arena + off);
    }
    void flush()
    {
        if (arena_used > arena_flushed)
            HIPCHK(hipMemcpyAsync(d_arena + arena_flushed, h_arena + arena_flushed, arena_used - arena_flushed,
                                  hipMemcpyHostToDevice, stream));
        arena_flushed = arena_used;
    }
    // called only when the stream is idle (after a synchronisation)
    void reset_arena() { arena_used = arena_flushed = 0; }

    // ---- batched kernels ----------------------------------------------------
    // twiddles of every request (A of its stage input) and, with fields E,
    // the operator merge of the model box
    void twiddles(const std::vector<std::pair<sbe_ctx *, StageInput>> &rq,
                  const std::vector<std::array<double, 3>> *E)
    {
        const int m = (int)rq.size();
        std::vector<TwiddleEntry> tw(m);
        long mx = 0;
        for (int j = 0; j < m; j++) {
            const sbe_ctx &c = *rq[j].first;
            tw[j].g = c.geom;
            tw[j].in = rq[j].second;
            tw[j].has_a = 1;
            tw[j].use_shift = 1;
            tw[j].tw = c.tw;
            tw[j].mm = E ? c.merge_of((*E)[j].data()) : ModelMerge{};
            long cnt = (long)c.geom.Nk[0] * c.geom.W[0] + (long)c.geom.Nk[1] * c.geom.W[1] +
                       (long)c.geom.Nk[2] * c.geom.W[2];
            if (tw[j].mm.out) cnt += tw[j].mm.nbox * 2L * tw[j].mm.nn;
            mx = std::max(mx, cnt);
        }
        TwiddleEntry *d = add(tw.data(), m);
        flush();
        HIPCHK(launch_twiddles_batched(d, m, (int)mx, stream));
    }
    // partial phase sums (ops 2 with per-request field E, or observer ops)
    void partial(const std::vector<sbe_ctx *> &cs, const std::vector<std::array<double, 3>> &E, int ops, bool obs)
    {
        const int m = (int)cs.size();
        std::vector<std::vector<ContractStage>> st(2, std::vector<ContractStage>(m));
        for (int j = 0; j < m; j++) {
            ContractStage tmp[2];
            sbe_ctx &c = *cs[j];
            int k = phase_partial_plan(c.geom, c.pbufs(), ops, E[j][0], E[j][1], E[j][2], obs ? c.Yobs : c.Y,
                                       obs ? nullptr : c.merged(), tmp);
            if (k != (obs ? ns_obs : ns)) fail(SBE_ERR_ARG, "inconsistent phase-sum plans in the scan");
            for (int d = 0; d < k; d++) st[d][j] = tmp[d];
        }
        for (int d = 0; d < (obs ? ns_obs : ns); d++) {
            std::vector<long> b0(m);
            long tot = 0;
            for (int j = 0; j < m; j++) {
                b0[j] = tot;
                tot += st[d][j].blocks_x * st[d][j].blocks_y;
            }
            ContractStage *ds = add(st[d].data(), m);
            long *db = add(b0.data(), m);
            flush();
            HIPCHK(launch_contract_batched(ds, db, m, tot, stream));
        }
    }

    // dxdt = f(rho_in, t) for every request
    void rhs(const std::vector<RhsReq> &rq)
    {
        const int m = (int)rq.size();
        if (m == 0) return;
        std::vector<std::pair<sbe_ctx *, StageInput>> tw(m);
        std::vector<sbe_ctx *> cs(m);
        std::vector<std::array<double, 3>> E(m);
        for (int j = 0; j < m; j++) {
            tw[j] = {rq[j].c, rq[j].in};
            cs[j] = rq[j].c;
            rq[j].c->E_at(rq[j].t, E[j].data());
        }
        twiddles(tw, &E);
        partial(cs, E, 2, false);
        std::vector<StageArgs> args(m);
        for (int j = 0; j < m; j++) {
            sbe_ctx &c = *rq[j].c;
            StageArgs &a = args[j];
            a.in = rq[j].in;
            a.xout = rq[j].xout;
            a.dxdt = rq[j].dxdt;
            a.ph = c.last(c.Y);
            a.nk = c.nk;
            a.n = c.n;
            a.P = c.P;
            a.dim = c.dim;
            a.sweeps = c.prm.sweep_count;
            a.F2 = c.prm.relaxation_F2;
            a.w = c.prm.soothing_width;
            for (int d = 0; d < 3; d++) a.E[d] = E[j][d];
            a.dephase = c.prm.relaxation_F2 > 0;
            a.ucache = c.ucache;
            a.warm = (c.ucache_valid && !(c.prm.flags & SBE_FLAG_COLD_JACOBI)) ? 1 : 0;
            a.sweeps = c.sweeps_for(a.warm);
        }
        if (!lr) tune_stage(args);
        const int kpb = stage_k_per_block(N, lr);
        std::vector<long> b0(m);
        long tot = 0;
        for (int j = 0; j < m; j++) {
            b0[j] = tot;
            tot += (args[j].nk + kpb - 1) / kpb;
        }
        for (int j = 0; j < m; j++)
            if (rq[j].err) {
                // the attempt's error terms, per stage block, in this (last) stage
                if (tot > partial_cap) fail(SBE_ERR_ARG, "error partial buffer too small");
                const sbe_ctx &c = *rq[j].c;
                StageArgs &a = args[j];
                a.err_partial = d_partial + b0[j];
                for (int q = 0; q < 6; q++) a.err_e[q] = rq[j].err_e[q];
                a.err_ea = c.prm.eps_abs;
                a.err_er = c.prm.eps_rel;
                a.err_dt = rq[j].err_dt;
            }
        StageArgs *da = add(args.data(), m);
        long *db = add(b0.data(), m);
        flush();
        stage_b0 = db;
        stage_blocks = tot;
        HIPCHK(launch_rhs_stage_batched(N, lr, da, db, m, tot, stream));
        for (int j = 0; j < m; j++) {
            if (args[j].dephase) rq[j].c->ucache_valid = true;
            rq[j].c->rhs_calls++;
        }
    }

    // time every layout variant on this batched launch (cold Jacobi, 3 timed
    // launches after a warm-up) and keep the fastest; SBEWH_STAGE_LR forces one
    void tune_stage(std::vector<StageArgs> args)
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
        const int m = (int)args.size();
        for (auto &a : args) a.warm = 0;
        StageArgs *da = add(args.data(), m);
        flush();
        float best = 1e30f;
        tune_count = nv;
        for (int v = 0; v < nv; v++) {
            const int kpb = stage_k_per_block(N, lrs[v]);
            std::vector<long> b0(m);
            long tot = 0;
            for (int j = 0; j < m; j++) {
                b0[j] = tot;
                tot += (args[j].nk + kpb - 1) / kpb;
            }
            long *db = add(b0.data(), m);
            flush();
            HIPCHK(launch_rhs_stage_batched(N, lrs[v], da, db, m, tot, stream));
            HIPCHK(hipEventRecord(ev[0], stream));
            for (int it = 0; it < 3; it++) HIPCHK(launch_rhs_stage_batched(N, lrs[v], da, db, m, tot, stream));
            HIPCHK(hipEventRecord(ev[1], stream));
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

    // j(t) of the listed instances into their output slots
    void observe(const std::vector<int> &ids)
    {
        const int m = (int)ids.size();
        if (m == 0) return;
        std::vector<std::pair<sbe_ctx *, StageInput>> tw(m);
        std::vector<sbe_ctx *> cs(m);
        std::vector<std::array<double, 3>> E(m, std::array<double, 3>{0, 0, 0});
        std::vector<ObsArgs> oa(m);
        std::vector<long> b0(m);
        long tot = 0;
        const int oc = output_count;
        for (int j = 0; j < m; j++) {
            sbe_ctx &c = *inst[ids[j]];
            StageInput in{};
            in.x = c.x();
            tw[j] = {&c, in};
            cs[j] = &c;
            ObsArgs &o = oa[j];
            o.x = c.x();
            o.ph = c.last(c.Yobs);
            o.nk = c.nk;
            o.n = c.n;
            o.P = c.P;
            o.dim = c.dim;
            o.partial = c.partial;
            o.j_out = d_jout + 3 * ((long)ids[j] * oc + ctl[ids[j]].out);
            o.pstride = 3;
            b0[j] = tot;
            tot += c.nk;
        }
        twiddles(tw, nullptr);
        partial(cs, E, inst[ids[0]]->obs_ops, true);
        ObsArgs *da = add(oa.data(), m);
        long *db = add(b0.data(), m);
        flush();
        HIPCHK(launch_observe_j_batched(da, db, m, tot, stream));
    }

    // one Dormand-Prince attempt of every listed instance; errors into h_err[j]
    void attempt(const std::vector<int> &ids)
    {
        const int m = (int)ids.size();
        static const double bt[6][5] = {{B21}, {B31, B32}, {B41, B42, B43}, {B51, B52, B53, B54},
                                        {B61, B62, B63, B64, B65}, {C1, C3, C4, C5, C6}};
        static const double ct[6] = {A2, A3, A4, A5, 1.0, 1.0};
        for (int s = 0; s < 6; s++) {
            std::vector<RhsReq> rq(m);
            for (int j = 0; j < m; j++) {
                sbe_ctx &c = *inst[ids[j]];
                const double dt = ctl[ids[j]].dt, t = ctl[ids[j]].start;
                RhsReq &r = rq[j];
                r.c = &c;
                r.in = StageInput{};
                r.in.x = c.st[0];
                if (s < 5) {
                    for (int q = 0; q <= s; q++) {
                        r.in.k[q] = c.st[1 + q];
                        r.in.c[q] = dt * bt[s][q];
                    }
                    r.in.nt = s + 1;
                    r.dxdt = c.st[2 + s];
                    r.xout = nullptr;
                } else {
                    const int kk[5] = {1, 3, 4, 5, 6};
                    for (int q = 0; q < 5; q++) {
                        r.in.k[q] = c.st[kk[q]];
                        r.in.c[q] = dt * bt[5][q];
                    }
                    r.in.nt = 5;
                    r.dxdt = c.st[7];
                    r.xout = c.st[9];
                    const double dc[6] = {DC1, DC3, DC4, DC5, DC6, DC7};
                    r.err = true;
                    for (int q = 0; q < 6; q++) r.err_e[q] = dt * dc[q];
                    r.err_dt = dt;
                }
                r.t = (s < 4) ? t + dt * ct[s] : t + dt;
            }
            rhs(rq);
        }
        // the last stage left each instance's error maxima per stage block
        HIPCHK(launch_error_final_batched(stage_b0, m, stage_blocks, d_partial, d_err, stream));
        HIPCHK(hipMemcpyAsync(h_err, d_err, sizeof(double) * m, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
    }
};

namespace {

template <class F>
int scan_guarded(sbe_scan *s, F &&f)
{
    try {
        f();
        return SBE_OK;
    } catch (const HipFail &h) {
        char buf[512];
        snprintf(buf, sizeof buf, "HIP error %d (%s) in %s", (int)h.e, hipGetErrorString(h.e), h.what);
        if (s) s->err = buf;
        else g_create_err = buf;
        return SBE_ERR_HIP;
    } catch (const ArgFail &a) {
        if (s) s->err = a.msg;
        else g_create_err = a.msg;
        return a.code;
    } catch (const std::exception &e) {
        if (s) s->err = e.what();
        else g_create_err = e.what();
        return SBE_ERR_ARG;
    }
}

// table bytes one batched RHS, observe or error launch needs per instance
size_t arena_bytes_per_instance()
{
    return sizeof(TwiddleEntry) + 2 * sizeof(ContractStage) + 3 * sizeof(long) + sizeof(StageArgs) +
           sizeof(ObsArgs) + 6 * 256;
}

}  // namespace

extern "C" {

int sbe_scan_create(const sbe_model *model, const sbe_grid *grid, const sbe_params *params,
                    const sbe_instance *instances, int count, const sbe_subpulse *pulses, int total_pulses,
                    sbe_scan **out)
{
    if (!out) return SBE_ERR_ARG;
    *out = nullptr;
    sbe_scan *sc = new sbe_scan();
    int rc = scan_guarded(nullptr, [&] {
        if (!model || !grid || !params || !instances || !pulses) fail(SBE_ERR_ARG, "null argument");
        if (count < 1) fail(SBE_ERR_ARG, "a scan needs at least one instance");
        HIPCHK(hipStreamCreateWithFlags(&sc->stream, hipStreamNonBlocking));
        for (auto &e : sc->ev) HIPCHK(hipEventCreate(&e));
        long partial_blocks = 0;
        for (int i = 0; i < count; i++) {
            const sbe_instance &in = instances[i];
            if (in.pulse_offset < 0 || in.num_pulses < 1 || in.pulse_offset + in.num_pulses > total_pulses)
                fail(SBE_ERR_ARG, "instance %d: sub-pulses [%d, %d) outside the pulse array (%d)", i,
                     in.pulse_offset, in.pulse_offset + in.num_pulses, total_pulses);
            sbe_grid g = *grid;
            for (int d = 0; d < 3; d++) g.Nk[d] = in.Nk[d];
            sbe_params p = *params;
            p.relaxation_F2 = in.relaxation_F2;
            sbe_ctx *c = new sbe_ctx();
            sc->inst.push_back(c);
            try {
                build(c, model, &g, &p, pulses + in.pulse_offset, in.num_pulses, true);
            } catch (ArgFail &a) {
                a.msg = "instance " + std::to_string(i) + ": " + a.msg;
                throw;
            }
            c->stream = sc->stream;
            if (c->geom.W[c->dim - 1] > 16 || (c->dim > 1 && c->geom.W[0] > 16) || (c->dim > 2 && c->geom.W[1] > 16))
                fail(SBE_ERR_ARG, "batched scans support R extents up to 16 per axis");
            partial_blocks += c->nk;  // stage blocks hold at least one k-point
        }
        sc->n = sc->inst[0]->n;
        sc->N = sc->inst[0]->N;
        sc->dim = sc->inst[0]->dim;
        sc->ns = sc->dim - 1;                        // the merged box feeds the first contraction
        sc->ns_obs = sc->dim > 1 ? sc->dim - 1 : 1;  // dim 1: one operator-box pass
        sc->arena_cap = 12 * arena_bytes_per_instance() * (size_t)count + (1 << 20);
        HIPCHK(hipHostMalloc((void **)&sc->h_arena, sc->arena_cap, hipHostMallocDefault));
        HIPCHK(hipMalloc((void **)&sc->d_arena, sc->arena_cap));
        sc->partial_cap = partial_blocks;
        sc->d_partial = dalloc<double>(partial_blocks);
        sc->d_err = dalloc<double>(count);
        HIPCHK(hipHostMalloc((void **)&sc->h_err, sizeof(double) * count, hipHostMallocDefault));
        sc->ctl.resize(count);
    });
    if (rc) {
        delete sc;
        return rc;
    }
    *out = sc;
    return SBE_OK;
}

void sbe_scan_destroy(sbe_scan *scan) { delete scan; }

const char *sbe_scan_last_error(const sbe_scan *scan) { return scan ? scan->err.c_str() : g_create_err.c_str(); }

int sbe_scan_run(sbe_scan *scan, int occupied_bands, int output_count, double *t_out, double *j_out, long *steps,
                 long *rhs_calls, double *seconds)
{
    if (!scan || !t_out || !j_out || !steps) return SBE_ERR_ARG;
    return scan_guarded(scan, [&] {
        sbe_scan &S = *scan;
        const int B = (int)S.inst.size();
        if (output_count < 2) fail(SBE_ERR_ARG, "output_count must be >= 2");
        const long need = 3L * B * output_count;
        if (S.jout_cap < need) {
            if (S.d_jout) HIPCHK(hipFree(S.d_jout));
            S.d_jout = dalloc<double>(need);
            S.jout_cap = need;
        }
        S.output_count = output_count;
        HIPCHK(hipMemsetAsync(S.d_jout, 0, sizeof(double) * need, S.stream));
        auto tic = std::chrono::steady_clock::now();
        // ground states: Fermi level and rho_0 once per distinct grid (the
        // reference reuses one initial state for all pulses of a multi-run)
        std::map<std::tuple<int, int, int>, int> first_of_grid;
        for (int i = 0; i < B; i++) {
            sbe_ctx &c = *S.inst[i];
            auto key = std::make_tuple(c.geom.Nk[0], c.geom.Nk[1], c.geom.Nk[2]);
            auto it = first_of_grid.find(key);
            if (it == first_of_grid.end()) {
                double mu = 0;
                int rc = sbe_fermi_level(&c, occupied_bands, 0, &mu);
                if (rc) fail(rc, "instance %d: %s", i, c.err.c_str());
                initial_state_dev(c, mu, c.x(), S.stream);
                first_of_grid[key] = i;
            } else {
                HIPCHK(hipMemcpyAsync(c.x(), S.inst[it->second]->x(), 16 * c.S, hipMemcpyDeviceToDevice, S.stream));
            }
            c.ucache_valid = false;
            c.rhs_calls = 0;
            Ctl &q = S.ctl[i];
            q = Ctl{};
            q.t0 = c.t_start();
            q.t1 = c.t_end();
            q.dt_out = (q.t1 - q.t0) / (output_count - 1);
            q.t_end = q.t1 + q.dt_out / 2;
            q.time = q.t0;
            q.dt = q.dt_out;
        }
        HIPCHK(hipStreamSynchronize(S.stream));
        std::vector<double> tv((size_t)B * output_count, 0.0);
        int active = B;
        // SBEWH_PROGRESS=<seconds>: a progress line on stderr at that interval
        const char *penv = std::getenv("SBEWH_PROGRESS");
        const double pint = penv ? std::atof(penv) : 0.0;
        auto plast = std::chrono::steady_clock::now();
        long iters = 0;
        while (active > 0) {
            if (pint > 0 && std::chrono::duration<double>(std::chrono::steady_clock::now() - plast).count() > pint) {
                plast = std::chrono::steady_clock::now();
                long done_steps = 0;
                for (const Ctl &q : S.ctl) done_steps += q.steps;
                fprintf(stderr, "[scan] %.1f s: %d of %d instances in flight, %ld iterations, %ld steps\n",
                        std::chrono::duration<double>(plast - tic).count(), active, B, iters, done_steps);
                fflush(stderr);
            }
            iters++;
            S.reset_arena();
            // interval boundaries: observe, open the next interval or finish
            std::vector<int> obs, firsts, att;
            for (int i = 0; i < B; i++) {
                Ctl &q = S.ctl[i];
                if (q.done || q.open) continue;
                if (less_eq_with_sign(q.time + q.dt_out, q.t_end)) {
                    if (q.out < output_count) {
                        tv[(size_t)i * output_count + q.out] = q.time;
                        obs.push_back(i);
                    }
                    q.open = true;
                    q.first = true;
                    q.start = q.time;
                    q.end = q.time + q.dt_out;
                    if (less_with_sign(q.end, q.start + q.dt)) q.dt = q.end - q.start;
                } else {
                    if (q.out < output_count) {
                        tv[(size_t)i * output_count + q.out] = q.time;
                        obs.push_back(i);
                    }
                    q.done = true;
                    active--;
                }
            }
            S.observe(obs);
            for (int i : obs) S.ctl[i].out++;
            std::vector<RhsReq> k1;
            for (int i = 0; i < B; i++) {
                Ctl &q = S.ctl[i];
                if (!q.open) continue;
                att.push_back(i);
                if (q.first) {
                    sbe_ctx &c = *S.inst[i];
                    RhsReq r{};
                    r.c = &c;
                    r.in.x = c.x();
                    r.dxdt = c.st[1];
                    r.t = q.start;
                    k1.push_back(r);
                    q.first = false;
                }
            }
            S.rhs(k1);
            if (att.empty()) {
                HIPCHK(hipStreamSynchronize(S.stream));
                continue;
            }
            S.attempt(att);
            for (size_t j = 0; j < att.size(); j++) {
                const int i = att[j];
                Ctl &q = S.ctl[i];
                sbe_ctx &c = *S.inst[i];
                const double err = S.h_err[j];
                if (!(err <= 1.0)) {
                    if (std::isnan(err) || ++q.fails >= 500) {
                        // this instance stops; the others go on
                        q.status = SBE_ERR_STEP;
                        q.open = false;
                        q.done = true;
                        active--;
                        continue;
                    }
                    double f = 0.9 * std::pow(err, -1.0 / 3.0);
                    q.dt *= f > 0.2 ? f : 0.2;
                    continue;
                }
                q.start += q.dt;
                if (err < 0.5) {
                    double e = std::max(std::pow(5.0, -5.0), err);
                    q.dt *= 0.9 * std::pow(e, -1.0 / 5.0);
                }
                q.fails = 0;
                c.accept();
                q.steps++;
                if (!less_with_sign(q.start, q.end)) {
                    q.open = false;
                    q.step++;
                    q.time = q.t0 + (double)q.step * q.dt_out;
                } else if (less_with_sign(q.end, q.start + q.dt)) {
                    q.dt = q.end - q.start;
                }
            }
        }
        HIPCHK(hipMemcpyAsync(j_out, S.d_jout, sizeof(double) * need, hipMemcpyDeviceToHost, S.stream));
        HIPCHK(hipStreamSynchronize(S.stream));
        auto toc = std::chrono::steady_clock::now();
        if (seconds) *seconds = std::chrono::duration<double>(toc - tic).count();
        std::memcpy(t_out, tv.data(), sizeof(double) * tv.size());
        for (int i = 0; i < B; i++) {
            steps[i] = S.ctl[i].status ? S.ctl[i].status : S.ctl[i].steps;
            if (rhs_calls) rhs_calls[i] = S.inst[i]->rhs_calls;
        }
    });
}

int sbe_scan_final_state(const sbe_scan *scan, int index, double *state)
{
    if (!scan || !state || index < 0 || index >= (int)scan->inst.size()) return SBE_ERR_ARG;
    sbe_scan *s = const_cast<sbe_scan *>(scan);
    return scan_guarded(s, [&] {
        sbe_ctx &c = *s->inst[index];
        HIPCHK(hipMemcpyAsync(state, c.x(), 16 * c.S, hipMemcpyDeviceToHost, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
    });
}

long sbe_scan_state_size(const sbe_scan *scan, int index)
{
    if (!scan || index < 0 || index >= (int)scan->inst.size()) return SBE_ERR_ARG;
    return scan->inst[index]->S;
}

int sbe_scan_tuning(const sbe_scan *scan, int *lanes_per_row, int *count, int *variants, double *ms)
{
    if (!scan || !lanes_per_row) return SBE_ERR_ARG;
    *lanes_per_row = scan->lr;
    if (count) *count = scan->tune_count;
    for (int v = 0; v < scan->tune_count; v++) {
        if (variants) variants[v] = scan->tune_lr[v];
        if (ms) ms[v] = scan->tune_ms[v];
    }
    return SBE_OK;
}

}  // extern "C"
