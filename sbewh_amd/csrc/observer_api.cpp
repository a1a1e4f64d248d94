// C ABI of the full observer (include/sbewh_hip.h, "full observer"): region
// samples, occupations and region-averaged density matrices recorded on the
// device at each output time, collected in the layout of the reference's
// ExpectationValues_t (code/src/ExpectationValues.h:9-28,
// code/src/ObserverBase.cpp:66-112).
#include "observer.hpp"

using namespace sbe_host;

extern "C" {

int sbe_observer_create(sbe_ctx *ctx, const sbe_region *regions, int num_regions, double smearing_width,
                        int density_stride, int capacity, sbe_observer **out)
{
    if (!ctx || !out) return SBE_ERR_ARG;
    *out = nullptr;
    sbe_observer *o = new sbe_observer();
    int rc = guarded(ctx, [&] {
        if (num_regions < 1 || !regions) fail(SBE_ERR_ARG, "at least one region is needed");
        if (capacity < 1) fail(SBE_ERR_ARG, "capacity must be >= 1");
        if (density_stride < 0) fail(SBE_ERR_ARG, "density_stride must be >= 0");
        sbe_ctx &c = *ctx;
        o->ctx = ctx;
        o->R = num_regions;
        o->Q = 3 + 2 * c.n;
        o->Qmax = o->Q + 1 + (density_stride > 0 ? 2 * c.nn : 0);
        o->stride = density_stride;
        o->capacity = capacity;
        o->smear = smearing_width;
        o->rec_size = 3 + (long)o->R * o->Qmax + 2 * c.n;
        // fractional k_global_shift from its atomic-unit form: frac_d = k . a_d / 2 pi
        for (int d = 0; d < 3; d++) {
            const double *ad = c.geom.lattice + 3 * d;
            o->kshift_frac[d] =
                (c.geom.kshift_au[0] * ad[0] + c.geom.kshift_au[1] * ad[1] + c.geom.kshift_au[2] * ad[2]) / (2 * M_PI);
        }
        std::vector<RegionDev> rd(num_regions);
        for (int r = 0; r < num_regions; r++) {
            const sbe_region &g = regions[r];
            if (!(g.fwhm > 0)) fail(SBE_ERR_ARG, "region %d: fwhm must be > 0 (INFINITY: whole zone)", r);
            for (int d = 0; d < 3; d++) rd[r].k[d] = g.k[d];
            rd[r].inv_sigma2 = std::isinf(g.fwhm) ? 0.0 : 8 * std::log(2.0) / (g.fwhm * g.fwhm);
            rd[r].moving = g.moving_frame ? 1 : 0;
        }
        o->reg = dalloc<RegionDev>(num_regions);
        HIPCHK(hipMemcpy(o->reg, rd.data(), rd.size() * sizeof(RegionDev), hipMemcpyHostToDevice));
        o->chunk = 4096;
        o->chunks = (int)((c.nk + o->chunk - 1) / o->chunk);
        o->X = dalloc<double>((size_t)c.nk * o->Q);
        o->part = dalloc<double>((size_t)o->R * o->chunks * o->Qmax);
        o->pmin = dalloc<double>((size_t)o->chunks * c.n);
        o->pmax = dalloc<double>((size_t)o->chunks * c.n);
        o->rec = dalloc<double>((size_t)capacity * o->rec_size);
    });
    if (rc) {
        delete o;
        return rc;
    }
    *out = o;
    return SBE_OK;
}

void sbe_observer_destroy(sbe_observer *obs) { delete obs; }

int sbe_observer_reset(sbe_observer *obs)
{
    if (!obs) return SBE_ERR_ARG;
    obs->count = 0;
    obs->t.clear();
    obs->dens.clear();
    return SBE_OK;
}

int sbe_observer_record(sbe_observer *obs, const double *state, double t)
{
    if (!obs || !state) return SBE_ERR_ARG;
    return guarded(obs->ctx, [&] {
        sbe_ctx &c = *obs->ctx;
        HIPCHK(hipMemcpyAsync(c.st[8], state, 16 * c.S, hipMemcpyHostToDevice, c.stream));
        obs->record_dev(c.st[8], t, c.stream);
        HIPCHK(hipStreamSynchronize(c.stream));
    });
}

int sbe_observer_record_device(sbe_observer *obs, const double *d_state, double t, void *stream)
{
    if (!obs || !d_state) return SBE_ERR_ARG;
    return guarded(obs->ctx, [&] {
        hipStream_t s = stream ? static_cast<hipStream_t>(stream) : obs->ctx->stream;
        obs->record_dev(d_state, t, s);
    });
}

int sbe_observer_count(const sbe_observer *obs, int *density_count)
{
    if (!obs) return SBE_ERR_ARG;
    if (density_count) *density_count = (int)obs->dens.size();
    return obs->count;
}

int sbe_observer_results(sbe_observer *obs, sbe_observations *out)
{
    if (!obs || !out) return SBE_ERR_ARG;
    return guarded(obs->ctx, [&] {
        sbe_ctx &c = *obs->ctx;
        const int cnt = obs->count, R = obs->R, n = c.n, Q = obs->Q, nn = c.nn;
        std::vector<double> h((size_t)cnt * obs->rec_size);
        HIPCHK(hipStreamSynchronize(c.stream));
        if (cnt) HIPCHK(hipMemcpy(h.data(), obs->rec, h.size() * sizeof(double), hipMemcpyDeviceToHost));
        int di = 0;
        for (int i = 0; i < cnt; i++) {
            const double *rec = h.data() + (size_t)i * obs->rec_size;
            if (out->t) out->t[i] = obs->t[i];
            if (out->E) {
                double E[3];
                c.E_at(obs->t[i], E);
                for (int d = 0; d < 3; d++) out->E[3 * i + d] = d < c.dim ? E[d] : 0.0;
            }
            if (out->A)
                for (int d = 0; d < 3; d++) out->A[3 * i + d] = rec[d];
            const bool has_dens = di < (int)obs->dens.size() && obs->dens[di] == i;
            for (int r = 0; r < R; r++) {
                const double *row = rec + 3 + (size_t)r * obs->Qmax;
                const size_t ir = (size_t)i * R + r;
                if (out->weight_sum) out->weight_sum[ir] = row[Q];
                if (out->j)
                    for (int d = 0; d < 3; d++) out->j[3 * ir + d] = row[d];
                if (out->occ_ham)
                    for (int b = 0; b < n; b++) out->occ_ham[ir * n + b] = row[3 + b];
                if (out->occ_wan)
                    for (int b = 0; b < n; b++) out->occ_wan[ir * n + b] = row[3 + n + b];
                if (has_dens && out->wan)
                    std::memcpy(out->wan + 2 * ((size_t)di * R + r) * nn, row + Q + 1, 2 * nn * sizeof(double));
            }
            const double *mm = rec + 3 + (size_t)R * obs->Qmax;
            if (out->occ_ham_min) std::memcpy(out->occ_ham_min + (size_t)i * n, mm, n * sizeof(double));
            if (out->occ_ham_max) std::memcpy(out->occ_ham_max + (size_t)i * n, mm + n, n * sizeof(double));
            if (has_dens) {
                if (out->density_index) out->density_index[di] = i;
                di++;
            }
        }
    });
}

long sbe_integrate_observed(sbe_ctx *ctx, double *state, int output_count, sbe_observer *obs, double *t_out,
                            double *j_out, sbe_steplog *log, const sbe_steplog *replay, long *rhs_calls)
{
    if (!ctx || !state || !obs || obs->ctx != ctx) return SBE_ERR_ARG;
    return integrate_impl(ctx, state, output_count, t_out, j_out, log, replay, rhs_calls, obs);
}

}  // extern "C"
