// The full observer of a context (ObserverCPU, code/src/ObserverCPU.cpp:137-308,
// ObserverBase, code/src/ObserverBase.cpp:3-112): k-space region samples,
// Hamiltonian / Wannier occupations, optional region-averaged density
// matrices.  Records stay on the device until they are collected; recording
// is asynchronous on the given stream.
#pragma once
#include "context.hpp"

struct sbe_observer {
    sbe_ctx *ctx = nullptr;
    int R = 0, Q = 0, Qmax = 0, stride = 0, capacity = 0, count = 0;
    long rec_size = 0;
    double smear = -1;
    int chunks = 0;
    long chunk = 0;
    std::vector<double> t;      // output times of the records
    std::vector<int> dens;      // records that hold density matrices
    RegionDev *reg = nullptr;
    double *X = nullptr, *part = nullptr, *pmin = nullptr, *pmax = nullptr, *rec = nullptr;

    ~sbe_observer()
    {
        if (reg) (void)hipFree(reg);
        double *p[] = {X, part, pmin, pmax, rec};
        for (double *q : p)
            if (q) (void)hipFree(q);
    }

    // one output time of the device state xs: current per k, occupations per
    // k, region sums into record `count`
    void record_dev(const double *xs, double time, hipStream_t s)
    {
        sbe_ctx &c = *ctx;
        if (count >= capacity) fail(SBE_ERR_ARG, "observer capacity (%d records) exhausted", capacity);
        StageInput in{};
        in.x = xs;
        HIPCHK(launch_twiddles(c.geom, &in, 1, c.tw, s));
        HIPCHK(launch_phase_partial(c.geom, c.pbufs(), c.obs_ops, 0, 0, 0, c.Yobs, nullptr, s));
        ObsArgs o{};
        o.x = xs;
        o.ph = c.last(c.Yobs);
        o.nk = c.nk;
        o.n = c.n;
        o.P = c.P;
        o.dim = c.dim;
        o.partial = X;
        o.pstride = Q;
        HIPCHK(launch_observe_jk(o, s));
        OccArgs oa{};
        oa.x = xs;
        oa.ph = c.last(c.Yobs);
        oa.M = c.obs_ops * c.nn;
        oa.nk = c.nk;
        oa.n = c.n;
        oa.P = c.P;
        oa.sweeps = c.sweeps_for(0);
        oa.smear = smear;
        oa.X = X;
        oa.Q = Q;
        HIPCHK(launch_occupations(c.N, c.lr ? c.lr : sbe_ctx::default_lr(c.N), oa, s));
        const bool with_density = stride > 0 && count % stride == 0;
        RegionArgs ra{};
        ra.x = xs;
        ra.X = X;
        ra.reg = reg;
        ra.R = R;
        ra.Q = Q;
        ra.n = c.n;
        ra.P = c.P;
        ra.density = with_density ? 1 : 0;
        ra.nk = c.nk;
        ra.ko = c.korder;
        for (int d = 0; d < 3; d++) {
            ra.Nk[d] = c.geom.Nk[d];
            ra.shift[d] = kshift_frac[d];
        }
        std::memcpy(ra.lat, c.geom.lattice, sizeof ra.lat);
        ra.dim = c.dim;
        ra.chunks = chunks;
        ra.chunk = chunk;
        ra.part = part;
        ra.pmin = pmin;
        ra.pmax = pmax;
        ra.rec = rec + (long)count * rec_size;
        ra.rstride = Qmax;
        HIPCHK(launch_region_sums(ra, s));
        t.push_back(time);
        if (with_density) dens.push_back(count);
        count++;
    }

    double kshift_frac[3] = {0, 0, 0};
};

namespace sbe_host {
// integrate_const of sbe_integrate, with an optional full observer recording
// every output time (sbe_api.cpp)
long integrate_impl(sbe_ctx *ctx, double *state, int output_count, double *t_out, double *j_out, sbe_steplog *log,
                    const sbe_steplog *replay, long *rhs_calls, sbe_observer *obs);
}  // namespace sbe_host
