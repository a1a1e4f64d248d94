// Test driver (not part of the library): drives sbewh::PropagatorHIP, the C++
// functor with the reference's odeint System signature
//   operator()(const State &state, State &dsdt, const double t)
// (code/src/PropagatorCPU.h:55, code/src/CuPropagator.h:34), through a plain
// fixed-step Dormand-Prince 5(4) loop written here, with `m` equal steps per
// output interval of integrate_const's output grid, and sbewh::ObserverHIP at
// every output time.  The state container is CPUstate (complex, CPU layout)
// or InterleavedState (interleaved doubles) for a context created with
// SBE_FLAG_GPU_LAYOUT (the reference GPU path's unpadded layout).
#include <cstring>

#include "sbewh_propagator.hpp"

namespace {

template <class State>
void axpy_sum(State &out, const State &x, const State *k[], const double *c, int nt)
{
    for (size_t i = 0; i < x.size(); i++) {
        auto v = x[i];
        for (int j = 0; j < nt; j++) v += c[j] * (*k[j])[i];
        out[i] = v;
    }
}

template <class State>
int run(const sbewh::Context &ctx, double mu, double t0, double dt_out, int outputs, int m, double *state_out,
        double *j_out)
{
    sbewh::PropagatorHIP prop(ctx);
    sbewh::ObserverHIP obs(ctx);
    State x = prop.calcInitialState<State>(mu);
    State k[7], xt = x;
    for (auto &v : k) v = x;
    // Dormand-Prince 5(4) tableau
    const double a[6] = {1.0 / 5, 3.0 / 10, 4.0 / 5, 8.0 / 9, 1.0, 1.0};
    const double b[6][5] = {{1.0 / 5},
                            {3.0 / 40, 9.0 / 40},
                            {44.0 / 45, -56.0 / 15, 32.0 / 9},
                            {19372.0 / 6561, -25360.0 / 2187, 64448.0 / 6561, -212.0 / 729},
                            {9017.0 / 3168, -355.0 / 33, 46732.0 / 5247, 49.0 / 176, -5103.0 / 18656},
                            {35.0 / 384, 0.0, 500.0 / 1113, 125.0 / 192, -2187.0 / 6784}};
    const double c6 = 11.0 / 84;
    for (int o = 0; o + 1 < outputs; o++) {
        const double start = t0 + o * dt_out;
        obs(x, start);
        prop(x, k[0], start);  // a fresh first derivative per output interval, as integrate_const
        const double h = dt_out / m;
        for (int s = 0; s < m; s++) {
            const double t = start + s * h;
            for (int q = 0; q < 6; q++) {
                const State *ks[6];
                double cs[6];
                int nt = 0;
                for (int j = 0; j <= q && j < 5; j++) {
                    ks[nt] = &k[j];
                    cs[nt++] = h * b[q][j];
                }
                if (q == 5) {
                    ks[nt] = &k[5];
                    cs[nt++] = h * c6;
                }
                axpy_sum(xt, x, ks, cs, nt);
                prop(xt, k[q + 1], t + a[q] * h);
            }
            x = xt;          // 5th-order solution (stage 7 input)
            k[0] = k[6];     // FSAL
        }
    }
    obs(x, t0 + (outputs - 1) * dt_out);
    std::memcpy(state_out, sbewh::detail::as_doubles(x), 16 * sbewh::detail::complex_count(x));
    for (size_t i = 0; i < obs.j.size(); i++)
        for (int d = 0; d < 3; d++) j_out[3 * i + d] = obs.j[i][d];
    return 0;
}

}  // namespace

extern "C" int functor_dopri5(const sbe_model *model, const sbe_grid *grid, const sbe_params *params,
                              const sbe_subpulse *pulses, int num_pulses, double mu, double t0, double dt_out,
                              int outputs, int m, double *state_out, double *j_out)
{
    try {
        sbewh::Context ctx(*model, *grid, *params, std::vector<sbe_subpulse>(pulses, pulses + num_pulses));
        if (params->flags & SBE_FLAG_GPU_LAYOUT)
            return run<sbewh::InterleavedState>(ctx, mu, t0, dt_out, outputs, m, state_out, j_out);
        return run<sbewh::CPUstate>(ctx, mu, t0, dt_out, outputs, m, state_out, j_out);
    } catch (const sbewh::Error &e) {
        return e.code;
    }
}
