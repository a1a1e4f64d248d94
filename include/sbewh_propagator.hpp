// Header-only C++ mirror of the reference's propagation interface over the C
// ABI of libsbewh_hip.so.
//
//   sbewh::PropagatorHIP  <->  PropagatorCPU (code/src/PropagatorCPU.h:52-55)
//       calcInitialState(double fermiLevel) -> CPUstate
//       operator()(const CPUstate &state, CPUstate &dsdt, double t)   (odeint System)
//   sbewh::ObserverHIP    <->  ObserverCPU (code/src/ObserverCPU.h:43) for j(t)
//   sbewh::integrate_const(...) <-> the odeint call of propagateCPUintertwined
//       (code/src/main/runSBE.cpp:170-183), device resident
//
// CPUstate is std::vector<std::complex<double>> in the reference's layout
// (code/src/StateContext.h:7-25), so the functor drops into the reference's
// odeint call unchanged.  A context created with SBE_FLAG_GPU_LAYOUT takes
// states in the layout of the reference's GPU path instead (8 special slots,
// then rho_k unpadded, interleaved doubles: GPUstate of
// code/src/CuPropagator.h:34); the functor accepts any contiguous container
// of complex<double> or double of the context's state size.  Errors throw
// sbewh::Error (the ABI itself never throws).
#pragma once

#include <array>
#include <complex>
#include <stdexcept>
#include <string>
#include <vector>

#include "sbewh_hip.h"

namespace sbewh {

using CPUstate = std::vector<std::complex<double>>;
// host mirror of the reference's GPUstate (interleaved re/im doubles)
using InterleavedState = std::vector<double>;

namespace detail {
inline const double *as_doubles(const CPUstate &s) { return reinterpret_cast<const double *>(s.data()); }
inline double *as_doubles(CPUstate &s) { return reinterpret_cast<double *>(s.data()); }
inline const double *as_doubles(const InterleavedState &s) { return s.data(); }
inline double *as_doubles(InterleavedState &s) { return s.data(); }
inline size_t complex_count(const CPUstate &s) { return s.size(); }
inline size_t complex_count(const InterleavedState &s) { return s.size() / 2; }
inline void resize_complex(CPUstate &s, size_t n) { s.resize(n); }
inline void resize_complex(InterleavedState &s, size_t n) { s.resize(2 * n); }
}  // namespace detail

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

inline void check(long rc, const sbe_ctx *ctx)
{
    if (rc < 0) throw Error((int)rc, std::string("sbewh: ") + sbe_last_error(ctx));
}

// Owns one device context (model, grid, parameters, pulse).
class Context {
    sbe_ctx *ctx_ = nullptr;

public:
    Context(const sbe_model &m, const sbe_grid &g, const sbe_params &p, const std::vector<sbe_subpulse> &pulses)
    {
        int rc = sbe_create(&m, &g, &p, pulses.data(), (int)pulses.size(), &ctx_);
        if (rc < 0) throw Error(rc, std::string("sbe_create: ") + sbe_last_error(nullptr));
    }
    ~Context() { sbe_destroy(ctx_); }
    Context(const Context &) = delete;
    Context &operator=(const Context &) = delete;
    sbe_ctx *get() const { return ctx_; }
    long stateSize() const { return sbe_state_size(ctx_); }
    int matrixStride() const { return sbe_matrix_stride(ctx_); }
    // grid indices (i_0, i_1, i_2) of every density matrix in state order
    std::vector<std::array<int, 3>> kIndexMap() const
    {
        std::vector<std::array<int, 3>> g((size_t)sbe_num_k(ctx_));
        check(sbe_k_index_map(ctx_, g.empty() ? nullptr : g[0].data()), ctx_);
        return g;
    }
    double fermiLevel(int occupiedBands, bool belowGap = false) const
    {
        double mu = 0;
        check(sbe_fermi_level(ctx_, occupiedBands, belowGap ? 1 : 0, &mu), ctx_);
        return mu;
    }
    std::array<double, 2> timeWindow() const
    {
        std::array<double, 2> w{};
        check(sbe_time_window(ctx_, &w[0], &w[1]), ctx_);
        return w;
    }
};

// The odeint System functor (drop-in for PropagatorCPU).
class PropagatorHIP {
    const Context &c_;

public:
    explicit PropagatorHIP(const Context &c) : c_(c) {}
    void reset() {}
    template <class State = CPUstate>
    State calcInitialState(double fermiLevel) const
    {
        State s;
        detail::resize_complex(s, (size_t)c_.stateSize());
        check(sbe_initial_state(c_.get(), fermiLevel, detail::as_doubles(s)), c_.get());
        return s;
    }
    // State: CPUstate or InterleavedState, in the context's layout
    template <class State>
    void operator()(const State &state, State &dsdt, const double t) const
    {
        if ((long)detail::complex_count(state) != c_.stateSize())
            throw Error(SBE_ERR_ARG, "sbewh: state size does not match the context's layout");
        detail::resize_complex(dsdt, detail::complex_count(state));
        check(sbe_rhs(c_.get(), detail::as_doubles(state), detail::as_doubles(dsdt), t), c_.get());
    }
};

// Current observer: records (t, j) at every call.
class ObserverHIP {
    const Context &c_;

public:
    std::vector<double> t;
    std::vector<std::array<double, 3>> j;
    explicit ObserverHIP(const Context &c) : c_(c) {}
    void reset()
    {
        t.clear();
        j.clear();
    }
    template <class State>
    void operator()(const State &state, double time)
    {
        std::array<double, 3> v{};
        check(sbe_observe(c_.get(), detail::as_doubles(state), v.data()), c_.get());
        t.push_back(time);
        j.push_back(v);
    }
};

// integrate_const over the pulse window with outputCount outputs, state
// resident on the GPU for the whole run; fills the observer; returns the
// number of accepted steps.
template <class State>
long integrate_const(const Context &c, State &state, int outputCount, ObserverHIP &obs, long *rhsCalls = nullptr)
{
    std::vector<double> t(outputCount), j(3 * (size_t)outputCount);
    long steps = sbe_integrate(c.get(), detail::as_doubles(state), outputCount, t.data(), j.data(), nullptr, nullptr,
                               rhsCalls);
    check(steps, c.get());
    obs.reset();
    for (int i = 0; i < outputCount; i++) {
        obs.t.push_back(t[i]);
        obs.j.push_back({j[3 * i], j[3 * i + 1], j[3 * i + 2]});
    }
    return steps;
}

}  // namespace sbewh
