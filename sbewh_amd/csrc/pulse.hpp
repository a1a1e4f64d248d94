// Pulse fields E(t) for host and device code: the sub-pulse shapes of
// code/src/pulse/{Sin2Pulse1D,GaussianPulse1D,PureGauss1D,Sin2RampPulse1D}.cpp
// and their sum over polarizations (code/src/pulse/Pulse3D.cpp:15-21).  The
// arithmetic is kept free of FMA contraction, so host and device differ only
// where their cos / sin / exp do.
#pragma once
#include <hip/hip_runtime.h>

#include <math.h>

#include "../../include/sbewh_types.h"

namespace sbe {

struct SubPulse {
    sbe_subpulse p;
    double pol[3];  // normalized polarization
};

__host__ __device__ __forceinline__ double sin2_env_field(double E0, double w, double cep, double beta, double t)
{
#pragma clang fp contract(off)
    double cp = w * t + cep, ec = cos(beta * t);
    return E0 * (cos(cp) * ec * ec - 2 * beta / w * sin(cp) * ec * sin(beta * t));
}

__host__ __device__ __forceinline__ double sub_E(const sbe_subpulse &p, double t)
{
#pragma clang fp contract(off)
    const double w = p.omega, F = p.fwhm, E0 = p.E0;
    switch (p.type) {
    case SBE_PULSE_SIN2:
        if (t < -F || t > F) return 0;
        return sin2_env_field(E0, w, p.cep, M_PI / 2 / F, t);
    case SBE_PULSE_GAUSS: {
        double beta = 4 * log(2.0) / (F * F), cp = w * t + p.cep;
        return E0 * (-2 * beta / w * t * sin(cp) + cos(cp)) * exp(-beta * t * t);
    }
    case SBE_PULSE_PURE_GAUSS: {
        double beta = 4 * log(2.0) / (F * F);
        return E0 * exp(-beta * t * t);
    }
    case SBE_PULSE_SIN2_RAMP: {
        double rc = p.rising_cycles, beta = w / 4 / rc;
        int co = p.cycles_on;
        double rising = 2 * M_PI / w * rc, on = M_PI / w * co, half = rising + M_PI / w * co;
        if (t < -half) return 0;
        if (t < -on) return sin2_env_field(E0, w, 0.0, beta, t + on);
        if (t < on) return ((co % 2 == 0) ? 1 : -1) * E0 * cos(w * t);
        if (t < half) return sin2_env_field(E0, w, 0.0, beta, t - on);
        return 0;
    }
    }
    return 0;
}

__host__ __device__ inline double sub_start(const sbe_subpulse &p)
{
#pragma clang fp contract(off)
    if (p.type == SBE_PULSE_SIN2_RAMP) return -2 * M_PI / p.omega * (p.rising_cycles + 0.5 * p.cycles_on);
    return p.t_start_fwhm * p.fwhm;
}
__host__ __device__ inline double sub_end(const sbe_subpulse &p)
{
#pragma clang fp contract(off)
    if (p.type == SBE_PULSE_SIN2_RAMP) return 2 * M_PI / p.omega * (p.rising_cycles + 0.5 * p.cycles_on);
    return p.t_end_fwhm * p.fwhm;
}

// E(t) = sum over sub-pulses of E_p(t - centre_p) pol_p
__host__ __device__ __forceinline__ void pulse_E(const SubPulse *ps, int np, double t, double E[3])
{
#pragma clang fp contract(off)
    E[0] = E[1] = E[2] = 0;
    for (int q = 0; q < np; q++) {
        const double e = sub_E(ps[q].p, t - ps[q].p.center);
        for (int d = 0; d < 3; d++) E[d] += e * ps[q].pol[d];
    }
}

}  // namespace sbe
