// Device functions of the phase-sum preparation (twiddles, merged model box),
// shared by phase_kernels.hip and the resident integrator.
#pragma once
#include "device_common.hpp"
#include "kernels.hpp"

namespace sbe {

__device__ __forceinline__ void twiddle_body(const PhaseGeom &g, const StageInput &in, int has_a, int use_shift,
                                             double *tw, int t, const AOverride *ao = nullptr)
{
    int n0 = g.Nk[0] * g.W[0], n1 = g.Nk[1] * g.W[1], n2 = g.Nk[2] * g.W[2];
    if (t >= n0 + n1 + n2) return;
    int d = t < n0 ? 0 : (t < n0 + n1 ? 1 : 2);
    int local = t - (d == 0 ? 0 : (d == 1 ? n0 : n0 + n1));
    int i = local / g.W[d], r = local % g.W[d];
    double k_au[3];
    for (int e = 0; e < 3; e++) {
        double A = 0.0;
        if (has_a && e < g.dim) {
            // same operation order as the stage input combination
            A = in.x[2 * e];
#pragma unroll
            for (int j = 0; j < 6; j++)
                if (j < in.nt) A += in.c[j] * ((ao && ((ao->mask >> j) & 1)) ? ao->v[j][e] : in.k[j][2 * e]);
        }
        k_au[e] = (use_shift ? g.kshift_au[e] : 0.0) + A;
    }
    // toFracK: (k . a_d) / 2 pi  (code/src/WhFourierTransform.cpp:46-52)
    const double *ad = g.lattice + 3 * d;
    double off = (k_au[0] * ad[0] + k_au[1] * ad[1] + k_au[2] * ad[2]) / (2 * M_PI);
    double ph = ((double)i / g.Nk[d] + off) * (double)(r + g.rmin[d]);
    double s, c;
    sincospi(2 * ph, &s, &c);
    reinterpret_cast<double2 *>(tw)[t] = make_double2(c, s);
}

__device__ __forceinline__ int twiddle_count(const PhaseGeom &g)
{
    return g.Nk[0] * g.W[0] + g.Nk[1] * g.W[1] + g.Nk[2] * g.W[2];
}

// entry t of the merged model box (see ModelMerge)
__device__ __forceinline__ void merge_body(const ModelMerge &mm, long t)
{
    const long M = 2L * mm.nn;
    if (!mm.out || t >= mm.nbox * M) return;
    const long b = t / M;
    const int m = (int)(t - b * M), op = m >= mm.nn, e = m - op * mm.nn;
    const double2 h = reinterpret_cast<const double2 *>(mm.Hb)[b * mm.nn + e];
    double re = h.x, im = h.y;
    if (op)
        for (int d = 0; d < mm.dim; d++) {
            const double2 r = reinterpret_cast<const double2 *>(mm.Db)[((long)d * mm.nbox + b) * mm.nn + e];
            re += mm.E[d] * r.x;
            im += mm.E[d] * r.y;
        }
    reinterpret_cast<double2 *>(mm.out)[t] = make_double2(re, im);
}

__device__ __forceinline__ void prep_body(const PhaseGeom &g, const StageInput &in, int has_a, int use_shift,
                                          double *tw, const ModelMerge &mm, long t)
{
    const int ntw = twiddle_count(g);
    if (t < ntw) twiddle_body(g, in, has_a, use_shift, tw, (int)t);
    else merge_body(mm, t - ntw);
}

}  // namespace sbe
