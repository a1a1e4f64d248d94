// The per-k current of the observer (see observer_kernels.hip), shared by
// observer_kernels.hip and the resident integrator.
#pragma once
#include "device_common.hpp"
#include "kernels.hpp"

namespace sbe {

constexpr int OBS_MAXN = 16;

__device__ __forceinline__ void jk_body(const ObsArgs &a, const long k)
{
    __shared__ cd Hs[OBS_MAXN * OBS_MAXN], Rs[OBS_MAXN * OBS_MAXN];
    const int n = a.n, nn = n * n, M = (1 + 2 * a.dim) * nn;
    const double *rho = a.x + 2 * (8 + k * a.P);
    for (int e = threadIdx.x; e < nn; e += 64) {
        Hs[e] = phase_last(a.ph, k, M, e);  // H0(k)
        Rs[e] = mk(rho[2 * e], rho[2 * e + 1]);
    }
    __syncthreads();
    double jl[3] = {0, 0, 0};
    for (int e = threadIdx.x; e < nn; e += 64) {
        int aa = e / n, bb = e % n;
        // T[b][a]
        cd t = mk(0, 0);
        for (int l = 0; l < n; l++) {
            t = cfma(Hs[bb * n + l], Rs[l * n + aa], t);
            t = csub(t, cmul(Rs[bb * n + l], Hs[l * n + aa]));
        }
        cd rba = Rs[bb * n + aa];
        for (int d = 0; d < a.dim; d++) {
            const cd Dv = phase_last(a.ph, k, M, (1 + d) * nn + e);
            const cd dHv = phase_last(a.ph, k, M, (1 + a.dim + d) * nn + e);
            cd dt = cmul(Dv, t);      // D_ab T_ba
            cd hr = cmul(dHv, rba);   // dH_ab rho_ba
            // Re(i * dt - hr) = -dt.y - hr.x
            jl[d] += -dt.y - hr.x;
        }
    }
    for (int d = 0; d < 3; d++) {
        double v = seg_sum<64>(jl[d]);
        if (threadIdx.x == 0) a.partial[(long)a.pstride * k + d] = v;
    }
}

}  // namespace sbe
