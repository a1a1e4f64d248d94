// Current expectation value at an output time (default k-space region, w_k = 1):
//   j_d = (1/N_k) sum_k Re sum_ab [ i D_d(k)_ab T_ba - dH0/dk_d(k)_ab rho_ba ],
//   T = H0(k) rho - rho H0(k),
// with H0, D_d, dH0/dk_d evaluated at k + A (code/src/ObserverCPU.cpp:188-210,
// normalisation :293-308).  The operators come from a partial phase sum whose
// last axis is contracted here.  Two passes, fixed summation order
// (deterministic).
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
        if (threadIdx.x == 0) a.partial[3 * k + d] = v;
    }
}

__global__ __launch_bounds__(64) void jk_kernel(ObsArgs a) { jk_body(a, blockIdx.x); }

// one block per k-point; entry j owns blocks [block0[j], block0[j] + nk_j)
__global__ __launch_bounds__(64) void jk_batched_kernel(const ObsArgs *args, const long *block0, int m)
{
    const long gb = blockIdx.x;
    const int j = find_entry(block0, m, gb);
    const ObsArgs a = args[j];
    jk_body(a, gb - block0[j]);
}

__device__ __forceinline__ void jreduce_body(const double *partial, long nk, double *j_out)
{
    __shared__ double s[3][256];
    double acc[3] = {0, 0, 0};
    for (long k = threadIdx.x; k < nk; k += 256)
        for (int d = 0; d < 3; d++) acc[d] += partial[3 * k + d];
    for (int d = 0; d < 3; d++) s[d][threadIdx.x] = acc[d];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            for (int d = 0; d < 3; d++) s[d][threadIdx.x] += s[d][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 3) j_out[threadIdx.x] = s[threadIdx.x][0] / (double)nk;
}

__global__ __launch_bounds__(256) void jreduce_kernel(const double *partial, long nk, double *j_out)
{
    jreduce_body(partial, nk, j_out);
}

__global__ __launch_bounds__(256) void jreduce_batched_kernel(const ObsArgs *args)
{
    const ObsArgs &a = args[blockIdx.x];
    jreduce_body(a.partial, a.nk, a.j_out);
}

hipError_t launch_observe_j(const ObsArgs &a, hipStream_t s)
{
    if (a.n > OBS_MAXN) return hipErrorInvalidValue;
    hipLaunchKernelGGL(jk_kernel, dim3((unsigned)a.nk), dim3(64), 0, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(jreduce_kernel, dim3(1), dim3(256), 0, s, a.partial, a.nk, a.j_out);
    return hipGetLastError();
}

hipError_t launch_observe_j_batched(const ObsArgs *d_args, const long *d_block0, int m, long blocks, hipStream_t s)
{
    if (m <= 0 || blocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(jk_batched_kernel, dim3((unsigned)blocks), dim3(64), 0, s, d_args, d_block0, m);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(jreduce_batched_kernel, dim3((unsigned)m), dim3(256), 0, s, d_args);
    return hipGetLastError();
}

}  // namespace sbe
