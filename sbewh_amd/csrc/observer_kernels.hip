// Observer kernels.
//
// Current expectation value at an output time (default k-space region, w_k = 1):
//   j_d = (1/N_k) sum_k Re sum_ab [ i D_d(k)_ab T_ba - dH0/dk_d(k)_ab rho_ba ],
//   T = H0(k) rho - rho H0(k),
// with H0, D_d, dH0/dk_d evaluated at k + A (code/src/ObserverCPU.cpp:188-210,
// normalisation :293-308).  The operators come from a partial phase sum whose
// last axis is contracted here.  Two passes, fixed summation order
// (deterministic).
#include "device_common.hpp"
#include "kernels.hpp"
#include "observer_device.hpp"

namespace sbe {

__global__ __launch_bounds__(64) void jk_kernel(ObsArgs a) { jk_body(a, blockIdx.x); }

// one block per k-point; entry j owns blocks [block0[j], block0[j] + nk_j)
__global__ __launch_bounds__(64) void jk_batched_kernel(const ObsArgs *args, const long *block0, int m)
{
    const long gb = blockIdx.x;
    const int j = find_entry_wave(block0, m, gb);
    const ObsArgs a = args[j];
    jk_body(a, gb - block0[j]);
}

__device__ __forceinline__ void jreduce_body(const double *partial, long nk, double *j_out, int pstride)
{
    __shared__ double s[3][256];
    double acc[3] = {0, 0, 0};
    for (long k = threadIdx.x; k < nk; k += 256)
        for (int d = 0; d < 3; d++) acc[d] += partial[(long)pstride * k + d];
    for (int d = 0; d < 3; d++) s[d][threadIdx.x] = acc[d];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            for (int d = 0; d < 3; d++) s[d][threadIdx.x] += s[d][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 3) j_out[threadIdx.x] = s[threadIdx.x][0] / (double)nk;
}

__global__ __launch_bounds__(256) void jreduce_kernel(const double *partial, long nk, double *j_out, int pstride)
{
    jreduce_body(partial, nk, j_out, pstride);
}

__global__ __launch_bounds__(256) void jreduce_batched_kernel(const ObsArgs *args)
{
    const ObsArgs &a = args[blockIdx.x];
    jreduce_body(a.partial, a.nk, a.j_out, a.pstride);
}

hipError_t launch_observe_j(const ObsArgs &a, hipStream_t s)
{
    if (a.n > OBS_MAXN) return hipErrorInvalidValue;
    hipLaunchKernelGGL(jk_kernel, dim3((unsigned)a.nk), dim3(64), 0, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(jreduce_kernel, dim3(1), dim3(256), 0, s, a.partial, a.nk, a.j_out, a.pstride);
    return hipGetLastError();
}

hipError_t launch_observe_jk(const ObsArgs &a, hipStream_t s)
{
    if (a.n > OBS_MAXN || a.pstride < 3) return hipErrorInvalidValue;
    hipLaunchKernelGGL(jk_kernel, dim3((unsigned)a.nk), dim3(64), 0, s, a);
    return hipGetLastError();
}

// ---- k-space regions ------------------------------------------------------
// Weight of k-point k in region r (ObserverCPU::calcWeight,
// code/src/ObserverCPU.cpp:137-147; the position is getFFTfracShift,
// code/src/WhFourierTransform.cpp:163-174): p = grid(k)/Nk + shift + frac(A)
// (+ frac(-A) in the moving frame) - centre, w = prod_d exp(-c_d^2 inv),
// c_d = p_d - round(p_d).
__device__ __forceinline__ double region_w(const RegionArgs &a, const RegionDev &r, long k, const double off[3],
                                           const double negA[3])
{
    const KOrder ko{{a.ko.Nf[0], a.ko.Nf[1], a.ko.Nf[2]}, {a.ko.Nrep[0], a.ko.Nrep[1], a.ko.Nrep[2]}};
    int g[3];
    k_to_grid(ko, k, g[0], g[1], g[2]);
    double w = 1;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        double pos = (double)g[d] / a.Nk[d] + off[d];
        if (r.moving) pos += negA[d];
        pos -= r.k[d];
        const double c = pos - rint(pos);
        w *= exp(-c * c * r.inv_sigma2);
    }
    return w;
}

// frac(shift + A) and frac(-A) from the state's special slots (the reference
// converts k_global_shift + A to fractional coordinates in one step,
// code/src/ObserverCPU.cpp:168-171)
__device__ __forceinline__ void region_offsets(const RegionArgs &a, double off[3], double negA[3])
{
    double A[3];
    for (int d = 0; d < 3; d++) A[d] = a.x[2 * d];
    for (int d = 0; d < 3; d++) {
        const double *ad = a.lat + 3 * d;
        off[d] = a.shift[d] + (A[0] * ad[0] + A[1] * ad[1] + A[2] * ad[2]) / (2 * M_PI);
        negA[d] = (-A[0] * ad[0] + -A[1] * ad[1] + -A[2] * ad[2]) / (2 * M_PI);
    }
}

__device__ __forceinline__ double region_src(const RegionArgs &a, long k, int q)
{
    if (q < a.Q) return a.X[k * a.Q + q];
    if (q == a.Q) return 1.0;
    return a.x[2 * (8 + k * (long)a.P) + (q - a.Q - 1)];
}

// grid (R, chunks, ceil(Qtot / 64)), 256 threads: wavefront w of the block
// takes every fourth group of 64 k-points of the chunk; lane l evaluates the
// weight of k-point l of the group, then every lane accumulates its quantity
// q over the 64 k-points with the weights broadcast lane by lane.  Block 0 of
// region 0 also keeps min / max over k of the smeared occupations.
__global__ __launch_bounds__(256) void region_sum_kernel(RegionArgs a)
{
    const int r = blockIdx.x, ch = blockIdx.y;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int Qtot = a.Q + 1 + (a.density ? 2 * a.n * a.n : 0);
    const int q = blockIdx.z * 64 + lane;
    const RegionDev rg = a.reg[r];
    double off[3], negA[3];
    region_offsets(a, off, negA);
    const long k0 = (long)ch * a.chunk, k1 = min(a.nk, k0 + a.chunk);
    const bool mm = (r == 0 && blockIdx.z == 0 && q >= 3 && q < 3 + a.n);
    double acc = 0, mn = INFINITY, mx = -INFINITY;
    for (long g = k0 + 64 * wv; g < k1; g += 256) {
        const long kl = g + lane;
        const double wl = kl < k1 ? region_w(a, rg, kl, off, negA) : 0.0;
        const int cnt = (int)min((long)64, k1 - g);
        for (int m = 0; m < cnt; m++) {
            const double w = __shfl(wl, m);
            if (q < Qtot) {
                const double v = region_src(a, g + m, q);
                acc += w * v;
                if (mm) {
                    mn = fmin(mn, v);
                    mx = fmax(mx, v);
                }
            }
        }
    }
    __shared__ double sh[3][256];
    sh[0][threadIdx.x] = acc;
    sh[1][threadIdx.x] = mn;
    sh[2][threadIdx.x] = mx;
    __syncthreads();
    if (wv == 0 && q < Qtot) {
        double t = sh[0][lane], lo = sh[1][lane], hi = sh[2][lane];
        for (int w2 = 1; w2 < 4; w2++) {
            t += sh[0][lane + 64 * w2];
            lo = fmin(lo, sh[1][lane + 64 * w2]);
            hi = fmax(hi, sh[2][lane + 64 * w2]);
        }
        a.part[((long)r * a.chunks + ch) * Qtot + q] = t;
        if (mm) {
            a.pmin[(long)ch * a.n + (q - 3)] = lo;
            a.pmax[(long)ch * a.n + (q - 3)] = hi;
        }
    }
}

// one block per region: sums the chunks in order, divides by the weight sum
// (normalizeExpectationValues, code/src/ObserverCPU.cpp:293-308) and writes
// the record; block 0 also writes A and the min / max over k
__global__ __launch_bounds__(256) void region_final_kernel(RegionArgs a)
{
    const int r = blockIdx.x;
    const int Qtot = a.Q + 1 + (a.density ? 2 * a.n * a.n : 0);
    double ws = 0;
    for (int c = 0; c < a.chunks; c++) ws += a.part[((long)r * a.chunks + c) * Qtot + a.Q];
    double *out = a.rec + 3 + (long)r * a.rstride;
    for (int q = threadIdx.x; q < Qtot; q += 256) {
        double t = 0;
        for (int c = 0; c < a.chunks; c++) t += a.part[((long)r * a.chunks + c) * Qtot + q];
        out[q] = (q == a.Q) ? t : t / ws;
    }
    if (r == 0) {
        if (threadIdx.x < 3) a.rec[threadIdx.x] = a.x[2 * threadIdx.x];
        double *mnmx = a.rec + 3 + (long)a.R * a.rstride;
        for (int b = threadIdx.x; b < a.n; b += 256) {
            double lo = INFINITY, hi = -INFINITY;
            for (int c = 0; c < a.chunks; c++) {
                lo = fmin(lo, a.pmin[(long)c * a.n + b]);
                hi = fmax(hi, a.pmax[(long)c * a.n + b]);
            }
            mnmx[b] = lo;
            mnmx[a.n + b] = hi;
        }
    }
}

hipError_t launch_region_sums(const RegionArgs &a, hipStream_t s)
{
    const int Qtot = a.Q + 1 + (a.density ? 2 * a.n * a.n : 0);
    if (a.R < 1 || a.chunks < 1 || a.nk < 1 || a.chunk * a.chunks < a.nk || a.rstride < Qtot)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(region_sum_kernel, dim3((unsigned)a.R, (unsigned)a.chunks, (unsigned)((Qtot + 63) / 64)),
                       dim3(256), 0, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(region_final_kernel, dim3((unsigned)a.R), dim3(256), 0, s, a);
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
