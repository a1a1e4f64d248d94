// Dormand-Prince error norm (FP64).  The stage inputs x + (dt b_1) k_1 + ...
// are formed by the stage kernel that consumes them, and the error terms by
// the attempt's last stage kernel (per-block maxima, reduced here: one
// instance, or the instances of a batched launch; an attempt whose six stages
// share one launch reduces them in that launch's last block).
//
// Boost.Odeint's runge_kutta_dopri5 forms the error estimate as
// sum_j (dt dc_j) k_j, normalised by
//   eps_abs + eps_rel (|x_i| + dt |dxdt_i|)  and reduced with the max norm
// (default_error_checker; call site code/src/main/runSBE.cpp:178-180).
// |.| is the complex modulus of each state entry, as for the reference's CPU
// state std::vector<std::complex<double>>.
#include "device_common.hpp"
#include "kernels.hpp"

namespace sbe {

// max over the 256 threads of a block (result valid in thread 0)
__device__ __forceinline__ double block_max(double v)
{
    __shared__ double red[4];
    v = seg_max<64>(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    return fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

// Final reduction of per-block maxima (the last stage kernel of an attempt
// leaves one per block; the kernel boundary orders them): stores the bit
// pattern of the max with a system-scope store, so that a host polling mapped
// memory sees it without waiting for the stream to drain.  One block of 1024
// threads, eight independent loads in flight per thread (110 592 maxima of
// the 48^3 16-band grid: one latency chain of 14 loads per thread); the max
// is exact, so the order does not matter.
__global__ __launch_bounds__(1024) void error_final_kernel(const double *partial, int nb, unsigned long long *out)
{
    constexpr int U = 8;
    double v[U];
#pragma unroll
    for (int u = 0; u < U; u++) v[u] = 0;
    int i = threadIdx.x;
    for (; i + (U - 1) * 1024 < nb; i += U * 1024) {
        double x[U];
#pragma unroll
        for (int u = 0; u < U; u++) x[u] = partial[i + u * 1024];
#pragma unroll
        for (int u = 0; u < U; u++) v[u] = fmax(v[u], x[u]);
    }
    for (; i < nb; i += 1024) v[0] = fmax(v[0], partial[i]);
#pragma unroll
    for (int u = 1; u < U; u++) v[0] = fmax(v[0], v[u]);
    __shared__ double red[16];
    double w = seg_max<64>(v[0]);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = w;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = red[0];
        for (int k = 1; k < 16; k++) m = fmax(m, red[k]);
        __hip_atomic_store(out, (unsigned long long)__double_as_longlong(m), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// per-instance final reduction of a batched last stage's per-block maxima:
// block j reduces partial[block0[j] .. block0[j+1]) (the last entry up to
// blocks) into out[j]
__global__ __launch_bounds__(256) void error_final_batched_kernel(const long *block0, int m, long blocks,
                                                                  const double *partial, double *out)
{
    const int j = blockIdx.x;
    const long b1 = j + 1 < m ? block0[j + 1] : blocks;
    double v = 0;
    for (long b = block0[j] + threadIdx.x; b < b1; b += 256) v = fmax(v, partial[b]);
    v = block_max(v);
    if (threadIdx.x == 0) out[j] = v;
}

hipError_t launch_error_final(const double *partial, long nb, unsigned long long *out, hipStream_t s)
{
    if (nb < 1 || nb > 0x7fffffffL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(error_final_kernel, dim3(1), dim3(1024), 0, s, partial, (int)nb, out);
    return hipGetLastError();
}

hipError_t launch_error_final_batched(const long *d_block0, int m, long blocks, const double *partial, double *out,
                                     hipStream_t s)
{
    if (m <= 0 || blocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(error_final_batched_kernel, dim3((unsigned)m), dim3(256), 0, s, d_block0, m, blocks, partial, out);
    return hipGetLastError();
}

}  // namespace sbe
