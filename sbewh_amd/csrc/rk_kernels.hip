// Dormand-Prince error norm (FP64).  The stage inputs x + (dt b_1) k_1 + ...
// are formed by the stage kernel that consumes them, and a single instance's
// error terms by its last stage kernel (per-block maxima, reduced here); the
// batched form below covers the instances of a scan.
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

// relative error of state entry i
template <class A>
__device__ __forceinline__ double err_term(const A &a, long i)
{
    double2 kv[6];
#pragma unroll
    for (int j = 0; j < 6; j++) kv[j] = reinterpret_cast<const double2 *>(a.k[j])[i];
    return err_entry(reinterpret_cast<const double2 *>(a.x)[i], kv, a.e, a.ea, a.er, a.dt);
}

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
// memory sees it without waiting for the stream to drain
__global__ __launch_bounds__(256) void error_final_kernel(const double *partial, int nb, unsigned long long *out)
{
    double v = 0;
    for (int i = threadIdx.x; i < nb; i += 256) v = fmax(v, partial[i]);
    v = block_max(v);
    if (threadIdx.x == 0)
        __hip_atomic_store(out, (unsigned long long)__double_as_longlong(v), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_SYSTEM);
}

// batched pass 1: block lb of entry j covers entries [lb, lb + 1) * ERR_CHUNK of its state
__global__ __launch_bounds__(256) void error_batched_kernel(const ErrEntry *e, const long *block0, int m,
                                                            double *partial)
{
    const long gb = blockIdx.x;
    const int j = find_entry(block0, m, gb);
    const ErrEntry q = e[j];
    const long i0 = (gb - block0[j]) * ERR_CHUNK;
    const long i1 = i0 + ERR_CHUNK < q.S ? i0 + ERR_CHUNK : q.S;
    double v = 0;
    for (long i = i0 + threadIdx.x; i < i1; i += 256) v = fmax(v, err_term(q, i));
    v = block_max(v);
    if (threadIdx.x == 0) partial[gb] = v;
}

// batched pass 2: block j reduces the partials of entry j into *err_out
__global__ __launch_bounds__(256) void error_batched_final_kernel(const ErrEntry *e, const long *block0, int m,
                                                                  long blocks, const double *partial)
{
    const int j = blockIdx.x;
    const long b1 = j + 1 < m ? block0[j + 1] : blocks;
    double v = 0;
    for (long b = block0[j] + threadIdx.x; b < b1; b += 256) v = fmax(v, partial[b]);
    v = block_max(v);
    if (threadIdx.x == 0) *e[j].err_out = v;
}

hipError_t launch_error_final(const double *partial, long nb, unsigned long long *out, hipStream_t s)
{
    if (nb < 1 || nb > 0x7fffffffL) return hipErrorInvalidValue;
    hipLaunchKernelGGL(error_final_kernel, dim3(1), dim3(256), 0, s, partial, (int)nb, out);
    return hipGetLastError();
}

hipError_t launch_error_norm_batched(const ErrEntry *d_entries, const long *d_block0, int m, long blocks,
                                     double *partial, hipStream_t s)
{
    if (m <= 0 || blocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(error_batched_kernel, dim3((unsigned)blocks), dim3(256), 0, s, d_entries, d_block0, m, partial);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(error_batched_final_kernel, dim3((unsigned)m), dim3(256), 0, s, d_entries, d_block0, m, blocks,
                       (const double *)partial);
    return hipGetLastError();
}

}  // namespace sbe
