// Dormand-Prince error norm over the whole state (FP64, HBM-bound).  The
// stage inputs x + (dt b_1) k_1 + (dt b_2) k_2 + ... are formed by the stage
// kernel that consumes them.
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

struct ErrArgs {
    const double *x;
    const double *k[6];   // k1 (= dxdt at the step start), k3, k4, k5, k6, k7
    double e[6];
    double ea, er, dt;
    long S;
    double *partial;      // [ERR_BLOCKS]
};

constexpr int ERR_BLOCKS = 1024;

// relative error of state entry i
template <class A>
__device__ __forceinline__ double err_term(const A &a, long i)
{
    double re = 0, im = 0;
#pragma unroll
    for (int j = 0; j < 6; j++) {
        double2 kv = reinterpret_cast<const double2 *>(a.k[j])[i];
        re += a.e[j] * kv.x;
        im += a.e[j] * kv.y;
    }
    double2 xv = reinterpret_cast<const double2 *>(a.x)[i];
    double2 d1 = reinterpret_cast<const double2 *>(a.k[0])[i];
    double num = sqrt(re * re + im * im);
    double den = a.ea + a.er * (sqrt(xv.x * xv.x + xv.y * xv.y) + a.dt * sqrt(d1.x * d1.x + d1.y * d1.y));
    return num / den;
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

// pass 1: grid-stride max over the state, one partial per block
__global__ __launch_bounds__(256) void error_kernel(ErrArgs a)
{
    double v = 0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < a.S; i += (long)gridDim.x * blockDim.x)
        v = fmax(v, err_term(a, i));
    v = block_max(v);
    if (threadIdx.x == 0) a.partial[blockIdx.x] = v;
}

// pass 2: one block reduces the partials (the kernel boundary orders them)
// and stores the bit pattern of the max with a system-scope store, so that a
// host polling mapped memory sees it without waiting for the stream to drain
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

hipError_t launch_error_norm(long S, const double *x, const double *const *k, const double *e, double ea,
                             double er, double dt, double *partial, unsigned long long *out, hipStream_t s)
{
    ErrArgs a{};
    a.x = x;
    for (int j = 0; j < 6; j++) {
        a.k[j] = k[j];
        a.e[j] = e[j];
    }
    a.ea = ea;
    a.er = er;
    a.dt = dt;
    a.S = S;
    a.partial = partial;
    long need = (S + 255) / 256;
    int nb = (int)(need < ERR_BLOCKS ? need : ERR_BLOCKS);
    if (nb < 1) nb = 1;
    hipLaunchKernelGGL(error_kernel, dim3(nb), dim3(256), 0, s, a);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(error_final_kernel, dim3(1), dim3(256), 0, s, partial, nb, out);
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
