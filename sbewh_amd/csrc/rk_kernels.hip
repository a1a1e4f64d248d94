// Dormand-Prince stage algebra over the whole state (FP64, HBM-bound).
//
// Boost.Odeint's runge_kutta_dopri5 forms each stage input as
//   x_tmp = 1.0 * x + (dt b_1) k_1 + (dt b_2) k_2 + ...      (scale_sumN, left to right)
// and the error estimate as sum_j (dt dc_j) k_j, normalised by
//   eps_abs + eps_rel (|x_i| + dt |dxdt_i|)  and reduced with the max norm
// (default_error_checker; call site code/src/main/runSBE.cpp:178-180).
// |.| is the complex modulus of each state entry, as for the reference's CPU
// state std::vector<std::complex<double>>.
#include "device_common.hpp"
#include "kernels.hpp"

namespace sbe {

constexpr int MAXT = 7;

struct CombArgs {
    const double *x;
    const double *k[MAXT];
    double c[MAXT];
    int nt;
    long S;
    double *out;
};

template <int NT>
__global__ __launch_bounds__(256) void combine_kernel(CombArgs a)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.S) return;
    const double2 *x = reinterpret_cast<const double2 *>(a.x);
    double2 v = x[i];
    double re = v.x, im = v.y;
#pragma unroll
    for (int j = 0; j < NT; j++) {
        double2 kv = reinterpret_cast<const double2 *>(a.k[j])[i];
        re += a.c[j] * kv.x;
        im += a.c[j] * kv.y;
    }
    reinterpret_cast<double2 *>(a.out)[i] = make_double2(re, im);
}

hipError_t launch_combine(long S, const double *x, int nt, const double *const *k, const double *c,
                          double *out, hipStream_t s)
{
    if (nt < 1 || nt > MAXT) return hipErrorInvalidValue;
    CombArgs a{};
    a.x = x;
    a.nt = nt;
    a.S = S;
    a.out = out;
    for (int j = 0; j < nt; j++) {
        a.k[j] = k[j];
        a.c[j] = c[j];
    }
    dim3 grid((unsigned)((S + 255) / 256)), blk(256);
    switch (nt) {
    case 1: hipLaunchKernelGGL(combine_kernel<1>, grid, blk, 0, s, a); break;
    case 2: hipLaunchKernelGGL(combine_kernel<2>, grid, blk, 0, s, a); break;
    case 3: hipLaunchKernelGGL(combine_kernel<3>, grid, blk, 0, s, a); break;
    case 4: hipLaunchKernelGGL(combine_kernel<4>, grid, blk, 0, s, a); break;
    case 5: hipLaunchKernelGGL(combine_kernel<5>, grid, blk, 0, s, a); break;
    case 6: hipLaunchKernelGGL(combine_kernel<6>, grid, blk, 0, s, a); break;
    case 7: hipLaunchKernelGGL(combine_kernel<7>, grid, blk, 0, s, a); break;
    }
    return hipGetLastError();
}

struct ErrArgs {
    const double *x;
    const double *k[6];   // k1 (= dxdt at the step start), k3, k4, k5, k6, k7
    double e[6];
    double ea, er, dt;
    long S;
    double *partial;      // [ERR_BLOCKS]
    unsigned long long *err_bits;
};

constexpr int ERR_BLOCKS = 1024;

// pass 1: grid-stride max over the state, one partial per block
__global__ __launch_bounds__(256) void error_kernel(ErrArgs a)
{
    __shared__ double red[4];
    double v = 0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < a.S; i += (long)gridDim.x * blockDim.x) {
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
        v = fmax(v, num / den);
    }
    v = seg_max<64>(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) a.partial[blockIdx.x] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

// pass 2: one block reduces the partials; writes the bit pattern of the max
__global__ __launch_bounds__(256) void error_final_kernel(const double *partial, int nb, unsigned long long *err_bits)
{
    __shared__ double red[4];
    double v = 0;
    for (int i = threadIdx.x; i < nb; i += 256) v = fmax(v, partial[i]);
    v = seg_max<64>(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
        *err_bits = (unsigned long long)__double_as_longlong(m);
    }
}

hipError_t launch_error_norm(long S, const double *x, const double *const *k, const double *e, double ea,
                             double er, double dt, double *partial, unsigned long long *err_bits, hipStream_t s)
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
    a.err_bits = err_bits;
    long need = (S + 255) / 256;
    int nb = (int)(need < ERR_BLOCKS ? need : ERR_BLOCKS);
    if (nb < 1) nb = 1;
    hipLaunchKernelGGL(error_kernel, dim3(nb), dim3(256), 0, s, a);
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
    hipLaunchKernelGGL(error_final_kernel, dim3(1), dim3(256), 0, s, partial, nb, err_bits);
    return hipGetLastError();
}

}  // namespace sbe
