// Device helpers shared by the CDNA4 kernels: FP64 complex arithmetic and
// wavefront (64-lane) exchanges.
#pragma once
#include <hip/hip_runtime.h>

namespace sbe {

struct cd {
    double x, y;
};

__device__ __forceinline__ cd mk(double x, double y) { return cd{x, y}; }
__device__ __forceinline__ cd cadd(cd a, cd b) { return cd{a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ cd csub(cd a, cd b) { return cd{a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ cd cmul(cd a, cd b) { return cd{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ __forceinline__ cd cconj(cd a) { return cd{a.x, -a.y}; }
__device__ __forceinline__ cd cscale(cd a, double s) { return cd{a.x * s, a.y * s}; }
// acc + a*b
__device__ __forceinline__ cd cfma(cd a, cd b, cd acc)
{
    return cd{fma(a.x, b.x, fma(-a.y, b.y, acc.x)), fma(a.x, b.y, fma(a.y, b.x, acc.y))};
}
// acc + conj(a)*b
__device__ __forceinline__ cd cfma_cj(cd a, cd b, cd acc)
{
    return cd{fma(a.x, b.x, fma(a.y, b.y, acc.x)), fma(a.x, b.y, fma(-a.y, b.x, acc.y))};
}
// acc + a*conj(b)
__device__ __forceinline__ cd cfma_jc(cd a, cd b, cd acc)
{
    return cd{fma(a.x, b.x, fma(a.y, b.y, acc.x)), fma(-a.x, b.y, fma(a.y, b.x, acc.y))};
}
__device__ __forceinline__ double cabs2(cd a) { return a.x * a.x + a.y * a.y; }

__device__ __forceinline__ cd shfl_xor(cd v, int mask)
{
    return cd{__shfl_xor(v.x, mask), __shfl_xor(v.y, mask)};
}

// sum over aligned segments of `seg` lanes (seg a power of two <= 64)
template <int seg>
__device__ __forceinline__ double seg_sum(double v)
{
#pragma unroll
    for (int m = 1; m < seg; m <<= 1) v += __shfl_xor(v, m);
    return v;
}

template <int seg>
__device__ __forceinline__ double seg_max(double v)
{
#pragma unroll
    for (int m = 1; m < seg; m <<= 1) v = fmax(v, __shfl_xor(v, m));
    return v;
}

// k index (state / Fourier-block order) <-> global grid indices.
// kidx = c + Nf2 (b + Nf1 (a + Nf0 blk)), blk = cr + Nrep2 (br + Nrep1 ar),
// i0 = a Nrep0 + ar, ...  (code/src/WhFourierTransform.cpp:137-149,303-312)
struct KOrder {
    int Nf[3];
    int Nrep[3];
};

__device__ __forceinline__ void k_to_grid(const KOrder &o, long k, int &i0, int &i1, int &i2)
{
    long fft = (long)o.Nf[0] * o.Nf[1] * o.Nf[2];
    long blk = k / fft, f = k - blk * fft;
    int a = (int)(f / (o.Nf[1] * o.Nf[2]));
    int b = (int)((f / o.Nf[2]) % o.Nf[1]);
    int c = (int)(f % o.Nf[2]);
    int ar = (int)(blk / (o.Nrep[1] * o.Nrep[2]));
    int br = (int)((blk / o.Nrep[2]) % o.Nrep[1]);
    int cr = (int)(blk % o.Nrep[2]);
    i0 = a * o.Nrep[0] + ar;
    i1 = b * o.Nrep[1] + br;
    i2 = c * o.Nrep[2] + cr;
}

}  // namespace sbe
