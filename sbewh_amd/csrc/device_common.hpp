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
// Relative error term of one state entry of a Dormand-Prince attempt,
//   |sum_j e[j] k_j| / (ea + er (|x| + dt |k_0|)),  k_0..k_5 = k1, k3, k4, k5, k6, k7
// (the controller's error checker, code/src/main/runSBE.cpp:178-180 via
// Boost.Odeint), with a fixed operation order: every kernel that evaluates
// it gets the same bits.
__device__ __forceinline__ double err_entry(double2 xv, const double2 (&kv)[6], const double (&e)[6], double ea,
                                            double er, double dt)
{
    double re = 0, im = 0;
#pragma unroll
    for (int j = 0; j < 6; j++) {
        re = fma(e[j], kv[j].x, re);
        im = fma(e[j], kv[j].y, im);
    }
    const double num = sqrt(fma(re, re, im * im));
    const double ax = sqrt(fma(xv.x, xv.x, xv.y * xv.y)), ak = sqrt(fma(kv[0].x, kv[0].x, kv[0].y * kv[0].y));
    return num / fma(er, fma(dt, ak, ax), ea);
}

// Last-axis contraction of a partial phase sum (see PhaseLast in kernels.hpp)
// for k-point k, entry e of an M-wide operator row.
template <class PL>
__device__ __forceinline__ cd phase_last(const PL &p, long k, int M, int e)
{
    const int2 ph = reinterpret_cast<const int2 *>(p.kphase)[k];
    const double2 *y = reinterpret_cast<const double2 *>(p.Y) + (long)ph.x * M + e;
    const double2 *t = reinterpret_cast<const double2 *>(p.twL) + ph.y;
    cd v = mk(0, 0);
#pragma unroll 4
    for (int q = 0; q < p.WL; q++) {
        const double2 tq = t[q], yq = y[(long)q * M];
        v = cfma(mk(tq.x, tq.y), mk(yq.x, yq.y), v);
    }
    return v;
}

// Entry j of a batched launch owns the logical blocks [b0[j], b0[j+1]):
// the largest j with b0[j] <= gb (b0 ascending, b0[0] == 0).
__device__ __forceinline__ int find_entry(const long *b0, int m, long gb)
{
    int lo = 0, hi = m - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (b0[mid] <= gb) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// The same for m <= 64 with every lane of the wavefront active: one load per
// lane and a ballot instead of log2(m) dependent loads (b0 ascending from 0,
// so the entries with b0[j] <= gb are a prefix).
__device__ __forceinline__ int find_entry_wave(const long *b0, int m, long gb)
{
    if (m > 64) return find_entry(b0, m, gb);
    const int l = threadIdx.x & 63;
    const bool le = l < m && b0[l] <= gb;
    return __popcll(__ballot(le)) - 1;
}

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup b to XCD
// b % 8).  This bijection gives each XCD one contiguous range of logical
// blocks, so neighbouring k-points (which share phase-sum rows) meet in the
// same XCD's L2.
__device__ __forceinline__ long xcd_block(long b, long nb)
{
    const long q = nb >> 3, r = nb & 7, x = b & 7;
    return x * q + (x < r ? x : r) + (b >> 3);
}

__device__ __forceinline__ double cabs2(cd a) { return a.x * a.x + a.y * a.y; }

__device__ __forceinline__ cd shfl_xor(cd v, int mask)
{
    return cd{__shfl_xor(v.x, mask), __shfl_xor(v.y, mask)};
}

// Lane exchange lane ^ MASK for MASK < 16 with DPP moves (no LDS round trip):
// quad_perm gives xor 1, 2, 3; row_half_mirror is xor 7 within 8 lanes and
// row_mirror xor 15 within 16 lanes; other masks are compositions.
template <int CTRL>
__device__ __forceinline__ int dpp_i(int v)
{
    return __builtin_amdgcn_mov_dpp(v, CTRL, 0xF, 0xF, true);
}

template <int MASK>
__device__ __forceinline__ int xor_dpp_i(int v)
{
    static_assert(MASK > 0 && MASK < 16, "DPP exchange covers masks 1..15");
    if constexpr (MASK == 1) return dpp_i<0xB1>(v);        // quad_perm [1,0,3,2]
    else if constexpr (MASK == 2) return dpp_i<0x4E>(v);   // quad_perm [2,3,0,1]
    else if constexpr (MASK == 3) return dpp_i<0x1B>(v);   // quad_perm [3,2,1,0]
    else if constexpr (MASK == 7) return dpp_i<0x141>(v);  // row_half_mirror
    else if constexpr (MASK == 15) return dpp_i<0x140>(v); // row_mirror
    else if constexpr (MASK < 8) return dpp_i<0x141>(xor_dpp_i<MASK ^ 7>(v));
    else return dpp_i<0x140>(xor_dpp_i<MASK ^ 15>(v));
}

template <int MASK>
__device__ __forceinline__ double xor_dpp(double v)
{
    long long b = __double_as_longlong(v);
    int lo = (int)(b & 0xffffffffll), hi = (int)(b >> 32);
    lo = xor_dpp_i<MASK>(lo);
    hi = xor_dpp_i<MASK>(hi);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// lane ^ MASK exchange (MASK < 64) through the LDS crossbar (ds_bpermute,
// no LDS storage); l4 = 4 x this lane's index in the wavefront.  One address
// for the four dwords and no range check (every lane ^ MASK is a lane of the
// wavefront).
template <int MASK>
__device__ __forceinline__ cd xchg_bp(cd v, int l4)
{
    static_assert(MASK > 0 && MASK < 64, "lane mask within one wavefront");
    const int addr = l4 ^ (MASK << 2);
    const long long bx = __double_as_longlong(v.x), by = __double_as_longlong(v.y);
    const int xl = __builtin_amdgcn_ds_bpermute(addr, (int)bx), xh = __builtin_amdgcn_ds_bpermute(addr, (int)(bx >> 32));
    const int yl = __builtin_amdgcn_ds_bpermute(addr, (int)by), yh = __builtin_amdgcn_ds_bpermute(addr, (int)(by >> 32));
    return cd{__longlong_as_double(((long long)xh << 32) | (unsigned int)xl),
              __longlong_as_double(((long long)yh << 32) | (unsigned int)yl)};
}

// lane ^ MASK exchange of a complex value: DPP for MASK < 16, ds_bpermute otherwise
template <int MASK>
__device__ __forceinline__ cd xchg(cd v, int l4)
{
    if constexpr (MASK < 16) return cd{xor_dpp<MASK>(v.x), xor_dpp<MASK>(v.y)};
    else return xchg_bp<MASK>(v, l4);
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
