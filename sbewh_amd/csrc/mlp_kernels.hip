// Surrogate MLP kernels for CDNA4 (gfx950): BF16 x BF16 -> FP32 GEMMs on the
// matrix cores (v_mfma_f32_32x32x16_bf16) with fused bias / SiLU / SiLU'
// epilogues, batch gather, MSE loss gradient, bias column sums and Adam.
//
// GEMM tiling: a 256-thread block owns a 64 x 64 tile of C, each of its four
// wavefronts a 32 x 32 sub-tile accumulated in 16 FP32 registers per lane.
// Per MFMA step a lane supplies 8 consecutive k of row (l mod 32) of opA and
// the same 8 k of column (l mod 32) of opB, k = k0 + 8 (l / 32) + i; the
// product sums over all 16 k of the step whatever order the unit pairs them
// in, as long as A and B slots agree.  The 32 x 32 result sits in lane l,
// register r at column l mod 32, row 8 (r / 4) + 4 (l / 32) + r mod 4.
// Operands are read straight from global memory (the matrices are at most
// 4096 x 256; rows are reused from L1/L2), out-of-range k, m, n read as 0.
#include "device_common.hpp"
#include "mlp.hpp"

namespace sbe {

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v16f __attribute__((ext_vector_type(16)));

__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ __forceinline__ float dsilu(float z)
{
    const float s = 1.0f / (1.0f + __expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

template <int TA, int TB>
__global__ __launch_bounds__(256) void gemm_kernel(GemmArgs g)
{
    const int e = blockIdx.z;
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int m0 = blockIdx.y * 64 + (w >> 1) * 32, n0 = blockIdx.x * 64 + (w & 1) * 32;
    const bf16 *A = g.A + e * g.strideA, *B = g.B + e * g.strideB;
    const int row = m0 + (l & 31), col = n0 + (l & 31), kg = 8 * (l >> 5);
    v16f acc;
#pragma unroll
    for (int r = 0; r < 16; r++) acc[r] = 0.0f;
    const bf16 zero = (bf16)0.0f;
    for (int k0 = 0; k0 < g.K; k0 += 16) {
        v8bf a, b;
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const int k = k0 + kg + i;
            const bool kin = k < g.K;
            a[i] = (kin && row < g.M) ? (TA ? A[(long)k * g.lda + row] : A[(long)row * g.lda + k]) : zero;
            b[i] = (kin && col < g.N) ? (TB ? B[(long)col * g.ldb + k] : B[(long)k * g.ldb + col]) : zero;
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
    const int n = col;
    if (n >= g.N) return;
    const float bias = (g.mode == 1 || g.mode == 2) ? g.bias[e * g.strideBias + n] : 0.0f;
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int m = m0 + 8 * (r >> 2) + 4 * (l >> 5) + (r & 3);
        if (m >= g.M) continue;
        const float v = acc[r];
        if (g.mode == 0) {
            g.C[e * g.strideC + (long)m * g.ldc + n] = v;
        } else if (g.mode == 1) {
            const float z = v + bias;
            g.C[e * g.strideC + (long)m * g.ldc + n] = z;
            g.H[e * g.strideH + (long)m * g.ldh + n] = (bf16)silu(z);
        } else if (g.mode == 2) {
            g.C[e * g.strideC + (long)m * g.ldc + n] = v + bias;
        } else {
            const float z = g.Zin[e * g.strideZ + (long)m * g.ldz + n];
            g.H[e * g.strideH + (long)m * g.ldh + n] = (bf16)(v * dsilu(z));
        }
    }
}

hipError_t launch_gemm(const GemmArgs &g, int ta, int tb, int members, hipStream_t s)
{
    dim3 grid((unsigned)((g.N + 63) / 64), (unsigned)((g.M + 63) / 64), (unsigned)members), blk(256);
    switch (ta * 2 + tb) {
    case 0: hipLaunchKernelGGL((gemm_kernel<0, 0>), grid, blk, 0, s, g); break;
    case 1: hipLaunchKernelGGL((gemm_kernel<0, 1>), grid, blk, 0, s, g); break;
    case 2: hipLaunchKernelGGL((gemm_kernel<1, 0>), grid, blk, 0, s, g); break;
    default: hipLaunchKernelGGL((gemm_kernel<1, 1>), grid, blk, 0, s, g); break;
    }
    return hipGetLastError();
}

// grid (batch, members), block 64
__global__ void gather_kernel(const float *Xall, const float *Tall, int din, int dout, const int *perm, long n_perm,
                              const int *counter, int spe, int batch, int cap, bf16 *X, float *T)
{
    const int b = blockIdx.x, e = blockIdx.y;
    const long idx = perm[e * n_perm + (long)(*counter % spe) * batch + b];
    const long row = (long)e * cap + b;
    for (int i = threadIdx.x; i < din; i += 64) X[row * din + i] = (bf16)Xall[idx * din + i];
    for (int i = threadIdx.x; i < dout; i += 64) T[row * dout + i] = Tall[idx * dout + i];
}

hipError_t launch_gather(const float *Xall, const float *Tall, int din, int dout, const int *perm, long n_perm,
                         const int *counter, int spe, int batch, int cap, int members, bf16 *X, float *T,
                         hipStream_t s)
{
    hipLaunchKernelGGL(gather_kernel, dim3((unsigned)batch, (unsigned)members), dim3(64), 0, s, Xall, Tall, din, dout,
                       perm, n_perm, counter, spe, batch, cap, X, T);
    return hipGetLastError();
}

__device__ __forceinline__ float block_sum256(float v)
{
    __shared__ float red[256];
    red[threadIdx.x] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    return red[0];
}

// one block per member, fixed summation order
__global__ __launch_bounds__(256) void loss_grad_kernel(const float *Y, const float *T, int batch, int cap, int dout,
                                                        bf16 *dY, float *loss, int loss_stride, const int *counter)
{
    const int e = blockIdx.x;
    const long cnt = (long)batch * dout, off = (long)e * cap * dout;
    const float sc = 2.0f / (float)cnt;
    float acc = 0.0f;
    for (long i = threadIdx.x; i < cnt; i += 256) {
        const float d = Y[off + i] - T[off + i];
        dY[off + i] = (bf16)(sc * d);
        acc += d * d;
    }
    acc = block_sum256(acc);
    if (loss && threadIdx.x == 0) loss[(long)e * loss_stride + *counter] = acc / (float)cnt;
}

hipError_t launch_loss_grad(const float *Y, const float *T, int batch, int cap, int dout, int members, bf16 *dY,
                            float *loss, int loss_stride, const int *counter, hipStream_t s)
{
    hipLaunchKernelGGL(loss_grad_kernel, dim3((unsigned)members), dim3(256), 0, s, Y, T, batch, cap, dout, dY, loss,
                       loss_stride, counter);
    return hipGetLastError();
}

// grid (ceil(n / 256), members): thread per column, sum over the batch in order
__global__ __launch_bounds__(256) void colsum_kernel(const bf16 *dZ, int batch, int n, long strideZ, float *db,
                                                     long strideDb)
{
    const int c = blockIdx.x * 256 + threadIdx.x, e = blockIdx.y;
    if (c >= n) return;
    const bf16 *z = dZ + e * strideZ;
    float acc = 0.0f;
    for (int b = 0; b < batch; b++) acc += (float)z[(long)b * n + c];
    db[e * strideDb + c] = acc;
}

hipError_t launch_colsum(const bf16 *dZ, int batch, int n, long strideZ, int members, float *db, long strideDb,
                         hipStream_t s)
{
    hipLaunchKernelGGL(colsum_kernel, dim3((unsigned)((n + 255) / 256), (unsigned)members), dim3(256), 0, s, dZ,
                       batch, n, strideZ, db, strideDb);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void adam_kernel(float *p, bf16 *pb, float *m, float *v, const float *g, long count,
                                                   float lr, float b1, float b2, float eps, long total,
                                                   const int *counter)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i >= count) return;
    const int t = *counter + 1;
    const float frac = fminf((float)(t - 1) / (float)total, 1.0f);
    const float lr_t = lr * 0.5f * (1.0f + cospif(frac));
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float mh = mi / (1.0f - powf(b1, (float)t)), vh = vi / (1.0f - powf(b2, (float)t));
    const float pi = p[i] - lr_t * mh / (sqrtf(vh) + eps);
    p[i] = pi;
    pb[i] = (bf16)pi;
}

hipError_t launch_adam(float *p, bf16 *pb, float *m, float *v, const float *g, long count, float lr, float b1,
                       float b2, float eps, long total, const int *counter, hipStream_t s)
{
    hipLaunchKernelGGL(adam_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, p, pb, m, v, g, count, lr,
                       b1, b2, eps, total, counter);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void to_bf16_kernel(const float *p, bf16 *pb, long count)
{
    const long i = (long)blockIdx.x * 256 + threadIdx.x;
    if (i < count) pb[i] = (bf16)p[i];
}

hipError_t launch_to_bf16(const float *p, bf16 *pb, long count, hipStream_t s)
{
    hipLaunchKernelGGL(to_bf16_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, p, pb, count);
    return hipGetLastError();
}

__global__ void increment_kernel(int *counter) { *counter += 1; }

hipError_t launch_increment(int *counter, hipStream_t s)
{
    hipLaunchKernelGGL(increment_kernel, dim3(1), dim3(1), 0, s, counter);
    return hipGetLastError();
}

}  // namespace sbe
