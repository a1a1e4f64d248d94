// Launchers of the surrogate MLP kernels (mlp_kernels.hip): BF16 operands,
// FP32 accumulation on the CDNA4 matrix cores, SiLU hidden activations.
// Every launcher covers all ensemble members at once (member = blockIdx.z,
// consecutive members `stride*` elements apart) and enqueues on `s` only.
#pragma once
#include <hip/hip_runtime.h>

namespace sbe {

typedef __bf16 bf16;

struct GemmArgs {
    int M, N, K;
    // C(m, n) = sum_k opA(m, k) opB(k, n);  opA(m, k) = ta ? A[k*lda + m] : A[m*lda + k],
    //                                         opB(k, n) = tb ? B[n*ldb + k] : B[k*ldb + n]
    const bf16 *A;
    long lda, strideA;
    const bf16 *B;
    long ldb, strideB;
    // epilogue
    int mode;              // 0: C = acc;  1: Z = acc + bias, H = silu(Z);  2: C = acc + bias;
                           // 3: H = acc * silu'(Zin)
    const float *bias;
    long strideBias;
    float *C;              // fp32 output (modes 0, 1 (Z), 2)
    long ldc, strideC;
    bf16 *H;               // bf16 output (modes 1, 3)
    long ldh, strideH;
    const float *Zin;      // mode 3
    long ldz, strideZ;
};
hipError_t launch_gemm(const GemmArgs &g, int ta, int tb, int members, hipStream_t s);

// X[e][b][:] = bf16(Xall[perm[e][(step % spe) * batch + b]]), T likewise (fp32);
// step = *counter.  perm: [members][n_perm]; member e's rows start at row e * cap.
hipError_t launch_gather(const float *Xall, const float *Tall, int din, int dout, const int *perm, long n_perm,
                         const int *counter, int spe, int batch, int cap, int members, bf16 *X, float *T,
                         hipStream_t s);
// dY = 2 (Y - T) / (batch dout) in bf16; loss[e][*counter] = mean (Y - T)^2 (if loss != null);
// member e's rows start at row e * cap
hipError_t launch_loss_grad(const float *Y, const float *T, int batch, int cap, int dout, int members, bf16 *dY,
                            float *loss, int loss_stride, const int *counter, hipStream_t s);
// db[e][n] = sum_b dZ[e][b][n]
hipError_t launch_colsum(const bf16 *dZ, int batch, int n, long strideZ, int members, float *db, long strideDb,
                         hipStream_t s);
// Adam on fp32 master parameters (all members, count = members * P) with
// lr_t = lr * 0.5 (1 + cos(pi step / total)); writes the bf16 copies; step = *counter + 1
hipError_t launch_adam(float *p, bf16 *pb, float *m, float *v, const float *g, long count, float lr, float b1,
                       float b2, float eps, long total, const int *counter, hipStream_t s);
hipError_t launch_to_bf16(const float *p, bf16 *pb, long count, hipStream_t s);
hipError_t launch_increment(int *counter, hipStream_t s);

}  // namespace sbe
