// Fused SBE right-hand-side stage for CDNA4 (gfx950), FP64.
//
// Per k-point:  d rho/dt = -i [H_E, rho]  +  U^dagger (D o (U rho U^dagger)) U
// where U H_0 U^dagger = diag(eps) comes from a wavefront-level complex Jacobi
// eigensolver and D_ab = -F2 (1 - exp(-((eps_a - eps_b)/w)^2)).
// Reference semantics: code/src/PropagatorCPU.cpp:73-155,208-256 (RHS),
// code/src/util/DiagonalizatorSingle.cpp:57-122 (rotation formula, threshold).
//
// Data layout in registers ("xrow"): the k-point's N x N matrix is spread over
// LK = N*LR lanes; lane (i, h) owns row i, columns j = i ^ (h*C + c) for its
// C = N/LR registers c.  With this XOR-wrapped layout and the XOR pairing
// (p, p ^ r) of Jacobi round r, every register index in the rotation and
// product loops is a compile-time constant and every cross-lane exchange is
// a fixed lane-xor.  One sweep = rounds r = 1..N-1 (each unordered pair once).
// The reference's cyclic order and this parallel order converge to the same
// decomposition; the dephasing term depends only on the converged
// eigen-decomposition (invariant under eigenvector phases and under
// rotations within degenerate subspaces, where gamma -> 0).
#include "device_common.hpp"
#include "kernels.hpp"

namespace sbe {

template <int N>
struct Geo {
    static constexpr int LR = (N <= 8) ? 1 : N / 8;  // lanes per row
    static constexpr int C = N / LR;                 // matrix entries per lane
    static constexpr int LK = N * LR;                // lanes per k-point
    static constexpr int KPW = 64 / LK;              // k-points per wavefront
    static constexpr int NN = N * N;
};

struct Lane {
    int kk, i, h;
    long k;
    bool valid;
};

template <int N>
__device__ __forceinline__ Lane lane_id(long nk)
{
    using G = Geo<N>;
    Lane L;
    int lane = threadIdx.x;
    L.kk = lane / G::LK;
    int li = lane % G::LK;
    L.i = li / G::LR;
    L.h = li % G::LR;
    L.k = (long)blockIdx.x * G::KPW + L.kk;
    L.valid = L.k < nk;
    return L;
}

// wave-local sync: blocks are exactly one wavefront
__device__ __forceinline__ void wsync() { __syncthreads(); }

// coalesced copy of the wave's KPW matrices (n x n at src + k*stride + off,
// complex) into LDS natural layout [kk][N][N], zero padded.
template <int N>
__device__ __forceinline__ void load_mats(cd *lds, const double *src, long stride, long off, int n, long nk)
{
    using G = Geo<N>;
    for (int e = threadIdx.x; e < G::KPW * G::NN; e += 64) {
        int kk = e / G::NN, idx = e % G::NN;
        int a = idx / N, b = idx % N;
        long k = (long)blockIdx.x * G::KPW + kk;
        cd v = mk(0, 0);
        if (a < n && b < n && k < nk) {
            const double *p = src + 2 * (k * stride + off + a * n + b);
            v = mk(p[0], p[1]);
        }
        lds[e] = v;
    }
}

// product out[c] = sum_l X[i][l] Y[l][j_c] with X, Y in LDS (natural layout);
// op selects conj variants: XOP 0 X, 1 conj(X^T) (i.e. X[i][l] := conj(Xs[l][i]));
// YOP 0 Y, 1 conj(Y^T).
template <int N, int XOP, int YOP>
__device__ __forceinline__ void lds_product(const cd *X, const cd *Y, const Lane &L, cd *out)
{
    using G = Geo<N>;
#pragma unroll
    for (int c = 0; c < G::C; c++) out[c] = mk(0, 0);
#pragma unroll 4
    for (int l = 0; l < N; l++) {
        cd x = (XOP == 0) ? X[L.i * N + l] : cconj(X[l * N + L.i]);
#pragma unroll
        for (int c = 0; c < G::C; c++) {
            int j = L.i ^ (L.h * G::C + c);
            cd y = (YOP == 0) ? Y[l * N + j] : cconj(Y[j * N + l]);
            out[c] = cfma(x, y, out[c]);
        }
    }
}

template <int N>
__device__ __forceinline__ void store_xrow(cd *lds, const Lane &L, const cd *v)
{
    using G = Geo<N>;
#pragma unroll
    for (int c = 0; c < G::C; c++) lds[L.i * N + (L.i ^ (L.h * G::C + c))] = v[c];
}

__device__ __forceinline__ void cs_half(double y, double x, double &c, double &s)
{
    double t = y / (fabs(x) + sqrt(x * x + y * y));
    double cc = 1 / sqrt(1 + t * t);
    double ss = t * cc;
    if (x >= 0) {
        c = cc;
        s = ss;
    } else {
        c = fabs(ss);
        s = copysign(cc, y);
    }
}

// Wavefront Jacobi on the xrow matrix A (Hermitian) -> U (rows), eigenvalues
// in Dg[kk*N + i].  LDS scratch: Dg[KPW*N], Rp/Rq[KPW*N], Fa[KPW*N].
template <int N>
__device__ void wave_jacobi(cd *A, cd *U, const Lane &L, int sweeps, double *Dg, cd *Rp, cd *Rq, int *Fa)
{
    using G = Geo<N>;
    constexpr int C = G::C;
    const int base = L.kk * N;
#pragma unroll
    for (int c = 0; c < C; c++) U[c] = mk((L.h * C + c) == 0 ? 1.0 : 0.0, 0);
    if (L.h == 0) Dg[base + L.i] = A[0].x;
    wsync();
    for (int sw = 0; sw < sweeps; sw++) {
        double s = 0;
#pragma unroll
        for (int c = 0; c < C; c++) {
            int j = L.i ^ (L.h * C + c);
            if (L.i <= j) s += fabs(A[c].x) + fabs(A[c].y);
        }
        s = seg_sum<G::LK>(s);
        const double thr = 1e-15 * (0.5 * s);
        int anyrot = 0;
#pragma unroll
        for (int r = 1; r < N; r++) {
            const int rc = r % C, rh = r / C;
            const int mrow = r * G::LR + rh;
            // (a) rotation parameters, computed by the owner of element (p, p^r), p < p^r
            if (L.h == rh && L.i < (L.i ^ r)) {
                cd apq = A[rc];
                double app = Dg[base + L.i], aqq = Dg[base + (L.i ^ r)];
                double mod = hypot(apq.x, apq.y);
                int act = L.valid && (mod > thr);
                cd rpp = mk(0, 0), rpq = mk(0, 0);
                if (act) {
                    double c1, s1, c2, s2;
                    cs_half(-apq.x, apq.y, c1, s1);
                    cs_half(2 * mod, app - aqq, c2, s2);
                    rpp = mk(-s1 * s2 - 1, -c1 * s2);
                    rpq = mk(-c1 * c2, -s1 * c2);
                }
                Rp[base + L.i] = rpp;
                Rq[base + L.i] = rpq;
                Fa[base + L.i] = act;
            }
            wsync();
            const int prow = min(L.i, L.i ^ r);
            const cd rpp = Rp[base + prow], rpq = Rq[base + prow];
            const int arow = Fa[base + prow];
            const bool isp = L.i < (L.i ^ r);
            anyrot |= arow;
#pragma unroll
            for (int c = 0; c < C; c++) {
                const int c2 = c ^ rc;
                if (c2 < c) continue;
                // row update of the register pair (c, c2)
                cd a0 = A[c], a1 = A[c2], u0 = U[c], u1 = U[c2];
                cd b0 = shfl_xor(A[c2], mrow), b1 = shfl_xor(A[c], mrow);
                cd v0 = shfl_xor(U[c2], mrow), v1 = shfl_xor(U[c], mrow);
                if (arow) {
                    if (isp) {
                        // new_p = (1 + Rpp) x_p + Rpq x_q
                        a0 = cfma(rpq, b0, cadd(a0, cmul(rpp, a0)));
                        u0 = cfma(rpq, v0, cadd(u0, cmul(rpp, u0)));
                        if (c2 != c) {
                            a1 = cfma(rpq, b1, cadd(a1, cmul(rpp, a1)));
                            u1 = cfma(rpq, v1, cadd(u1, cmul(rpp, u1)));
                        }
                    } else {
                        // new_q = Rqp x_p + (1 + Rqq) x_q, Rqp = -conj(Rpq), Rqq = conj(Rpp)
                        const cd rqp = mk(-rpq.x, rpq.y), rqq = cconj(rpp);
                        a0 = cfma(rqp, b0, cadd(a0, cmul(rqq, a0)));
                        u0 = cfma(rqp, v0, cadd(u0, cmul(rqq, u0)));
                        if (c2 != c) {
                            a1 = cfma(rqp, b1, cadd(a1, cmul(rqq, a1)));
                            u1 = cfma(rqp, v1, cadd(u1, cmul(rqq, u1)));
                        }
                    }
                }
                U[c] = u0;
                if (c2 != c) U[c2] = u1;
                if (c2 == c) a1 = a0;
                // column update of the pair: partner column j ^ r holds register c ^ rc
                cd d0, d1;
                if (rh == 0) {
                    d0 = a1;
                    d1 = a0;
                } else {
                    d0 = shfl_xor(a1, rh);
                    d1 = shfl_xor(a0, rh);
                }
#pragma unroll
                for (int t = 0; t < 2; t++) {
                    if (t == 1 && c2 == c) break;
                    const int cc = t ? c2 : c;
                    const cd a = t ? a1 : a0, d = t ? d1 : d0;
                    const int g = L.h * C + cc;
                    const int j = L.i ^ g;
                    const int pc = min(j, j ^ r);
                    const int acol = Fa[base + pc];
                    cd res = a;
                    if (acol) {
                        const cd cpp = Rp[base + pc], cpq = Rq[base + pc];
                        if (j < (j ^ r)) {
                            // A[:,p] = (1 + Rqq) A[:,p] - Rqp A[:,q]
                            const cd cqq = cconj(cpp), cqp = mk(-cpq.x, cpq.y);
                            res = csub(cadd(a, cmul(cqq, a)), cmul(cqp, d));
                        } else {
                            // A[:,q] = -Rpq A[:,p] + (1 + Rpp) A[:,q]
                            res = csub(cadd(a, cmul(cpp, a)), cmul(cpq, d));
                        }
                        if (g == r) res = mk(0, 0);
                    }
                    A[cc] = res;
                }
            }
            if (L.h == 0) Dg[base + L.i] = A[0].x;
            wsync();
        }
        if (!__any(anyrot)) break;
    }
}

template <int N>
__global__ __launch_bounds__(64) void rhs_stage_kernel(StageArgs a)
{
    using G = Geo<N>;
    constexpr int C = G::C;
    __shared__ cd S0[G::KPW * G::NN];
    __shared__ cd S1[G::KPW * G::NN];
    __shared__ double Dg[G::KPW * N];
    __shared__ cd Rp[G::KPW * N], Rq[G::KPW * N];
    __shared__ int Fa[G::KPW * N];

    const Lane L = lane_id<N>(a.nk);
    const int n = a.n, nn = n * n;
    cd *s0 = S0 + L.kk * G::NN, *s1 = S1 + L.kk * G::NN;

    if (blockIdx.x == 0 && threadIdx.x < 8) {
        double v = threadIdx.x < a.dim ? -a.E[threadIdx.x] : 0.0;
        a.dxdt[2 * threadIdx.x] = v;
        a.dxdt[2 * threadIdx.x + 1] = 0.0;
    }

    // rho -> S0, H_E -> S1
    load_mats<N>(S0, a.x, a.P, 8, n, a.nk);
    load_mats<N>(S1, a.mesh, 2L * nn, nn, n, a.nk);
    wsync();

    // coherent part: -i (H_E rho - rho H_E)
    cd dr[C];
    {
        cd t1[C], t2[C];
        lds_product<N, 0, 0>(s1, s0, L, t1);
        lds_product<N, 0, 0>(s0, s1, L, t2);
#pragma unroll
        for (int c = 0; c < C; c++) {
            cd d = csub(t1[c], t2[c]);
            dr[c] = mk(d.y, -d.x);
        }
    }

    if (a.dephase) {
        wsync();
        load_mats<N>(S1, a.mesh, 2L * nn, 0, n, a.nk);  // H_0 -> S1
        wsync();
        cd A[C], U[C];
#pragma unroll
        for (int c = 0; c < C; c++) A[c] = s1[L.i * N + (L.i ^ (L.h * C + c))];
        wave_jacobi<N>(A, U, L, a.sweeps, Dg, Rp, Rq, Fa);
        // U -> S1 (H_0 is dead)
        store_xrow<N>(s1, L, U);
        wsync();
        cd T[C];
        lds_product<N, 0, 0>(s1, s0, L, T);  // T = U rho
        wsync();
        store_xrow<N>(s0, L, T);
        wsync();
        lds_product<N, 0, 1>(s0, s1, L, T);  // rhoH = T U^dagger
        const double di = Dg[L.kk * N + L.i];
#pragma unroll
        for (int c = 0; c < C; c++) {
            const int j = L.i ^ (L.h * C + c);
            double gam;
            if (a.w < 0) {
                gam = (L.i != j) ? 1.0 : 0.0;
            } else {
                double dE = (di - Dg[L.kk * N + j]) / a.w;
                gam = 1 - exp(-dE * dE);
            }
            T[c] = cscale(T[c], -gam * a.F2);
        }
        wsync();
        store_xrow<N>(s0, L, T);
        wsync();
        lds_product<N, 0, 0>(s0, s1, L, T);  // S = rhoH' U
        wsync();
        store_xrow<N>(s0, L, T);
        wsync();
        lds_product<N, 1, 0>(s1, s0, L, T);  // U^dagger S
#pragma unroll
        for (int c = 0; c < C; c++) dr[c] = cadd(dr[c], T[c]);
    }

    // write d rho/dt (with zero pads) through LDS for coalescing
    wsync();
    store_xrow<N>(s0, L, dr);
    wsync();
    const long k0 = (long)blockIdx.x * G::KPW;
    const int P = a.P;
    for (int e = threadIdx.x; e < G::KPW * P; e += 64) {
        int kk = e / P, p = e % P;
        long k = k0 + kk;
        if (k >= a.nk) break;
        cd v = mk(0, 0);
        if (p < nn) v = S0[kk * G::NN + (p / n) * N + (p % n)];
        double *o = a.dxdt + 2 * (8 + k * P + p);
        o[0] = v.x;
        o[1] = v.y;
    }
}

// initial Fermi-Dirac state rho = U^dagger f(eps) U, and/or eigenvalues
// (code/src/PropagatorCPU.cpp:31-70, code/src/main/runSBE.cpp:130-155)
template <int N>
__global__ __launch_bounds__(64) void init_state_kernel(InitArgs a)
{
    using G = Geo<N>;
    constexpr int C = G::C;
    __shared__ cd S0[G::KPW * G::NN];
    __shared__ cd S1[G::KPW * G::NN];
    __shared__ double Dg[G::KPW * N];
    __shared__ cd Rp[G::KPW * N], Rq[G::KPW * N];
    __shared__ int Fa[G::KPW * N];
    const Lane L = lane_id<N>(a.nk);
    const int n = a.n, nn = n * n;
    cd *s0 = S0 + L.kk * G::NN, *s1 = S1 + L.kk * G::NN;
    load_mats<N>(S1, a.mesh, (long)a.ops * nn, 0, n, a.nk);
    wsync();
    cd A[C], U[C];
#pragma unroll
    for (int c = 0; c < C; c++) A[c] = s1[L.i * N + (L.i ^ (L.h * C + c))];
    wave_jacobi<N>(A, U, L, a.sweeps, Dg, Rp, Rq, Fa);
    if (a.eig && L.valid && L.h == 0 && L.i < n) a.eig[L.k * n + L.i] = Dg[L.kk * N + L.i];
    if (a.U && L.valid && L.i < n) {
#pragma unroll
        for (int c = 0; c < C; c++) {
            int j = L.i ^ (L.h * C + c);
            if (j < n) {
                double *o = a.U + 2 * ((L.k * n + L.i) * n + j);
                o[0] = U[c].x;
                o[1] = U[c].y;
            }
        }
    }
    if (!a.state) return;
    // U -> S1, f(eps) U -> S0
    store_xrow<N>(s1, L, U);
    const double f = 1.0 / (1.0 + exp((Dg[L.kk * N + L.i] - a.mu) / a.kT));
    cd FU[C];
#pragma unroll
    for (int c = 0; c < C; c++) FU[c] = (L.i < n) ? cscale(U[c], f) : mk(0, 0);
    store_xrow<N>(s0, L, FU);
    wsync();
    // rho[a][b] = sum_i conj(U[i][a]) f_i U[i][b]
    cd R[C];
    lds_product<N, 1, 0>(s1, s0, L, R);
    wsync();
    store_xrow<N>(s0, L, R);
    wsync();
    if (blockIdx.x == 0 && threadIdx.x < 8) {
        a.state[2 * threadIdx.x] = 0;
        a.state[2 * threadIdx.x + 1] = 0;
    }
    const long k0 = (long)blockIdx.x * G::KPW;
    const int P = a.P;
    for (int e = threadIdx.x; e < G::KPW * P; e += 64) {
        int kk = e / P, p = e % P;
        long k = k0 + kk;
        if (k >= a.nk) break;
        cd v = mk(0, 0);
        if (p < nn) v = S0[kk * G::NN + (p / n) * N + (p % n)];
        double *o = a.state + 2 * (8 + k * P + p);
        o[0] = v.x;
        o[1] = v.y;
    }
}

template <int N>
static hipError_t launch_rhs_N(const StageArgs &a, hipStream_t s)
{
    using G = Geo<N>;
    long blocks = (a.nk + G::KPW - 1) / G::KPW;
    if (blocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(rhs_stage_kernel<N>, dim3((unsigned)blocks), dim3(64), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_rhs_stage(int N, const StageArgs &a, hipStream_t s)
{
    switch (N) {
    case 2: return launch_rhs_N<2>(a, s);
    case 4: return launch_rhs_N<4>(a, s);
    case 8: return launch_rhs_N<8>(a, s);
    case 16: return launch_rhs_N<16>(a, s);
    }
    return hipErrorInvalidValue;
}

template <int N>
static hipError_t launch_init_N(const InitArgs &a, hipStream_t s)
{
    using G = Geo<N>;
    long blocks = (a.nk + G::KPW - 1) / G::KPW;
    if (blocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(init_state_kernel<N>, dim3((unsigned)blocks), dim3(64), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_init_state(int N, const InitArgs &a, hipStream_t s)
{
    switch (N) {
    case 2: return launch_init_N<2>(a, s);
    case 4: return launch_init_N<4>(a, s);
    case 8: return launch_init_N<8>(a, s);
    case 16: return launch_init_N<16>(a, s);
    }
    return hipErrorInvalidValue;
}

}  // namespace sbe
