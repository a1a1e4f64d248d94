// Fused SBE right-hand-side stage for CDNA4 (gfx950), FP64: device code (the
// layouts, the wavefront Jacobi, the stage bodies and their kernels), shared by
// the launchers in stage_kernels.hip and the resident integrator.
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
//
// LR (lanes per row) trades registers per lane against lanes per k-point; the
// host tunes it per problem (one wavefront holds KPW = 64/(N LR) k-points).
#pragma once
#include <utility>

#include "device_common.hpp"
#include "kernels.hpp"

namespace sbe {


template <int N, int LR>
struct Geo {
    static_assert(N * LR <= 64 && LR <= N, "layout must fit one wavefront");
    static constexpr int C = N / LR;     // matrix entries per lane
    static constexpr int LK = N * LR;    // lanes per k-point
    static constexpr int KPW = 64 / LK;  // k-points per wavefront
    static constexpr int NN = N * N;
};

struct Lane {
    int kk, i, h;
    int l4;  // 4 x the lane's index in the wavefront (a ds_bpermute address)
    long k;
    bool valid;
};

template <int N, int LR>
__device__ __forceinline__ Lane lane_id(long nk, long blk)
{
    using G = Geo<N, LR>;
    Lane L;
    int lane = threadIdx.x;
    L.kk = lane / G::LK;
    int li = lane % G::LK;
    L.i = li / LR;
    L.h = li % LR;
    L.l4 = (lane & 63) << 2;
    L.k = blk * G::KPW + L.kk;
    L.valid = L.k < nk;
    return L;
}

// LDS position of element (a, b) of an N x N matrix: row a, column b ^ g(a),
// with g(a) = a * x in GF(2^m) (N = 2^m; x^2+x+1, x^3+x+1, x^4+x+1).  g and
// a -> a ^ g(a) are bijections, so the wavefront's column walks (X[i][l] over
// rows i), row walks (Y[l][j]) and XOR-wrapped row accesses (row i, column
// i ^ c) of one k-point all fall on distinct LDS banks.
template <int N>
__device__ __forceinline__ int sw(int a, int b)
{
    if constexpr (N == 2) {
        return a * N + b;
    } else {
        // reduction by x^m = x + 1: the same low bits (3) for m = 2, 3, 4
        const int g = ((a << 1) & (N - 1)) ^ ((a & (N >> 1)) ? 3 : 0);
        return a * N + (b ^ g);
    }
}

// LDS layouts of a k-point's N x N matrix (MS entries per k-point):
// SwLay is the swizzled natural layout above; SlotLay keeps element (a, b) in
// row a, slot a ^ b, with a row stride of N + 1 entries.  In SlotLay the
// entries a lane of the one-lane-per-row layout touches in a product,
// Y[i ^ c'][i ^ c] (slot c' ^ c) and its transposed reads (row i ^ c, slot c),
// sit at compile-time offsets from one row address, and the odd row stride
// spreads the rows of a k-point over the LDS banks.
template <int N>
struct SwLay {
    static constexpr int MS = N * N;
    static constexpr int CHUNK = 64;  // entries per lane whose global loads are in flight together
    __device__ static int at(int a, int b) { return sw<N>(a, b); }
};
template <int N>
struct SlotLay {
    static constexpr int RS = N + 1;
    static constexpr int MS = N * RS;
    static constexpr int CHUNK = 2;
    __device__ static int at(int a, int b) { return a * RS + (a ^ b); }
};

// f(integral_constant<int, c * CH>) for c = 0 .. IT/CH - 1, unrolled at compile
// time (arrays indexed by the chunk offset stay in registers)
template <int CH, class F, int... I>
__device__ __forceinline__ void chunks_impl(F &&f, std::integer_sequence<int, I...>)
{
    (f(std::integral_constant<int, I * CH>{}), ...);
}
template <int IT, int CH, class F>
__device__ __forceinline__ void for_chunks(F &&f)
{
    chunks_impl<CH>(f, std::make_integer_sequence<int, IT / CH>{});
}

// wave-local sync: blocks are exactly one wavefront
__device__ __forceinline__ void wsync() { __syncthreads(); }

// coalesced copy of the wave's KPW matrices (n x n at src + k*stride + off,
// complex) into LDS natural layout [kk][N][N], zero padded.
template <int N, int LR, class Lay = SwLay<N>>
__device__ __forceinline__ void load_mats(cd *lds, const double *src, long stride, long off, int n, long nk, long blk)
{
    using G = Geo<N, LR>;
    constexpr int IT = G::KPW * G::NN / 64;  // entries per lane
    constexpr int CH = N >= 16 ? 4 : (Lay::CHUNK < IT ? Lay::CHUNK : IT);     // entries whose loads are in flight together
    const double2 *s2 = reinterpret_cast<const double2 *>(src);
    // unconditional loads (masked entries read entry `off` of k-point 0)
    for_chunks<IT, CH>([&](auto i0c) {
        constexpr int i0 = decltype(i0c)::value;
        double2 v[CH];
        bool ok[CH];
#pragma unroll
        for (int u = 0; u < CH; u++) {
            const int e = threadIdx.x + 64 * (i0 + u);
            const int kk = e / G::NN, idx = e % G::NN;
            const int a = idx / N, b = idx % N;
            const long k = blk * G::KPW + kk;
            ok[u] = a < n && b < n && k < nk;
            v[u] = s2[ok[u] ? k * stride + off + a * n + b : off];
        }
#pragma unroll
        for (int u = 0; u < CH; u++) {
            const int e = threadIdx.x + 64 * (i0 + u);
            const int idx = e % G::NN;
            lds[(e / G::NN) * Lay::MS + Lay::at(idx / N, idx % N)] = ok[u] ? mk(v[u].x, v[u].y) : mk(0, 0);
        }
    });
}

// H(k) of operator op (0: H0, 1: H_E) for the wave's k-points into LDS
// natural layout: the last phase-sum contraction
//   H(k)[a][b] = sum_r tw_L[i_L(k)][r] Y[rest(k)][r][op][a][b]
// (the k-points of a wave share rest(k), so the Y rows are L1/L2 hits).
template <int N, int LR, class Lay = SwLay<N>>
__device__ __forceinline__ void phase_mats(cd *lds, const StageArgs &a, int op, long blk)
{
    using G = Geo<N, LR>;
    constexpr int IT = G::KPW * G::NN / 64;  // entries per lane
    constexpr int WQ = 16;                   // largest R extent of the last axis
    const int n = a.n, nn = n * n, M = 2 * nn, WL = a.ph.WL;
    if (WL > WQ) {
        // wide R box along the last axis: one entry at a time
        for (int e = threadIdx.x; e < G::KPW * G::NN; e += 64) {
            const int kk = e / G::NN, idx = e % G::NN;
            const int r = idx / N, c = idx % N;
            const long k = blk * G::KPW + kk;
            cd v = mk(0, 0);
            if (r < n && c < n && k < a.nk) v = phase_last(a.ph, k, M, op * nn + r * n + c);
            lds[kk * Lay::MS + Lay::at(r, c)] = v;
        }
        return;
    }
    const int2 *kph = reinterpret_cast<const int2 *>(a.ph.kphase);
    const double2 *Yb = reinterpret_cast<const double2 *>(a.ph.Y);
    const double2 *Tb = reinterpret_cast<const double2 *>(a.ph.twL);
    // the lane's entries in chunks of CH, all r of a chunk in one unrolled
    // block, so that the loads of different entries are in flight together
    constexpr int CH = N >= 16 ? 4 : (Lay::CHUNK < IT ? Lay::CHUNK : IT);
    for_chunks<IT, CH>([&](auto i0c) {
        constexpr int i0 = decltype(i0c)::value;
        cd acc[CH];
        const double2 *yp[CH];
        const double2 *tp[CH];
        bool ok[CH];
#pragma unroll
        for (int u = 0; u < CH; u++) {
            const int e = threadIdx.x + 64 * (i0 + u);
            const int kk = e / G::NN, idx = e % G::NN;
            const int r = idx / N, c = idx % N;
            const long k = blk * G::KPW + kk;
            // masked entries read valid addresses and are zeroed at the end:
            // the loads stay unconditional.  The k of an entry is uniform over
            // its lanes when NN >= 64, so is its phase map.
            const bool krc = r < n && c < n;
            ok[u] = krc && k < a.nk;
            const int2 ph = kph[k < a.nk ? k : 0];
            yp[u] = Yb + (long)ph.x * M + (krc ? op * nn + r * n + c : 0);
            tp[u] = Tb + ph.y;
            acc[u] = mk(0, 0);
        }
        if constexpr (N > 4) {
#pragma unroll
            for (int q = 0; q < WQ; q++) {
                if (q < WL) {
#pragma unroll
                    for (int u = 0; u < CH; u++) {
                        const double2 tq = tp[u][q], yq = yp[u][(long)q * M];
                        acc[u] = cfma(mk(tq.x, tq.y), mk(yq.x, yq.y), acc[u]);
                    }
                }
            }
        } else {
            // few bands, registers to spare: QG cells of the last axis per
            // trip, their 2 QG CH loads in flight together (one latency per
            // trip instead of one per cell), the same products added in the
            // same order
            constexpr int QG = 8 / CH > 1 ? 8 / CH : 1;
            int q = 0;
            for (; q + QG <= WL; q += QG) {
                double2 tq[QG][CH], yq[QG][CH];
#pragma unroll
                for (int t = 0; t < QG; t++)
#pragma unroll
                    for (int u = 0; u < CH; u++) {
                        tq[t][u] = tp[u][q + t];
                        yq[t][u] = yp[u][(long)(q + t) * M];
                    }
#pragma unroll
                for (int t = 0; t < QG; t++)
#pragma unroll
                    for (int u = 0; u < CH; u++)
                        acc[u] = cfma(mk(tq[t][u].x, tq[t][u].y), mk(yq[t][u].x, yq[t][u].y), acc[u]);
            }
            for (; q < WL; q++) {
#pragma unroll
                for (int u = 0; u < CH; u++) {
                    const double2 tq = tp[u][q], yq = yp[u][(long)q * M];
                    acc[u] = cfma(mk(tq.x, tq.y), mk(yq.x, yq.y), acc[u]);
                }
            }
        }
#pragma unroll
        for (int u = 0; u < CH; u++) {
            const int e = threadIdx.x + 64 * (i0 + u);
            const int idx = e % G::NN;
            lds[(e / G::NN) * Lay::MS + Lay::at(idx / N, idx % N)] = ok[u] ? acc[u] : mk(0, 0);
        }
    });
}

// Few bands (n <= 4): H_0 into LDS and H_E into registers in one pass over
// the phase-sum cells, so that H_E's loads ride in the trips of H_0's (the
// same sums in the same order as two phase_mats calls).  he[i] is the lane's
// entry i in phase_mats' lane -> entry map (zero where masked); false, with
// nothing done, for an R box wider than 16 cells along the last axis.
template <int N, int LR, class Lay = SwLay<N>>
__device__ __forceinline__ bool phase_mats_pair(cd *lds, const StageArgs &a, long blk, cd *he)
{
    static_assert(N <= 4, "registers for both operators");
    using G = Geo<N, LR>;
    constexpr int IT = G::KPW * G::NN / 64;  // entries per lane
    const int n = a.n, nn = n * n, M = 2 * nn, WL = a.ph.WL;
    if (WL > 16) return false;
    const int2 *kph = reinterpret_cast<const int2 *>(a.ph.kphase);
    const double2 *Yb = reinterpret_cast<const double2 *>(a.ph.Y);
    const double2 *Tb = reinterpret_cast<const double2 *>(a.ph.twL);
    constexpr int CH = Lay::CHUNK < IT ? Lay::CHUNK : IT;
    constexpr int QG = 8 / CH > 1 ? 8 / CH : 1;
    for_chunks<IT, CH>([&](auto i0c) {
        constexpr int i0 = decltype(i0c)::value;
        cd acc0[CH], acc1[CH];
        const double2 *yp[CH];
        const double2 *tp[CH];
        bool ok[CH];
#pragma unroll
        for (int u = 0; u < CH; u++) {
            const int e = threadIdx.x + 64 * (i0 + u);
            const int kk = e / G::NN, idx = e % G::NN;
            const int r = idx / N, c = idx % N;
            const long k = blk * G::KPW + kk;
            const bool krc = r < n && c < n;
            ok[u] = krc && k < a.nk;
            const int2 ph = kph[k < a.nk ? k : 0];
            yp[u] = Yb + (long)ph.x * M + (krc ? r * n + c : 0);  // operator 0; operator 1 at + nn
            tp[u] = Tb + ph.y;
            acc0[u] = mk(0, 0);
            acc1[u] = mk(0, 0);
        }
        int q = 0;
        for (; q + QG <= WL; q += QG) {
            double2 tq[QG][CH], y0[QG][CH], y1[QG][CH];
#pragma unroll
            for (int t = 0; t < QG; t++)
#pragma unroll
                for (int u = 0; u < CH; u++) {
                    tq[t][u] = tp[u][q + t];
                    y0[t][u] = yp[u][(long)(q + t) * M];
                    y1[t][u] = yp[u][(long)(q + t) * M + nn];
                }
#pragma unroll
            for (int t = 0; t < QG; t++)
#pragma unroll
                for (int u = 0; u < CH; u++) {
                    const cd tw = mk(tq[t][u].x, tq[t][u].y);
                    acc0[u] = cfma(tw, mk(y0[t][u].x, y0[t][u].y), acc0[u]);
                    acc1[u] = cfma(tw, mk(y1[t][u].x, y1[t][u].y), acc1[u]);
                }
        }
        for (; q < WL; q++) {
#pragma unroll
            for (int u = 0; u < CH; u++) {
                const double2 tq = tp[u][q], y0 = yp[u][(long)q * M], y1 = yp[u][(long)q * M + nn];
                acc0[u] = cfma(mk(tq.x, tq.y), mk(y0.x, y0.y), acc0[u]);
                acc1[u] = cfma(mk(tq.x, tq.y), mk(y1.x, y1.y), acc1[u]);
            }
        }
#pragma unroll
        for (int u = 0; u < CH; u++) {
            const int e = threadIdx.x + 64 * (i0 + u);
            const int idx = e % G::NN;
            lds[(e / G::NN) * Lay::MS + Lay::at(idx / N, idx % N)] = ok[u] ? acc0[u] : mk(0, 0);
            he[i0 + u] = ok[u] ? acc1[u] : mk(0, 0);
        }
    });
    return true;
}

// the lane's entries (phase_mats' lane -> entry map) into LDS natural layout
template <int N, int LR, class Lay = SwLay<N>>
__device__ __forceinline__ void store_entries(cd *lds, const cd *v)
{
    using G = Geo<N, LR>;
    constexpr int IT = G::KPW * G::NN / 64;
#pragma unroll
    for (int i = 0; i < IT; i++) {
        const int e = threadIdx.x + 64 * i;
        const int idx = e % G::NN;
        lds[(e / G::NN) * Lay::MS + Lay::at(idx / N, idx % N)] = v[i];
    }
}

// stage input rho = x + sum_j c_j k_j with the number of derivatives NT
// fixed at compile time, so that the loads of x and of every k_j for a chunk
// of entries are issued before the first combination (one memory latency per
// chunk instead of one per derivative); the same sums in the same order as
// load_rho.  For few bands, where the registers for (NT + 1) x CH loads in
// flight are free.
template <int N, int LR, class Lay, int NT>
__device__ __forceinline__ void load_rho_nt(cd *lds, const StageArgs &a, long blk)
{
    using G = Geo<N, LR>;
    const int n = a.n;
    const StageInput &in = a.in;
    constexpr int IT = G::KPW * G::NN / 64;  // entries per lane
    constexpr int CH = Lay::CHUNK < IT ? Lay::CHUNK : IT;
    const double2 *src[NT + 1];
    double cf[NT + 1];
    src[0] = reinterpret_cast<const double2 *>(in.x);
#pragma unroll
    for (int j = 0; j < NT; j++) {
        src[j + 1] = reinterpret_cast<const double2 *>(in.k[j]);
        cf[j + 1] = in.c[j];
    }
    for_chunks<IT, CH>([&](auto i0c) {
        constexpr int i0 = decltype(i0c)::value;
        long idx_[CH];
        bool ok[CH];
#pragma unroll
        for (int u = 0; u < CH; u++) {
            const int e = threadIdx.x + 64 * (i0 + u);
            const int kk = e / G::NN, idx = e % G::NN;
            const int r = idx / N, c = idx % N;
            const long k = blk * G::KPW + kk;
            ok[u] = r < n && c < n && k < a.nk;
            // masked entries read state entry 8 (valid) and are zeroed at the end
            idx_[u] = ok[u] ? 8 + k * a.P + r * n + c : 8;
        }
        double2 v[NT + 1][CH];
#pragma unroll
        for (int j = 0; j <= NT; j++)
#pragma unroll
            for (int u = 0; u < CH; u++) v[j][u] = src[j][idx_[u]];
#pragma unroll
        for (int u = 0; u < CH; u++) {
            double re = v[0][u].x, im = v[0][u].y;
#pragma unroll
            for (int j = 1; j <= NT; j++) {
                re += cf[j] * v[j][u].x;
                im += cf[j] * v[j][u].y;
            }
            const int e = threadIdx.x + 64 * (i0 + u);
            const int idx = e % G::NN;
            lds[(e / G::NN) * Lay::MS + Lay::at(idx / N, idx % N)] = ok[u] ? mk(re, im) : mk(0, 0);
        }
    });
}

// stage input rho = x + sum_j c_j k_j (the Dormand-Prince stage combination,
// left to right as in the reference's scale_sum) into LDS natural layout
template <int N, int LR, class Lay = SwLay<N>>
__device__ __forceinline__ void load_rho(cd *lds, const StageArgs &a, long blk)
{
    if constexpr (N <= 4) {
        switch (a.in.nt) {
        case 0: load_rho_nt<N, LR, Lay, 0>(lds, a, blk); return;
        case 1: load_rho_nt<N, LR, Lay, 1>(lds, a, blk); return;
        case 2: load_rho_nt<N, LR, Lay, 2>(lds, a, blk); return;
        case 3: load_rho_nt<N, LR, Lay, 3>(lds, a, blk); return;
        case 4: load_rho_nt<N, LR, Lay, 4>(lds, a, blk); return;
        case 5: load_rho_nt<N, LR, Lay, 5>(lds, a, blk); return;
        default: load_rho_nt<N, LR, Lay, 6>(lds, a, blk); return;
        }
    }
    using G = Geo<N, LR>;
    const int n = a.n;
    const StageInput &in = a.in;
    constexpr int IT = G::KPW * G::NN / 64;  // entries per lane
    constexpr int CH = N >= 16 ? 4 : (Lay::CHUNK < IT ? Lay::CHUNK : IT);     // entries whose loads are in flight together
    for_chunks<IT, CH>([&](auto i0c) {
        constexpr int i0 = decltype(i0c)::value;
        long idx_[CH];
        bool ok[CH];
#pragma unroll
        for (int u = 0; u < CH; u++) {
            const int e = threadIdx.x + 64 * (i0 + u);
            const int kk = e / G::NN, idx = e % G::NN;
            const int r = idx / N, c = idx % N;
            const long k = blk * G::KPW + kk;
            ok[u] = r < n && c < n && k < a.nk;
            // masked entries read state entry 8 (valid) and are zeroed at the end
            idx_[u] = ok[u] ? 8 + k * a.P + r * n + c : 8;
        }
        double re[CH], im[CH];
#pragma unroll
        for (int u = 0; u < CH; u++) {
            const double2 xv = reinterpret_cast<const double2 *>(in.x)[idx_[u]];
            re[u] = xv.x;
            im[u] = xv.y;
        }
#pragma unroll
        for (int j = 0; j < 6; j++)
            if (j < in.nt) {
                const double2 *kj = reinterpret_cast<const double2 *>(in.k[j]);
                const double cj = in.c[j];
#pragma unroll
                for (int u = 0; u < CH; u++) {
                    const double2 kv = kj[idx_[u]];
                    re[u] += cj * kv.x;
                    im[u] += cj * kv.y;
                }
            }
#pragma unroll
        for (int u = 0; u < CH; u++) {
            const int e = threadIdx.x + 64 * (i0 + u);
            const int idx = e % G::NN;
            lds[(e / G::NN) * Lay::MS + Lay::at(idx / N, idx % N)] = ok[u] ? mk(re[u], im[u]) : mk(0, 0);
        }
    });
}

// the wave's matrices from LDS natural layout into a state vector (reference
// layout, zero pads), coalesced
template <int N, int LR, class Lay = SwLay<N>>
__device__ __forceinline__ void store_state(double *dst, const cd *lds, int n, int P, long nk, long blk)
{
    using G = Geo<N, LR>;
    const int nn = n * n;
    const long k0 = blk * G::KPW;
    for (int e = threadIdx.x; e < G::KPW * P; e += 64) {
        int kk = e / P, p = e % P;
        long k = k0 + kk;
        if (k >= nk) break;
        cd v = mk(0, 0);
        if (p < nn) v = lds[kk * Lay::MS + Lay::at(p / n, p % n)];
        reinterpret_cast<double2 *>(dst)[8 + k * P + p] = make_double2(v.x, v.y);
    }
}

// out[c] = sum_l X[i][l] Y[l][j_c] with X, Y in LDS (natural layout);
// XOP 1 uses X[i][l] := conj(Xs[l][i]); YOP 1 uses Y[l][j] := conj(Ys[j][l]).
template <int N, int LR, int XOP, int YOP>
__device__ __forceinline__ void lds_product(const cd *X, const cd *Y, const Lane &L, cd *out)
{
    using G = Geo<N, LR>;
#pragma unroll
    for (int c = 0; c < G::C; c++) out[c] = mk(0, 0);
#pragma unroll 2
    for (int l = 0; l < N; l++) {
        cd x = (XOP == 0) ? X[sw<N>(L.i, l)] : cconj(X[sw<N>(l, L.i)]);
#pragma unroll
        for (int c = 0; c < G::C; c++) {
            int j = L.i ^ (L.h * G::C + c);
            cd y = (YOP == 0) ? Y[sw<N>(l, j)] : cconj(Y[sw<N>(j, l)]);
            out[c] = cfma(x, y, out[c]);
        }
    }
}

template <int N, int LR>
__device__ __forceinline__ void store_xrow(cd *lds, const Lane &L, const cd *v)
{
    using G = Geo<N, LR>;
#pragma unroll
    for (int c = 0; c < G::C; c++) lds[sw<N>(L.i, L.i ^ (L.h * G::C + c))] = v[c];
}

template <int N, int LR>
__device__ __forceinline__ void load_xrow(const cd *lds, const Lane &L, cd *v)
{
    using G = Geo<N, LR>;
#pragma unroll
    for (int c = 0; c < G::C; c++) v[c] = lds[sw<N>(L.i, L.i ^ (L.h * G::C + c))];
}

// The Jacobi's eigenvector registers: lane (i, h) register c holds
// U[i ^ (h C + c)][i] (the xrow layout of U^T).  A rotation of rows
// (j, j ^ r) of U then combines register c with register c ^ (r mod C) of the
// same lane (r < C) or of lane h ^ (r / C) in the row's own lane group, and it
// uses the coefficients the column update of A reads anyway, instead of an
// exchange with the partner row's lanes (for 16 bands at 4 lanes per row this
// replaces the U half of the ds_bpermute exchanges of rounds 4-15 by DPP moves
// inside the lane group: 2.14 ms per 48^3 launch against 2.20 with U kept in
// rows).  U registers <-> LDS natural layout:
template <int N, int LR>
__device__ __forceinline__ void store_u(cd *lds, const Lane &L, const cd *v)
{
    using G = Geo<N, LR>;
#pragma unroll
    for (int c = 0; c < G::C; c++) lds[sw<N>(L.i ^ (L.h * G::C + c), L.i)] = v[c];
}

template <int N, int LR>
__device__ __forceinline__ void load_u(const cd *lds, const Lane &L, cd *v)
{
    using G = Geo<N, LR>;
#pragma unroll
    for (int c = 0; c < G::C; c++) v[c] = lds[sw<N>(L.i ^ (L.h * G::C + c), L.i)];
}

// padded rows and columns (index >= n) of U are the identity
template <int N, int LR>
__device__ __forceinline__ void u_pad_identity(const Lane &L, int n, cd *v)
{
    using G = Geo<N, LR>;
#pragma unroll
    for (int c = 0; c < G::C; c++) {
        const int g = L.h * G::C + c;
        if (L.i >= n || (L.i ^ g) >= n) v[c] = mk(g == 0 ? 1.0 : 0.0, 0);
    }
}

// U registers of the Jacobi -> xrow layout (rows), through the k-point's
// N x N LDS region lds (free on entry, rewritten)
template <int N, int LR>
__device__ __forceinline__ void u_rows(cd *lds, const Lane &L, cd *v)
{
    store_u<N, LR>(lds, L, v);
    wsync();
    load_xrow<N, LR>(lds, L, v);
    wsync();
}

// ---- 16 x 16 complex products on the FP64 matrix cores (16-band layouts) --
// v_mfma_f64_16x16x4f64: A (16 x 4) lane l holds A[l % 16][l / 16], B (4 x 16)
// lane l holds B[l / 16][l % 16], the 16 x 16 accumulator lane l holds rows
// 4 j + l / 16, j = 0..3, of column l % 16.  A complex product is four real
// products of 16 x 16 x 16, each four MFMA steps along the inner index.
typedef double d4 __attribute__((ext_vector_type(4)));
struct CAcc {
    d4 r, i;
};

__device__ __forceinline__ void cacc_zero(CAcc &c)
{
    c.r = d4{0, 0, 0, 0};
    c.i = d4{0, 0, 0, 0};
}

// Operands of one complex product, four inner-index steps: left a[kc] =
// X[l % 16][4 kc + l / 16], right b[kc] = Y[4 kc + l / 16][l % 16].
// From one k-point's LDS matrix in natural layout (XOP 1: X[i][l] :=
// conj(Xs[l][i]); YOP 1: Y[l][j] := conj(Ys[j][l])):
template <int XOP>
__device__ __forceinline__ void lds_left(const cd *X, cd *a)
{
    const int l = threadIdx.x & 63, r = l & 15, q = l >> 4;
#pragma unroll
    for (int kc = 0; kc < 4; kc++) {
        const int k = 4 * kc + q;
        a[kc] = (XOP == 0) ? X[sw<16>(r, k)] : cconj(X[sw<16>(k, r)]);
    }
}
template <int YOP>
__device__ __forceinline__ void lds_right(const cd *Y, cd *b)
{
    const int l = threadIdx.x & 63, r = l & 15, q = l >> 4;
#pragma unroll
    for (int kc = 0; kc < 4; kc++) {
        const int k = 4 * kc + q;
        b[kc] = (YOP == 0) ? Y[sw<16>(k, r)] : cconj(Y[sw<16>(r, k)]);
    }
}
// From a finished accumulator M: accumulator register kc holds
// M[4 kc + l / 16][l % 16], which is the right operand M itself, and, conjugated,
// the left operand M^dagger; a product can so consume the previous one's
// result without a trip through LDS.
__device__ __forceinline__ void acc_right(const CAcc &m, cd *b)
{
#pragma unroll
    for (int kc = 0; kc < 4; kc++) b[kc] = mk(m.r[kc], m.i[kc]);
}
__device__ __forceinline__ void acc_left_dagger(const CAcc &m, cd *a)
{
#pragma unroll
    for (int kc = 0; kc < 4; kc++) a[kc] = mk(m.r[kc], -m.i[kc]);
}

// acc += X Y from operands a, b as above
__device__ __forceinline__ void mfma_cmul(const cd *a, const cd *b, CAcc &acc)
{
#pragma unroll
    for (int kc = 0; kc < 4; kc++) {
        acc.r = __builtin_amdgcn_mfma_f64_16x16x4f64(a[kc].x, b[kc].x, acc.r, 0, 0, 0);
        acc.i = __builtin_amdgcn_mfma_f64_16x16x4f64(a[kc].x, b[kc].y, acc.i, 0, 0, 0);
        acc.r = __builtin_amdgcn_mfma_f64_16x16x4f64(-a[kc].y, b[kc].y, acc.r, 0, 0, 0);
        acc.i = __builtin_amdgcn_mfma_f64_16x16x4f64(a[kc].y, b[kc].x, acc.i, 0, 0, 0);
    }
}

// acc += SX * X Y for one k-point's matrices in LDS natural layout; every
// operand is read before the first matrix instruction
template <int XOP, int YOP, int SX = 1>
__device__ __forceinline__ void mfma_cprod(const cd *X, const cd *Y, CAcc &acc)
{
    cd a[4], b[4];
    lds_left<XOP>(X, a);
    lds_right<YOP>(Y, b);
    if (SX < 0) {
#pragma unroll
        for (int kc = 0; kc < 4; kc++) a[kc] = mk(-a[kc].x, -a[kc].y);
    }
    mfma_cmul(a, b, acc);
}

// accumulator element j of this lane: row and column
__device__ __forceinline__ int cacc_row(int j) { return 4 * j + ((threadIdx.x & 63) >> 4); }
__device__ __forceinline__ int cacc_col() { return threadIdx.x & 15; }

// the accumulator into LDS natural layout (one k-point)
__device__ __forceinline__ void cacc_store(cd *lds, const CAcc &c)
{
#pragma unroll
    for (int j = 0; j < 4; j++) lds[sw<16>(cacc_row(j), cacc_col())] = mk(c.r[j], c.i[j]);
}

// 1/sqrt(x) for a positive normal x: the hardware estimate and one
// second-order correction y (1 + e/2 + 3 e^2/8), e = 1 - x y^2 (the device
// library's rsqrt without its handling of zero, infinity and NaN)
__device__ __forceinline__ double rsqrt_pos(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    const double e = fma(-x * y, y, 1.0);
    return fma(y * e, fma(e, 0.375, 0.5), y);
}

// an optimisation barrier: the N accumulators are complete here, and no
// memory access moves across it (keeps a product's loads per step bounded)
template <int N>
__device__ __forceinline__ void pin_rows(cd *v)
{
    if constexpr (N == 8) {
        asm volatile("" : "+v"(v[0].x), "+v"(v[0].y), "+v"(v[1].x), "+v"(v[1].y), "+v"(v[2].x), "+v"(v[2].y),
                     "+v"(v[3].x), "+v"(v[3].y), "+v"(v[4].x), "+v"(v[4].y), "+v"(v[5].x), "+v"(v[5].y),
                     "+v"(v[6].x), "+v"(v[6].y), "+v"(v[7].x), "+v"(v[7].y)::"memory");
    } else {
#pragma unroll
        for (int c = 0; c < N; c++) asm volatile("" : "+v"(v[c].x), "+v"(v[c].y)::"memory");
    }
}

// One round r of the parallel Jacobi sweep: rotations on the pairs (p, q = p ^ r).
// With a_pq = m w (m = |a_pq|, |w| = 1) the rotation of a pair is
//   J = [[c, s w], [-s conj(w), c]],  (c, s) = (cos t, sin t), tan 2t = 2m / (a_pp - a_qq),
// |t| <= pi/4, which zeroes (J A J^dagger)_pq; its diagonal is real.  It acts on rows,
// row p <- c row p + s w row q and row q <- c row q - s conj(w) row p, and as
// J^dagger on columns, column j <- J_jj column j + conj(J_j,j^r) column j ^ r.
// The lane of row p that holds a_pq computes the rotation and stores both
// rows' coefficients (J_jj, J_j,j^r) under index j (the identity when the
// pair does not rotate), so every lane reads its row's and its columns'
// coefficients without selects or divergent branches; an identity
// coefficient pair reproduces its operand exactly.
template <int N, int LR, int R>
__device__ __forceinline__ void jac_round(cd *A, cd *U, const Lane &L, double thr, double *Dg, cd *Ja, cd *Jb,
                                          int *Fa, int &anyrot)
{
    using G = Geo<N, LR>;
    constexpr int C = G::C;
    constexpr int rc = R % C, rh = R / C;
    constexpr int mrow = R * LR + rh;   // lane mask of the row partner (row i ^ R)
    const int base = L.kk * N;
    // (a) rotation of the pair (p, q = p ^ R), p < q: lane h == rh of row p
    // holds a_pq in register rc
    {
        const bool isp = L.i < (L.i ^ R);
        const cd e = A[rc];
        const double thr2 = thr * thr;
        const bool own = L.h == rh && isp;
        // a round in which no pair of the wavefront rotates leaves A and U
        // unchanged: skip it before computing any rotation (the test is
        // wave-uniform, so the exchanges below always run in all lanes: a DPP
        // read from a lane that is switched off returns 0)
        if (!__any(own && L.valid && fma(e.x, e.x, e.y * e.y) > thr2)) return;
        anyrot = 1;
        if (own) {
            const cd apq = e;
            const double m2 = fma(apq.x, apq.x, apq.y * apq.y);
            const int act = L.valid && (m2 > thr2);
            double ja = 1, jbx = 0, jby = 0;
            if (act) {
                // the rotation through |theta| <= pi/4 that zeroes a_pq:
                // tan 2 theta = 2 |a_pq| / d, d = a_pp - a_qq.  With
                // r = sqrt(d^2 + 4 |a_pq|^2) and D = |d| + r (no cancellation),
                // c = D / sqrt(D^2 + 4 |a_pq|^2) and s w = sign(d) 2 a_pq / sqrt(D^2 + 4 |a_pq|^2):
                // two reciprocal square roots, no division, and c^2 + |s w|^2 = 1
                // to rounding whatever the error of r
                const double d = Dg[base + L.i] - Dg[base + (L.i ^ R)];
                const double m4 = 4 * m2;
                const double X = fma(d, d, m4);
                const double D = fabs(d) + X * rsqrt_pos(X);
                const double q = rsqrt_pos(fma(D, D, m4));
                const double tq = copysign(2 * q, d);
                ja = D * q;
                jbx = tq * apq.x;
                jby = tq * apq.y;
            }
            // row p: (c, s w); row q: (c, -s conj(w))
            Ja[base + L.i] = mk(ja, 0);
            Jb[base + L.i] = mk(jbx, jby);
            Ja[base + (L.i ^ R)] = mk(ja, 0);
            Jb[base + (L.i ^ R)] = mk(-jbx, jby);
            Fa[base + L.i] = act;
        }
    }
    wsync();
    const int arow = Fa[base + min(L.i, L.i ^ R)];
    const double ca = Ja[base + L.i].x;
    const cd cb = Jb[base + L.i];
#pragma unroll
    for (int c = 0; c < C; c++) {
        const int c2 = c ^ rc;
        if (c2 < c) continue;
        // row update of the register pair (c, c2); the partner row holds them in (c2, c)
        cd a0 = A[c], a1 = A[c2];
        const cd b0 = xchg<mrow>(A[c2], L.l4), b1 = xchg<mrow>(A[c], L.l4);
        a0 = cfma(cb, b0, cscale(a0, ca));
        if (c2 != c) {
            a1 = cfma(cb, b1, cscale(a1, ca));
        } else {
            a1 = a0;
        }
        // column update: partner of column j is column j ^ R, register c ^ rc
        cd d0, d1;
        if constexpr (rh == 0) {
            d0 = a1;
            d1 = a0;
        } else {
            d0 = xchg<rh>(a1, L.l4);
            d1 = xchg<rh>(a0, L.l4);
        }
        const int j0 = L.i ^ (L.h * C + c);
        const double ja = Ja[base + j0].x;
        const cd jb = Jb[base + j0];
        {
            // U^T registers: row j0 of U (register c) pairs with row j0 ^ R (register c2 of lane h ^ rh)
            const cd u0 = U[c], u1 = U[c2];
            cd e0, e1;
            if constexpr (rh == 0) {
                e0 = u1;
                e1 = u0;
            } else {
                e0 = xchg<rh>(u1, L.l4);
                e1 = xchg<rh>(u0, L.l4);
            }
            U[c] = cfma(jb, e0, cscale(u0, ja));
            if (c2 != c) {
                if constexpr (rh == 0) {
                    // row j1 = j0 ^ R: (J_j1j1, J_j1j0) = (J_j0j0, -conj(J_j0j1))
                    U[c2] = cfma(mk(-jb.x, jb.y), e1, cscale(u1, ja));
                } else {
                    const int j1 = L.i ^ (L.h * C + c2);
                    U[c2] = cfma(Jb[base + j1], e1, cscale(u1, Ja[base + j1].x));
                }
            }
        }
        cd r0 = cfma_cj(jb, d0, cscale(a0, ja));
        // the rotated pivot (i, i ^ R) is set to zero: it sits in register rc of lanes h == rh
        if (c == rc && L.h == rh && arow) r0 = mk(0, 0);
        A[c] = r0;
        if (c2 != c) {
            cd r1;
            if constexpr (rh == 0) {
                // column j1 = j0 ^ R is j0's partner: (J_j1j1, J_j1j0) = (J_j0j0, -conj(J_j0j1))
                r1 = cfma(mk(-jb.x, -jb.y), d1, cscale(a1, ja));
            } else {
                // column j1 = i ^ (h C + c2) belongs to another pair (its partner is in lane h ^ rh)
                const int j1 = L.i ^ (L.h * C + c2);
                r1 = cfma_cj(Jb[base + j1], d1, cscale(a1, Ja[base + j1].x));
            }
            if (c2 == rc && L.h == rh && arow) r1 = mk(0, 0);
            A[c2] = r1;
        }
    }
    if (L.h == 0) Dg[base + L.i] = A[0].x;
    wsync();
}

template <int N, int LR, int... Rm1>
__device__ __forceinline__ void jac_sweep(std::integer_sequence<int, Rm1...>, cd *A, cd *U, const Lane &L, double thr,
                                          double *Dg, cd *Ja, cd *Jb, int *Fa, int &anyrot)
{
    (jac_round<N, LR, Rm1 + 1>(A, U, L, thr, Dg, Ja, Jb, Fa, anyrot), ...);
}

// Wavefront Jacobi on the xrow matrix A (Hermitian) -> U (the transposed
// register layout of store_u / load_u), eigenvalues in Dg[kk*N + i].  LDS
// scratch: Dg[KPW*N], Ja/Jb[KPW*N] (rotation coefficients), Fa[KPW*N].
template <int N, int LR>
__device__ __forceinline__ void wave_jacobi(cd *A, cd *U, const Lane &L0, int sweeps, double *Dg, cd *Ja, cd *Jb, int *Fa,
                            bool initU = true)
{
    using G = Geo<N, LR>;
    constexpr int C = G::C;
    if (initU) {
#pragma unroll
        for (int c = 0; c < C; c++) U[c] = mk((L0.h * C + c) == 0 ? 1.0 : 0.0, 0);
    }
    if (L0.h == 0) Dg[L0.kk * N + L0.i] = A[0].x;
    wsync();
    for (int sw = 0; sw < sweeps; sw++) {
        // re-materialise the lane coordinates each sweep so that the per-round
        // LDS addresses are not hoisted out of the sweep loop (register pressure)
        Lane L = L0;
        asm volatile("" : "+v"(L.i), "+v"(L.h), "+v"(L.kk), "+v"(L.l4));
        double s = 0, m2 = 0;
#pragma unroll
        for (int c = 0; c < C; c++) {
            int j = L.i ^ (L.h * C + c);
            if (L.i <= j) s += fabs(A[c].x) + fabs(A[c].y);
            if (L.i != j) m2 = fmax(m2, fma(A[c].x, A[c].x, A[c].y * A[c].y));
        }
        s = seg_sum<G::LK>(s);
        const double thr = 1e-15 * (0.5 * s);
        // no off-diagonal entry above the threshold: no pair of this sweep
        // rotates (a round without rotations changes nothing), so the sweep
        // ends here instead of testing its N - 1 rounds one by one
        if (!__any(L.valid && m2 > thr * thr)) break;
        int anyrot = 0;
        jac_sweep<N, LR>(std::make_integer_sequence<int, N - 1>{}, A, U, L, thr, Dg, Ja, Jb, Fa, anyrot);
        if (!__any(anyrot)) break;
    }
}

// eigenvector rows from LDS natural layout into the per-k cache [k][n][n],
// with the lane -> entry mapping of load_mats
template <int N, int LR, class Lay = SwLay<N>>
__device__ __forceinline__ void store_ucache(const StageArgs &a, const cd *lds, long blk)
{
    using G = Geo<N, LR>;
    const int n = a.n, nn = n * n;
#pragma unroll
    for (int e = threadIdx.x; e < G::KPW * G::NN; e += 64) {
        const int kk = e / G::NN, idx = e % G::NN;
        const int r = idx / N, c = idx % N;
        const long k = blk * G::KPW + kk;
        if (r < n && c < n && k < a.nk) {
            const cd v = lds[kk * Lay::MS + Lay::at(r, c)];
            reinterpret_cast<double2 *>(a.ucache)[k * nn + r * n + c] = make_double2(v.x, v.y);
        }
    }
}

// Error norm of the attempt over this block's k-points (last stage only): the
// stage input's x, k1, k3..k6 are read again (HBM that the VALU-bound kernel
// leaves idle), k7 = d rho/dt comes from LDS natural layout; same lane ->
// entry mapping as load_rho, two entries' loads in flight at a time.
template <int N, int LR, class Lay = SwLay<N>>
__device__ __forceinline__ double stage_err(const StageArgs &a, const cd *lds, long blk)
{
    using G = Geo<N, LR>;
    const int n = a.n;
    const long k0 = blk * G::KPW;
    const double2 *x2 = reinterpret_cast<const double2 *>(a.in.x);
    double v = 0;
    constexpr int IT = G::KPW * G::NN / 64;
#pragma unroll 2
    for (int u = 0; u < IT; u++) {
        const int e = threadIdx.x + 64 * u;
        const int kk = e / G::NN, idx = e % G::NN, r = idx / N, c = idx % N;
        const long k = k0 + kk;
        if (r < n && c < n && k < a.nk) {
            const long i = 8 + k * a.P + r * n + c;
            double2 kv[6];
#pragma unroll
            for (int j = 0; j < 5; j++) kv[j] = reinterpret_cast<const double2 *>(a.in.k[j])[i];
            const cd d = lds[kk * Lay::MS + Lay::at(r, c)];
            kv[5] = make_double2(d.x, d.y);
            v = fmax(v, err_entry(x2[i], kv, a.err_e, a.err_ea, a.err_er, a.err_dt));
        }
    }
    if (blk == 0 && threadIdx.x < 8) {
        // special slots: d/dt of A_t is -E_t (t < dim), the rest 0
        const int t = threadIdx.x;
        double2 kv[6];
#pragma unroll
        for (int j = 0; j < 5; j++) kv[j] = reinterpret_cast<const double2 *>(a.in.k[j])[t];
        kv[5] = make_double2(t < a.dim ? -a.E[t] : 0.0, 0.0);
        v = fmax(v, err_entry(x2[t], kv, a.err_e, a.err_ea, a.err_er, a.err_dt));
    }
    return v;
}

// special slots of the stage: d A_t/dt = -E_t (t < dim), and the stage input
// itself where it is stored (block 0)
__device__ __forceinline__ void stage_specials(const StageArgs &a, const long blk)
{
    if (blk == 0 && threadIdx.x < 8) {
        const int t = threadIdx.x;
        double v = t < a.dim ? -a.E[t] : 0.0;
        a.dxdt[2 * t] = v;
        a.dxdt[2 * t + 1] = 0.0;
        if (a.xout) {
            double re = a.in.x[2 * t], im = a.in.x[2 * t + 1];
#pragma unroll
            for (int j = 0; j < 6; j++)
                if (j < a.in.nt) {
                    re += a.in.c[j] * a.in.k[j][2 * t];
                    im += a.in.c[j] * a.in.k[j][2 * t + 1];
                }
            a.xout[2 * t] = re;
            a.xout[2 * t + 1] = im;
        }
    }
}

// The stage for 16-band layouts: the Jacobi in the XOR-wrapped register
// layout, every matrix product on the matrix cores, one k-point of the block
// at a time with all 64 lanes.  A product's result stays in its accumulator
// and feeds the next product as an operand (acc_right, acc_left_dagger), so
// two LDS matrices per k-point suffice (8.9 KB for one k-point per block:
// four wavefronts per SIMD fit).  Same operations as the generic body below.
template <int LR>
__device__ __forceinline__ void rhs_stage_body16(const StageArgs &a, const long blk)
{
    constexpr int N = 16;
    using G = Geo<N, LR>;
    constexpr int C = G::C, KPW = G::KPW;
    __shared__ cd S0[KPW * G::NN];
    __shared__ cd S1[KPW * G::NN];
    __shared__ double Dg[KPW * N];
    __shared__ cd Ja[KPW * N], Jb[KPW * N];
    __shared__ int Fa[KPW * N];

    const Lane L = lane_id<N, LR>(a.nk, blk);
    const int n = a.n, nn = n * n;
    cd *s0 = S0 + L.kk * G::NN, *s1 = S1 + L.kk * G::NN;
    stage_specials(a, blk);

    if (a.dephase) {
        cd A[C], U[C];
        if (a.warm) {
            // warm start: A = U' (H0 U'^dagger) with U' the previous eigenvectors
            load_mats<N, LR>(S0, a.ucache, nn, 0, n, a.nk, blk);
            phase_mats<N, LR>(S1, a, 0, blk);  // H_0 -> S1
            wsync();
            load_u<N, LR>(s0, L, U);
            u_pad_identity<N, LR>(L, n, U);
            CAcc T[KPW];
#pragma unroll
            for (int kk = 0; kk < KPW; kk++) {
                cd x[4], y[4];
                CAcc t;
                cacc_zero(t);
                mfma_cprod<0, 1>(S1 + kk * G::NN, S0 + kk * G::NN, t);  // H0 U'^dagger
                lds_left<0>(S0 + kk * G::NN, x);
                acc_right(t, y);
                cacc_zero(T[kk]);
                mfma_cmul(x, y, T[kk]);  // U' (H0 U'^dagger)
            }
            wsync();
#pragma unroll
            for (int kk = 0; kk < KPW; kk++) cacc_store(S0 + kk * G::NN, T[kk]);
            wsync();
            load_xrow<N, LR>(s0, L, A);
            wave_jacobi<N, LR>(A, U, L, a.sweeps, Dg, Ja, Jb, Fa, false);
        } else {
            phase_mats<N, LR>(S1, a, 0, blk);  // H_0 -> S1
            wsync();
            load_xrow<N, LR>(s1, L, A);
            wave_jacobi<N, LR>(A, U, L, a.sweeps, Dg, Ja, Jb, Fa);
        }
        store_u<N, LR>(s1, L, U);  // U -> S1
        wsync();
        store_ucache<N, LR>(a, S1, blk);
    }
    load_rho<N, LR>(S0, a, blk);  // rho -> S0 (S1 still holds U)
    wsync();
    if (a.xout) store_state<N, LR>(a.xout, S0, n, a.P, a.nk, blk);
    // the dephasing first, from rho in S0 and U in S1 through the
    // accumulators, then the coherent part, accumulated onto its result
    CAcc D[KPW];
#pragma unroll
    for (int kk = 0; kk < KPW; kk++) cacc_zero(D[kk]);
    if (a.dephase) {
        // U^dagger (D o (U rho U^dagger))^dagger U: the bracket is Hermitian,
        // and its conjugate transpose is what an accumulator hands over as a
        // left operand
#pragma unroll
        for (int kk = 0; kk < KPW; kk++) {
            const cd *sr = S0 + kk * G::NN, *su = S1 + kk * G::NN;
            cd x[4], y[4];
            CAcc t, h;
            cacc_zero(t);
            mfma_cprod<0, 1>(sr, su, t);  // t = rho U^dagger
            lds_left<0>(su, x);
            acc_right(t, y);
            cacc_zero(h);
            mfma_cmul(x, y, h);  // rhoH = U t
            const int b = cacc_col();
            const double eb = Dg[kk * N + b];
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int ar = cacc_row(j);
                double gam;
                if (a.w < 0) {
                    gam = (ar != b) ? 1.0 : 0.0;
                } else {
                    const double dE = (Dg[kk * N + ar] - eb) / a.w;
                    gam = 1 - exp(-dE * dE);
                }
                const double f = -gam * a.F2;
                h.r[j] *= f;
                h.i[j] *= f;
            }
            acc_left_dagger(h, x);
            lds_right<0>(su, y);
            cacc_zero(t);
            mfma_cmul(x, y, t);  // S = rhoH'^dagger U
            lds_left<1>(su, x);
            acc_right(t, y);
            mfma_cmul(x, y, D[kk]);  // U^dagger S
        }
        wsync();
    }
    phase_mats<N, LR>(S1, a, 1, blk);  // H_E -> S1
    wsync();
    // coherent part -i (H_E rho - rho H_E): D + (-i) X = D + (X.i, -X.r), so
    // it is accumulated as i (rho H_E - H_E rho) with the roles of the real and
    // imaginary accumulators exchanged
#pragma unroll
    for (int kk = 0; kk < KPW; kk++) {
        CAcc t;
        t.r = D[kk].i;
        t.i = -D[kk].r;
        mfma_cprod<0, 0, -1>(S1 + kk * G::NN, S0 + kk * G::NN, t);  // - H_E rho
        mfma_cprod<0, 0>(S0 + kk * G::NN, S1 + kk * G::NN, t);      // + rho H_E
        // t = i D + (rho H_E - H_E rho) = i (D - i (H_E rho - rho H_E))
        D[kk].r = -t.i;
        D[kk].i = t.r;
    }
    // d rho/dt (with zero pads) through LDS for coalescing
    wsync();
#pragma unroll
    for (int kk = 0; kk < KPW; kk++) cacc_store(S0 + kk * G::NN, D[kk]);
    wsync();
    store_state<N, LR>(a.dxdt, S0, n, a.P, a.nk, blk);
    if (a.err_partial) {
        const double v = seg_max<64>(stage_err<N, LR>(a, S0, blk));
        if (threadIdx.x == 0) a.err_partial[blk] = v;
    }
}

// ---- one lane per matrix row (LR = 1, N <= 8): one LDS matrix per k-point --
// Lane i of a k-point holds row i of a matrix in registers, register c the
// entry (i, i ^ c) ("rows"), or column i, register c the entry (i ^ c, i)
// ("columns").  A product takes its left operand from registers and its right
// operand from the k-point's LDS matrix in SlotLay, so the stage needs one
// LDS matrix per k-point instead of two (three waves per SIMD fit in LDS).
template <int N>
__device__ __forceinline__ void slot_store_rows(cd *s, int i, const cd *v)
{
#pragma unroll
    for (int c = 0; c < N; c++) s[i * SlotLay<N>::RS + c] = v[c];
}
template <int N>
__device__ __forceinline__ void slot_load_rows(const cd *s, int i, cd *v)
{
#pragma unroll
    for (int c = 0; c < N; c++) v[c] = s[i * SlotLay<N>::RS + c];
}
template <int N>
__device__ __forceinline__ void slot_store_cols(cd *s, int i, const cd *v)
{
#pragma unroll
    for (int c = 0; c < N; c++) s[(i ^ c) * SlotLay<N>::RS + c] = v[c];
}
template <int N>
__device__ __forceinline__ void slot_load_cols(const cd *s, int i, cd *v)
{
#pragma unroll
    for (int c = 0; c < N; c++) v[c] = s[(i ^ c) * SlotLay<N>::RS + c];
}

// out[c] = sum_l X[i][l] Y[l][i ^ c], l = i ^ c'.  X from registers: XOP 0,
// xr = rows of X; XOP 1, xr = columns of V and X = V^dagger.  Y from LDS
// (SlotLay): YOP 0, Y = Ys; YOP 1, Y = Ys^dagger.  One register's loads are
// in flight at a time (the scheduler may not hoist the next ones), which
// keeps the product's footprint at one row of operands.
template <int N, int XOP, int YOP>
__device__ __forceinline__ void reg_product(const cd *xr, const cd *s, int i, cd *out)
{
    constexpr int RS = SlotLay<N>::RS;
#pragma unroll
    for (int c = 0; c < N; c++) out[c] = mk(0, 0);
#pragma unroll
    for (int cp = 0; cp < N; cp++) {
        const cd x = XOP ? cconj(xr[cp]) : xr[cp];
        if constexpr (YOP == 0) {
            // Y[l][j], l = i ^ cp, j = i ^ c: row l, slot cp ^ c
            const cd *row = s + (i ^ cp) * RS;
#pragma unroll
            for (int c = 0; c < N; c++) out[c] = cfma(x, row[cp ^ c], out[c]);
        } else {
            // conj(Ys[j][l]): row j = i ^ c, slot c ^ cp
#pragma unroll
            for (int c = 0; c < N; c++) out[c] = cfma_jc(x, s[(i ^ c) * RS + (c ^ cp)], out[c]);
        }
        pin_rows<N>(out);
    }
}

template <int N, bool PAIR>
__device__ __forceinline__ void rhs_stage_body_r1(const StageArgs &a, const long blk)
{
    using G = Geo<N, 1>;
    using Lay = SlotLay<N>;
    constexpr int C = N, KPW = G::KPW, RS = Lay::RS;
    __shared__ cd S[KPW * Lay::MS];
    __shared__ double Dg[KPW * N];
    __shared__ cd Ja[KPW * N], Jb[KPW * N];
    __shared__ int Fa[KPW * N];

    const Lane L = lane_id<N, 1>(a.nk, blk);
    const int n = a.n, nn = n * n;
    cd *s = S + L.kk * Lay::MS;
    stage_specials(a, blk);

    cd U[C];   // eigenvectors, columns
    cd T1[C];  // U rho, rows
    // few bands: H_E formed with H0, kept in registers until it is needed
    constexpr int HE = N <= 4 ? KPW * G::NN / 64 : 1;
    cd he[HE];
    bool he_ready = false;
    auto form_h0 = [&]() {  // H0 -> S
        if constexpr (N <= 4 && PAIR) he_ready = phase_mats_pair<N, 1, Lay>(S, a, blk, he);
        if (!he_ready) phase_mats<N, 1, Lay>(S, a, 0, blk);
    };
    if (a.dephase) {
        cd A[C];
        if (a.warm) {
            // warm start: A = U' H0 U'^dagger with U' the previous eigenvectors
            load_mats<N, 1, Lay>(S, a.ucache, nn, 0, n, a.nk, blk);  // U' -> S
            wsync();
            cd T[C];
            slot_load_rows<N>(s, L.i, T);
            slot_load_cols<N>(s, L.i, U);
            // padded rows and columns of U' are the identity
            u_pad_identity<N, 1>(L, n, T);
            u_pad_identity<N, 1>(L, n, U);
            wsync();
            form_h0();
            wsync();
            reg_product<N, 0, 0>(T, s, L.i, A);  // U' H0
            wsync();
            slot_store_cols<N>(s, L.i, U);  // U' -> S
            wsync();
#pragma unroll
            for (int c = 0; c < C; c++) T[c] = A[c];
            reg_product<N, 0, 1>(T, s, L.i, A);  // (U' H0) U'^dagger
            wave_jacobi<N, 1>(A, U, L, a.sweeps, Dg, Ja, Jb, Fa, false);
        } else {
            form_h0();
            wsync();
            slot_load_rows<N>(s, L.i, A);
            wave_jacobi<N, 1>(A, U, L, a.sweeps, Dg, Ja, Jb, Fa);
        }
        wsync();
        slot_store_cols<N>(s, L.i, U);  // U -> S
        wsync();
        store_ucache<N, 1, Lay>(a, S, blk);
        slot_load_rows<N>(s, L.i, A);  // U rows
        wsync();
        load_rho<N, 1, Lay>(S, a, blk);  // rho -> S
        wsync();
        reg_product<N, 0, 0>(A, s, L.i, T1);  // U rho
    } else {
        load_rho<N, 1, Lay>(S, a, blk);  // rho -> S
        wsync();
    }
    if (a.xout) store_state<N, 1, Lay>(a.xout, S, n, a.P, a.nk, blk);
    cd dr[C];
    {
        // coherent part -i (H_E rho - rho H_E), with H_E rho = (rho H_E)^dagger
        cd rr[C], T[C];
        slot_load_rows<N>(s, L.i, rr);
        wsync();
        if (he_ready) store_entries<N, 1, Lay>(S, he);  // H_E -> S
        else phase_mats<N, 1, Lay>(S, a, 1, blk);
        wsync();
        reg_product<N, 0, 0>(rr, s, L.i, T);  // rho H_E
        wsync();
        slot_store_rows<N>(s, L.i, T);
        wsync();
#pragma unroll
        for (int c = 0; c < C; c++) {
            const cd th = cconj(s[(L.i ^ c) * RS + c]);
            const cd d = csub(th, T[c]);
            dr[c] = mk(d.y, -d.x);
        }
    }
    if (a.dephase) {
        // dephasing in the eigenbasis: U^dagger (D o (U rho U^dagger)) U
        cd T[C];
        wsync();
        slot_store_cols<N>(s, L.i, U);  // U -> S
        wsync();
        reg_product<N, 0, 1>(T1, s, L.i, T);  // rhoH = (U rho) U^dagger
        const double di = Dg[L.kk * N + L.i];
#pragma unroll
        for (int c = 0; c < C; c++) {
            const int j = L.i ^ c;
            double gam;
            if (a.w < 0) {
                gam = (L.i != j) ? 1.0 : 0.0;
            } else {
                const double dE = (di - Dg[L.kk * N + j]) / a.w;
                gam = 1 - exp(-dE * dE);
            }
            T[c] = cscale(T[c], -gam * a.F2);
        }
        reg_product<N, 0, 0>(T, s, L.i, T1);  // Q = rhoH' U
        wsync();
        slot_store_rows<N>(s, L.i, T1);  // Q -> S
        wsync();
        reg_product<N, 1, 0>(U, s, L.i, T);  // U^dagger Q
#pragma unroll
        for (int c = 0; c < C; c++) dr[c] = cadd(dr[c], T[c]);
    }
    // d rho/dt through LDS for coalescing
    wsync();
    slot_store_rows<N>(s, L.i, dr);
    wsync();
    store_state<N, 1, Lay>(a.dxdt, S, n, a.P, a.nk, blk);
    if (a.err_partial) {
        const double v = seg_max<64>(stage_err<N, 1, Lay>(a, S, blk));
        if (threadIdx.x == 0) a.err_partial[blk] = v;
    }
}

// the stage for the KPW k-points of logical block blk of one instance.
// PAIR: for n <= 4, H_E is formed together with H_0 and held in registers
// (fewer memory trips; more registers, which a caller through a call
// boundary may not want)
template <int N, int LR, bool PAIR = true>
__device__ __forceinline__ void rhs_stage_body(const StageArgs &a, const long blk)
{
    using G = Geo<N, LR>;
    constexpr int C = G::C;
    if constexpr (N == 16) {
        rhs_stage_body16<LR>(a, blk);
        return;
    } else if constexpr (LR == 1) {
        rhs_stage_body_r1<N, PAIR>(a, blk);
        return;
    } else {
    __shared__ cd S0[G::KPW * G::NN];
    __shared__ cd S1[G::KPW * G::NN];
    __shared__ double Dg[G::KPW * N];
    __shared__ cd Ja[G::KPW * N], Jb[G::KPW * N];
    __shared__ int Fa[G::KPW * N];

    const Lane L = lane_id<N, LR>(a.nk, blk);
    const int n = a.n, nn = n * n;
    cd *s0 = S0 + L.kk * G::NN, *s1 = S1 + L.kk * G::NN;

    if (blk == 0 && threadIdx.x < 8) {
        const int t = threadIdx.x;
        double v = t < a.dim ? -a.E[t] : 0.0;
        a.dxdt[2 * t] = v;
        a.dxdt[2 * t + 1] = 0.0;
        if (a.xout) {
            double re = a.in.x[2 * t], im = a.in.x[2 * t + 1];
#pragma unroll
            for (int j = 0; j < 6; j++)
                if (j < a.in.nt) {
                    re += a.in.c[j] * a.in.k[j][2 * t];
                    im += a.in.c[j] * a.in.k[j][2 * t + 1];
                }
            a.xout[2 * t] = re;
            a.xout[2 * t + 1] = im;
        }
    }

    cd dr[C];
    cd U[C];  // the eigenvectors, kept in registers from the Jacobi to the dephasing
    // few bands: H_E formed with H_0, kept in registers until it is needed
    constexpr int HE = N <= 4 ? G::KPW * G::NN / 64 : 1;
    cd he[HE];
    bool he_ready = false;
    auto form_h0 = [&]() {  // H_0 -> S1
        if constexpr (N <= 4 && PAIR) he_ready = phase_mats_pair<N, LR>(S1, a, blk, he);
        if (!he_ready) phase_mats<N, LR>(S1, a, 0, blk);
    };
    if (a.dephase) {
        cd A[C];
        if (a.warm) {
            // warm start: A = U' H0 U'^dagger with U' the previous eigenvectors
            load_mats<N, LR>(S0, a.ucache, nn, 0, n, a.nk, blk);
            form_h0();
            wsync();
            load_u<N, LR>(s0, L, U);
            // padded rows and columns of U' are the identity
            u_pad_identity<N, LR>(L, n, U);
            cd T[C];
            lds_product<N, LR, 0, 0>(s0, s1, L, T);  // U' H0
            wsync();
            store_xrow<N, LR>(s0, L, T);
            store_u<N, LR>(s1, L, U);
            wsync();
            lds_product<N, LR, 0, 1>(s0, s1, L, A);  // (U' H0) U'^dagger
            wave_jacobi<N, LR>(A, U, L, a.sweeps, Dg, Ja, Jb, Fa, false);
        } else {
            form_h0();
            wsync();
            load_xrow<N, LR>(s1, L, A);
            wave_jacobi<N, LR>(A, U, L, a.sweeps, Dg, Ja, Jb, Fa);
        }
        store_u<N, LR>(s1, L, U);  // U -> S1
        wsync();
        // the eigenvectors go to the per-k cache: warm start of the next call
        store_ucache<N, LR>(a, S1, blk);
    }
    load_rho<N, LR>(S0, a, blk);  // rho -> S0
    wsync();
    if (a.xout) store_state<N, LR>(a.xout, S0, n, a.P, a.nk, blk);
    // coherent part: -i (H_E rho - rho H_E)
    if (he_ready) store_entries<N, LR>(S1, he);  // H_E -> S1
    else phase_mats<N, LR>(S1, a, 1, blk);
    wsync();
    {
        cd t1[C];
        lds_product<N, LR, 0, 0>(s1, s0, L, t1);
#pragma unroll
        for (int c = 0; c < C; c++) dr[c] = mk(t1[c].y, -t1[c].x);
        lds_product<N, LR, 0, 0>(s0, s1, L, t1);
#pragma unroll
        for (int c = 0; c < C; c++) dr[c] = mk(dr[c].x - t1[c].y, dr[c].y + t1[c].x);
    }
    if (a.dephase) {
        // dephasing in the eigenbasis: U^dagger (D o (U rho U^dagger)) U
        wsync();
        store_u<N, LR>(s1, L, U);  // U -> S1 (padded rows and columns: the identity)
        wsync();
        cd T[C];
        lds_product<N, LR, 0, 0>(s1, s0, L, T);  // T = U rho
        wsync();
        store_xrow<N, LR>(s0, L, T);
        wsync();
        lds_product<N, LR, 0, 1>(s0, s1, L, T);  // rhoH = T U^dagger
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
        store_xrow<N, LR>(s0, L, T);
        wsync();
        lds_product<N, LR, 0, 0>(s0, s1, L, T);  // S = rhoH' U
        wsync();
        store_xrow<N, LR>(s0, L, T);
        wsync();
        lds_product<N, LR, 1, 0>(s1, s0, L, T);  // U^dagger S
#pragma unroll
        for (int c = 0; c < C; c++) dr[c] = cadd(dr[c], T[c]);
    }

    // write d rho/dt (with zero pads) through LDS for coalescing
    wsync();
    store_xrow<N, LR>(s0, L, dr);
    wsync();
    store_state<N, LR>(a.dxdt, S0, n, a.P, a.nk, blk);
    if (a.err_partial) {
        const double v = seg_max<64>(stage_err<N, LR>(a, S0, blk));
        if (threadIdx.x == 0) a.err_partial[blk] = v;
    }
    }
}

// minimum waves per SIMD the register allocation must allow: the 16-band
// kernel at 4 lanes per row fits four (at most 128 VGPRs; 8.9 KB of LDS per
// wavefront) when the allocator is told so, instead of two with part of the
// matrix-core operands in AGPRs; at 2 lanes per row a bound would only buy
// spills to scratch
template <int N, int LR>
constexpr int stage_min_waves() { return (N == 16 && LR == 4) ? 4 : 1; }

template <int N, int LR>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(stage_min_waves<N, LR>())))
void rhs_stage_kernel(StageArgs a)
{
    rhs_stage_body<N, LR>(a, xcd_block(blockIdx.x, gridDim.x));
}

// Several instances in one launch: entry j owns the logical blocks
// [block0[j], block0[j+1]); its arguments are read once into registers.
template <int N, int LR>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(stage_min_waves<N, LR>())))
void rhs_stage_batched_kernel(const StageArgs *args, const long *block0, int m)
{
    const long gb = xcd_block(blockIdx.x, gridDim.x);
    const int j = find_entry_wave(block0, m, gb);
    // read in place (a local copy of the 256-byte struct would live in scratch)
    rhs_stage_body<N, LR>(args[j], gb - block0[j]);
}

// initial Fermi-Dirac state rho = U^dagger f(eps) U, and/or eigenvalues and
// eigenvectors (code/src/PropagatorCPU.cpp:31-70, code/src/main/runSBE.cpp:130-155)
template <int N, int LR>
__global__ __launch_bounds__(64) void init_state_kernel(InitArgs a)
{
    using G = Geo<N, LR>;
    constexpr int C = G::C;
    __shared__ cd S0[G::KPW * G::NN];
    __shared__ cd S1[G::KPW * G::NN];
    __shared__ double Dg[G::KPW * N];
    __shared__ cd Ja[G::KPW * N], Jb[G::KPW * N];
    __shared__ int Fa[G::KPW * N];
    const long blk = blockIdx.x;
    const Lane L = lane_id<N, LR>(a.nk, blk);
    const int n = a.n, nn = n * n;
    cd *s0 = S0 + L.kk * G::NN, *s1 = S1 + L.kk * G::NN;
    load_mats<N, LR>(S1, a.mesh, (long)a.ops * nn, 0, n, a.nk, blk);
    wsync();
    cd A[C], U[C];
    load_xrow<N, LR>(s1, L, A);
    wave_jacobi<N, LR>(A, U, L, a.sweeps, Dg, Ja, Jb, Fa);
    wsync();
    u_rows<N, LR>(s1, L, U);
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
    store_xrow<N, LR>(s1, L, U);
    const double f = 1.0 / (1.0 + exp((Dg[L.kk * N + L.i] - a.mu) / a.kT));
    cd FU[C];
#pragma unroll
    for (int c = 0; c < C; c++) FU[c] = (L.i < n) ? cscale(U[c], f) : mk(0, 0);
    store_xrow<N, LR>(s0, L, FU);
    wsync();
    // rho[a][b] = sum_i conj(U[i][a]) f_i U[i][b]
    cd R[C];
    lds_product<N, LR, 1, 0>(s1, s0, L, R);
    wsync();
    store_xrow<N, LR>(s0, L, R);
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
        if (p < nn) v = S0[kk * G::NN + sw<N>(p / n, p % n)];
        double *o = a.state + 2 * (8 + k * P + p);
        o[0] = v.x;
        o[1] = v.y;
    }
}

// H0(k + A) of the observer's partial phase sum (op 0 of M-wide rows) for the
// wave's k-points into LDS natural layout (last-axis contraction per entry)
template <int N, int LR>
__device__ __forceinline__ void obs_h0_mats(cd *lds, const OccArgs &a, long blk)
{
    using G = Geo<N, LR>;
    const int n = a.n;
    for (int e = threadIdx.x; e < G::KPW * G::NN; e += 64) {
        const int kk = e / G::NN, idx = e % G::NN;
        const int r = idx / N, c = idx % N;
        const long k = blk * G::KPW + kk;
        cd v = mk(0, 0);
        if (r < n && c < n && k < a.nk) v = phase_last(a.ph, k, a.M, r * n + c);
        lds[kk * G::NN + sw<N>(r, c)] = v;
    }
}

// Hamiltonian and Wannier occupations per k-point (ObserverCPU::operator(),
// code/src/ObserverCPU.cpp:211-245): U H0(k + A) U^dagger = diag(e) by the
// wavefront Jacobi (from the identity), occupations (U rho U^dagger)_bb,
// bands in ascending energy (ties by index), smeared over energy with
// weights exp(-((e_s - e_q)/smear)^2) summed over q in ascending order.
template <int N, int LR>
__global__ __launch_bounds__(64) void occupation_kernel(OccArgs a)
{
    using G = Geo<N, LR>;
    constexpr int C = G::C;
    __shared__ cd S0[G::KPW * G::NN];
    __shared__ cd S1[G::KPW * G::NN];
    __shared__ double Dg[G::KPW * N];
    __shared__ cd Ja[G::KPW * N], Jb[G::KPW * N];
    __shared__ int Fa[G::KPW * N];
    const long blk = blockIdx.x;
    const Lane L = lane_id<N, LR>(a.nk, blk);
    const int n = a.n;
    cd *s0 = S0 + L.kk * G::NN, *s1 = S1 + L.kk * G::NN;
    obs_h0_mats<N, LR>(S1, a, blk);
    load_mats<N, LR>(S0, a.x, a.P, 8, n, a.nk, blk);
    wsync();
    cd A[C], U[C];
    load_xrow<N, LR>(s1, L, A);
    wave_jacobi<N, LR>(A, U, L, a.sweeps, Dg, Ja, Jb, Fa);
    wsync();
    u_rows<N, LR>(s1, L, U);
    store_xrow<N, LR>(s1, L, U);
    wsync();
    cd T[C];
    lds_product<N, LR, 0, 0>(s1, s0, L, T);  // T = U rho
    // (U rho U^dagger)_ii = sum_j T_ij conj(U_ij) over the row's lanes
    double o = 0;
#pragma unroll
    for (int c = 0; c < C; c++) o += T[c].x * U[c].x + T[c].y * U[c].y;
    o = seg_sum<LR>(o);
    const double ei = Dg[L.kk * N + L.i];
    int rank = 0;
    for (int j = 0; j < n; j++) {
        const double ej = Dg[L.kk * N + j];
        rank += (ej < ei) || (ej == ei && j < L.i);
    }
    wsync();
    if (L.h == 0 && L.i < n) Ja[L.kk * N + rank] = mk(ei, o);
    wsync();
    if (L.h == 0 && L.i < n && L.valid) {
        const int sidx = L.i;  // the band this lane reports (ascending energy)
        const cd es = Ja[L.kk * N + sidx];
        double v = es.y;
        if (a.smear > 0) {
            double ws = 0, acc = 0;
            for (int q = 0; q < n; q++) {
                const cd eq = Ja[L.kk * N + q];
                const double dE = (es.x - eq.x) / a.smear;
                const double w = exp(-dE * dE);
                ws += w;
                acc += eq.y * w;
            }
            v = acc / ws;
        }
        double *row = a.X + L.k * (long)a.Q;
        row[3 + sidx] = v;
        row[3 + n + sidx] = s0[sw<N>(sidx, sidx)].x;
    }
}


template <int N, int LR>
__global__ __launch_bounds__(64) void rhs_stage_set_kernel(
    StageSet ss)
{
    const long blk = xcd_block(blockIdx.x, gridDim.x);
    for (int q = 0; q < ss.m; q++) {
        rhs_stage_body<N, LR>(ss.st[q], blk);
        // the next stage loads, from other lanes, what this one stored
        __threadfence_block();
        wsync();
    }
    if (ss.ticket) {
        // thread 0 stored this block's error maximum; its ticket releases it.
        // The block that takes the last ticket reduces all of them, resets the
        // counter for the next launch and stores the norm's bit pattern with a
        // system-scope store (no block waits for another)
        unsigned t = 0;
        if (threadIdx.x == 0) t = __hip_atomic_fetch_add(ss.ticket, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        t = __shfl(t, 0);
        if (t + 1 == gridDim.x) {
            __threadfence();
            const double *part = ss.st[ss.m - 1].err_partial;
            double v = 0;
            for (unsigned i = threadIdx.x; i < gridDim.x; i += 64)
                v = fmax(v, __hip_atomic_load(part + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
            v = seg_max<64>(v);
            if (threadIdx.x == 0) {
                __hip_atomic_store(ss.ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(ss.err_out, (unsigned long long)__double_as_longlong(v), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_SYSTEM);
            }
        }
    }
}

}  // namespace sbe
