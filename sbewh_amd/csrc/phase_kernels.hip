// Wannier -> Bloch phase sum X(k) = sum_R exp(i 2 pi k.R) X(R) on CDNA4, FP64.
//
// The reference evaluates it as pre-phase + batched FFT + post-phase
// (code/src/WhFourierTransform.cpp:208-315; GPU: code/src/CuWhFourierTransform.cu:96-173).
// Here the k + A shift is folded into per-dimension twiddles
//   tw_d[i][r] = exp(i 2 pi (i/Nk_d + off_d) (r + rmin_d)),  off = frac(kGlobalShift + A),
// and the sum over the dense R box is contracted one axis at a time, i.e. a
// chain of small complex GEMMs [Nk_d x W_d] x [W_d x (rest)].  The operator
// merge H_E(R) = H(R) + sum_d E_d r_d(R) (code/src/WhFourierTransform.cpp:63-66)
// happens while the first contraction reads the model.
//
// Two forms:
//  * launch_phase_sum: every axis (R_2 -> i_2, R_1 -> i_1, R_0 -> i_0), the
//    last contraction scattering into the state's k order through the
//    grid -> k map (observer operators, initial state, Fermi level);
//  * launch_phase_partial: axes 0 .. dim-2 only (R_0 -> i_0 first), leaving
//    the axis along which consecutive k indices run to the stage kernel,
//    which contracts it while loading H(k) (no H(k) mesh in HBM).
#include "device_common.hpp"
#include "kernels.hpp"
#include "phase_device.hpp"

namespace sbe {

__global__ void twiddle_kernel(PhaseGeom g, StageInput in, int has_a, int use_shift, double *tw, ModelMerge mm)
{
    prep_body(g, in, has_a, use_shift, tw, mm, (long)blockIdx.x * blockDim.x + threadIdx.x);
}

// entry blockIdx.y of the set
__global__ void prep_set_kernel(PrepSet ps)
{
    const int j = blockIdx.y;
    const TwiddleEntry &q = ps.e[j];
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const int ntw = twiddle_count(q.g);
    if (t < ntw) twiddle_body(q.g, q.in, q.has_a, q.use_shift, q.tw, (int)t, &ps.a[j]);
    else merge_body(q.mm, t - ntw);
}

// one instance per blockIdx.y
__global__ void twiddle_batched_kernel(const TwiddleEntry *e)
{
    const TwiddleEntry &q = e[blockIdx.y];
    prep_body(q.g, q.in, q.has_a, q.use_shift, q.tw, q.mm, (long)blockIdx.x * blockDim.x + threadIdx.x);
}

constexpr int IB = 8;

__device__ __forceinline__ cd model_value(const ContractArgs &a, long b, int m)
{
    int op = m / a.nn, e = m - op * a.nn;
    const double *h = a.Hb + 2 * (b * a.nn + e);
    cd v = mk(h[0], h[1]);
    if (a.in_mode == 1) {
        if (op == 1) {
            double E[3] = {a.E0, a.E1, a.E2};
            for (int d = 0; d < a.dim; d++) {
                const double *p = a.Db + 2 * (((long)d * a.nbox + b) * a.nn + e);
                v.x += E[d] * p[0];
                v.y += E[d] * p[1];
            }
        }
        return v;
    }
    // observer operators: H0, D_0..D_{dim-1}, dH_0..dH_{dim-1}
    if (op == 0) return v;
    if (op <= a.dim) {
        const double *p = a.Db + 2 * (((long)(op - 1) * a.nbox + b) * a.nn + e);
        return mk(p[0], p[1]);
    }
    const double *p = a.dHb + 2 * (((long)(op - 1 - a.dim) * a.nbox + b) * a.nn + e);
    return mk(p[0], p[1]);
}

__device__ __forceinline__ void contract_body(const ContractArgs &a, long bx)
{
    const int nIB = (a.Nkd + IB - 1) / IB;
    long tid = bx * 256 + threadIdx.x;
    long total = a.outer * nIB * a.inner * a.M;
    if (tid >= total) return;
    int m = (int)(tid % a.M);
    long rest = tid / a.M;
    long in = rest % a.inner;
    rest /= a.inner;
    int ib = (int)(rest % nIB);
    long o = rest / nIB;
    int i0 = ib * IB;
    cd acc[IB];
#pragma unroll
    for (int q = 0; q < IB; q++) acc[q] = mk(0, 0);
    for (int r = 0; r < a.W; r++) {
        long src = (o * a.W + r) * a.inner + in;
        cd x;
        if (a.in_mode == 0) {
            const double *p = a.in + 2 * (src * a.M + m);
            x = mk(p[0], p[1]);
        } else {
            x = model_value(a, src, m);
        }
#pragma unroll
        for (int q = 0; q < IB; q++) {
            int i = i0 + q;
            if (i < a.Nkd) {
                const double *t = a.tw + 2 * ((long)i * a.W + r);
                acc[q] = cfma(mk(t[0], t[1]), x, acc[q]);
            }
        }
    }
#pragma unroll
    for (int q = 0; q < IB; q++) {
        int i = i0 + q;
        if (i >= a.Nkd) break;
        long dst = (o * a.Nkd + i) * a.inner + in;
        if (a.out_mode == 1) dst = a.kmap[dst];
        double *p = a.out + 2 * (dst * a.M + m);
        p[0] = acc[q].x;
        p[1] = acc[q].y;
    }
}

// Tiled form of the same contraction: a block owns 256 consecutive columns
// (o, in, m) and an i-tile of TI rows; each thread keeps its W inputs in
// registers, the TI x W twiddle tile sits in LDS (broadcast reads), outputs
// are written coalesced along m.
constexpr int TI = 16;
constexpr int WMAX = 16;

__device__ __forceinline__ void contract_tiled_body(const ContractArgs &a, long bx, int by)
{
    __shared__ cd tws[TI * WMAX];
    const long cols_per_o = a.inner * a.M;
    const long ncol = a.outer * cols_per_o;
    const long col = bx * 256 + threadIdx.x;
    const int i0 = by * a.ti;
    const int W = a.W;
    for (int e = threadIdx.x; e < a.ti * W; e += 256) {
        int ii = e / W, r = e % W;
        int i = i0 + ii;
        cd t = mk(0, 0);
        if (i < a.Nkd) {
            const double *p = a.tw + 2 * ((long)i * W + r);
            t = mk(p[0], p[1]);
        }
        tws[ii * WMAX + r] = t;
    }
    __syncthreads();
    if (col >= ncol) return;
    const long o = col / cols_per_o;
    const long rem = col - o * cols_per_o;
    const long in = rem / a.M;
    const int m = (int)(rem - in * a.M);
    cd x[WMAX];
#pragma unroll
    for (int r = 0; r < WMAX; r++) {
        if (r < W) {
            long src = (o * W + r) * a.inner + in;
            if (a.in_mode == 0) {
                const double *p = a.in + 2 * (src * a.M + m);
                x[r] = mk(p[0], p[1]);
            } else {
                x[r] = model_value(a, src, m);
            }
        } else {
            x[r] = mk(0, 0);
        }
    }
    const int iend = min(a.ti, a.Nkd - i0);
    for (int ii = 0; ii < iend; ii++) {
        cd acc = mk(0, 0);
#pragma unroll
        for (int r = 0; r < WMAX; r++)
            if (r < W) acc = cfma(tws[ii * WMAX + r], x[r], acc);
        long dst = (o * a.Nkd + (i0 + ii)) * a.inner + in;
        if (a.out_mode == 1) dst = a.kmap[dst];
        reinterpret_cast<double2 *>(a.out)[dst * a.M + m] = make_double2(acc.x, acc.y);
    }
}

// operator-merged model box Y[b][m] = op_m(R_b) (dim == 1, observer operators)
__device__ __forceinline__ void box_body(const ContractArgs &a, long bx)
{
    long t = bx * 256 + threadIdx.x;
    if (t >= a.nbox * a.M) return;
    cd v = model_value(a, t / a.M, (int)(t % a.M));
    reinterpret_cast<double2 *>(a.out)[t] = make_double2(v.x, v.y);
}

__device__ __forceinline__ void stage_body(const ContractStage &st, long lb)
{
    if (st.kind == 0) contract_tiled_body(st.a, lb % st.blocks_x, (int)(lb / st.blocks_x));
    else if (st.kind == 1) contract_body(st.a, lb);
    else box_body(st.a, lb);
}

__global__ __launch_bounds__(256) void contract_stage_kernel(ContractStage st)
{
    stage_body(st, (long)blockIdx.y * st.blocks_x + blockIdx.x);
}

// entry j of the set owns the blocks [block0[j], block0[j+1])
__global__ __launch_bounds__(256) void contract_set_kernel(ContractSet cs)
{
    const long gb = blockIdx.x;
    int j = 0;
    while (j + 1 < cs.m && cs.block0[j + 1] <= gb) j++;
    stage_body(cs.st[j], gb - cs.block0[j]);
}

// entry j owns the blocks [block0[j], block0[j+1])
__global__ __launch_bounds__(256) void contract_batched_kernel(const ContractStage *st, const long *block0, int m)
{
    const long gb = blockIdx.x;
    const int j = find_entry_wave(block0, m, gb);
    const ContractStage q = st[j];
    stage_body(q, gb - block0[j]);
}

hipError_t launch_phase_prep(const PhaseGeom &g, const StageInput *in, int use_shift, double *tw,
                             const ModelMerge &mm, hipStream_t s)
{
    long tot = (long)g.Nk[0] * g.W[0] + (long)g.Nk[1] * g.W[1] + (long)g.Nk[2] * g.W[2];
    if (mm.out) tot += mm.nbox * 2L * mm.nn;
    StageInput v{};
    if (in) v = *in;
    hipLaunchKernelGGL(twiddle_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, g, v, in ? 1 : 0,
                       use_shift, tw, mm);
    return hipGetLastError();
}

hipError_t launch_twiddles(const PhaseGeom &g, const StageInput *in, int use_shift, double *tw, hipStream_t s)
{
    return launch_phase_prep(g, in, use_shift, tw, ModelMerge{}, s);
}

long twiddle_offset(const PhaseGeom &g, int d)
{
    long o = 0;
    for (int e = 0; e < d; e++) o += (long)g.Nk[e] * g.W[e];
    return o;
}

// grid of one contraction stage: the tiled kernel when the R extent fits its
// register tile, the generic one otherwise
static ContractStage plan_contract(const ContractArgs &a0, long min_blocks = 2048)
{
    ContractStage st{};
    st.a = a0;
    ContractArgs &a = st.a;
    if (a.W <= WMAX) {
        long ncol = a.outer * a.inner * a.M;
        long cb = (ncol + 255) / 256;
        // largest i-tile (inputs held in registers are reused ti times) that
        // still gives >= min_blocks blocks (2048 fill 256 CUs)
        int ti = TI;
        while (ti > 1 && cb * ((a.Nkd + ti - 1) / ti) < min_blocks) ti >>= 1;
        a.ti = ti;
        st.kind = 0;
        st.blocks_x = cb;
        st.blocks_y = (a.Nkd + ti - 1) / ti;
    } else {
        long total = a.outer * ((a.Nkd + IB - 1) / IB) * a.inner * a.M;
        st.kind = 1;
        st.blocks_x = (total + 255) / 256;
        st.blocks_y = 1;
    }
    return st;
}

static void launch_stage(const ContractStage &st, hipStream_t s)
{
    hipLaunchKernelGGL(contract_stage_kernel, dim3((unsigned)st.blocks_x, (unsigned)st.blocks_y), dim3(256), 0, s,
                       st);
}

static void launch_contract(const ContractArgs &a, hipStream_t s) { launch_stage(plan_contract(a), s); }

// stage d runs iff it changes the data: W_d > 1 or Nk_d > 1 (stage 0 always runs)
static bool stage_runs(const PhaseGeom &g, int d) { return d == 0 || g.W[d] > 1 || g.Nk[d] > 1; }

hipError_t launch_phase_sum(const PhaseGeom &g, const PhaseBuffers &b, int ops, double E0, double E1,
                            double E2, double *mesh, hipStream_t s)
{
    const int nn = g.n * g.n;
    ContractArgs a{};
    a.M = ops * nn;
    a.nn = nn;
    a.dim = g.dim;
    a.nbox = (long)g.W[0] * g.W[1] * g.W[2];
    a.Hb = b.Hb;
    a.Db = b.Db;
    a.dHb = b.dHb;
    a.E0 = g.dim > 0 ? E0 : 0;
    a.E1 = g.dim > 1 ? E1 : 0;
    a.E2 = g.dim > 2 ? E2 : 0;
    a.kmap = b.kmap;
    const double *src = nullptr;
    bool first = true;
    double *bufs[2] = {b.y2, b.y1};
    int nb = 0;
    long twoff[3] = {0, (long)g.Nk[0] * g.W[0], (long)g.Nk[0] * g.W[0] + (long)g.Nk[1] * g.W[1]};
    for (int d = 2; d >= 0; d--) {
        if (!stage_runs(g, d)) continue;
        long outer = 1, inner = 1;
        for (int e = 0; e < d; e++) outer *= g.W[e];
        for (int e = d + 1; e < 3; e++) inner *= g.Nk[e];
        a.in = src;
        a.in_mode = first ? (ops == 2 ? 1 : 2) : 0;
        a.W = g.W[d];
        a.Nkd = g.Nk[d];
        a.outer = outer;
        a.inner = inner;
        a.tw = b.tw + 2 * twoff[d];
        bool last = (d == 0);
        a.out_mode = last ? 1 : 0;
        a.out = last ? mesh : bufs[nb];
        launch_contract(a, s);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        src = a.out;
        nb ^= 1;
        first = false;
    }
    return hipSuccess;
}

long phase_partial_size(const PhaseGeom &g, int ops)
{
    const long M = (long)ops * g.n * g.n;
    const int L = g.dim - 1;
    long v = M;
    for (int e = 0; e < 3; e++) v *= (e < L) ? g.Nk[e] : g.W[e];
    return v;
}

long phase_partial_scratch(const PhaseGeom &g, int ops)
{
    if (g.dim < 3) return 1;
    return (long)g.Nk[0] * g.W[1] * g.W[2] * ops * g.n * g.n;
}

int phase_partial_plan(const PhaseGeom &g, const PhaseBuffers &b, int ops, double E0, double E1, double E2,
                       double *Y, const double *merged, ContractStage *st, long min_blocks)
{
    const int nn = g.n * g.n;
    ContractArgs a{};
    a.M = ops * nn;
    a.nn = nn;
    a.dim = g.dim;
    a.nbox = (long)g.W[0] * g.W[1] * g.W[2];
    a.Hb = b.Hb;
    a.Db = b.Db;
    a.dHb = b.dHb;
    a.E0 = g.dim > 0 ? E0 : 0;
    a.E1 = g.dim > 1 ? E1 : 0;
    a.E2 = g.dim > 2 ? E2 : 0;
    const int model_mode = ops == 2 ? 1 : 2;
    a.in_mode = model_mode;
    const int L = g.dim - 1;
    if (L == 0 && merged) return 0;  // the merged box is Y
    if (L == 0) {
        a.out = Y;
        st[0] = ContractStage{};
        st[0].a = a;
        st[0].kind = 2;
        st[0].blocks_x = (a.nbox * a.M + 255) / 256;
        st[0].blocks_y = 1;
        return 1;
    }
    const double *src = nullptr;
    for (int d = 0; d < L; d++) {
        long outer = 1, inner = 1;
        for (int e = 0; e < d; e++) outer *= g.Nk[e];
        for (int e = d + 1; e < 3; e++) inner *= g.W[e];
        a.in = (d == 0 && merged) ? merged : src;
        a.in_mode = (d == 0 && !merged) ? model_mode : 0;
        a.W = g.W[d];
        a.Nkd = g.Nk[d];
        a.outer = outer;
        a.inner = inner;
        a.tw = b.tw + 2 * twiddle_offset(g, d);
        a.out_mode = 0;
        a.out = (d == L - 1) ? Y : b.y1;
        st[d] = plan_contract(a, min_blocks);
        src = a.out;
    }
    return L;
}

hipError_t launch_phase_partial(const PhaseGeom &g, const PhaseBuffers &b, int ops, double E0, double E1,
                                double E2, double *Y, const double *merged, hipStream_t s)
{
    ContractStage st[2];
    int ns = phase_partial_plan(g, b, ops, E0, E1, E2, Y, merged, st);
    for (int i = 0; i < ns; i++) {
        launch_stage(st[i], s);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_prep_set(const PrepSet &ps, hipStream_t s)
{
    if (ps.m <= 0) return hipSuccess;
    if (ps.m > SET_MAX) return hipErrorInvalidValue;
    hipLaunchKernelGGL(prep_set_kernel, dim3((unsigned)((ps.max_count + 255) / 256), (unsigned)ps.m), dim3(256), 0, s,
                       ps);
    return hipGetLastError();
}

hipError_t launch_contract_set(const ContractSet &cs, hipStream_t s)
{
    if (cs.m <= 0) return hipSuccess;
    if (cs.m > SET_MAX) return hipErrorInvalidValue;
    const long blocks = cs.block0[cs.m];
    if (blocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(contract_set_kernel, dim3((unsigned)blocks), dim3(256), 0, s, cs);
    return hipGetLastError();
}

hipError_t launch_twiddles_batched(const TwiddleEntry *d_entries, int m, int max_count, hipStream_t s)
{
    if (m <= 0) return hipSuccess;
    hipLaunchKernelGGL(twiddle_batched_kernel, dim3((unsigned)((max_count + 255) / 256), (unsigned)m), dim3(256), 0,
                       s, d_entries);
    return hipGetLastError();
}

hipError_t launch_contract_batched(const ContractStage *d_stages, const long *d_block0, int m, long blocks,
                                   hipStream_t s)
{
    if (m <= 0 || blocks <= 0) return hipSuccess;
    hipLaunchKernelGGL(contract_batched_kernel, dim3((unsigned)blocks), dim3(256), 0, s, d_stages, d_block0, m);
    return hipGetLastError();
}

}  // namespace sbe
