// Resident integrator: Boost.Odeint's integrate_const with a controlled
// dopri5 stepper (code/src/main/runSBE.cpp:170-183) for a small 1-D single
// instance, in one cooperative launch.
//
// Each block owns the k-points the stage kernel gives its index and runs the
// six stages of every attempt on them (rhs_stage_body, as the fused attempt
// kernel does).  It also forms the phase data its k-points read: the rows of
// the six twiddle sets (tw(i, r) = e^{i 2 pi i (r + rmin) / N} e^{i 2 pi off_q (r + rmin)},
// the first factor formed once, the second per set from the stage input's A)
// and its own copy of the six merged model boxes, so the grid meets only once
// per attempt: after the last stage has left one error maximum per block.
// Every block then reads
// all maxima in the same order and runs the same step-size control on the
// same bits, so all blocks take the same decisions and reach the same
// barriers.  Per output interval: the current j(t) and the fresh stepper's
// first derivative at its start, as the host controller does.  No host round
// trip per attempt; the launch ends after the last output, after an error
// norm that is NaN, 500 failed step sizes in a row, or an attempt limit.
#include <hip/hip_cooperative_groups.h>

#include "dopri5.hpp"
#include "observer_device.hpp"
#include "phase_device.hpp"
#include "pulse.hpp"
#include "stage_device.hpp"

namespace sbe {

namespace cg = cooperative_groups;

constexpr int RES_WMAX = 32;  // largest R extent of the chain the resident integrator takes

// the grid barrier, then an agent-scope acquire: this block's later loads see
// what other blocks stored before the barrier (twiddles, boxes, maxima)
__device__ __forceinline__ void grid_sync(cg::grid_group &grid)
{
    grid.sync();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

// stage input coefficient b_qj and node c_q of the Dormand-Prince tableau
__device__ __forceinline__ double dp_coef(int q, int j)
{
    switch (q) {
    case 0: return B21;
    case 1: return j == 0 ? B31 : B32;
    case 2: return j == 0 ? B41 : (j == 1 ? B42 : B43);
    case 3: return j == 0 ? B51 : (j == 1 ? B52 : (j == 2 ? B53 : B54));
    case 4: return j == 0 ? B61 : (j == 1 ? B62 : (j == 2 ? B63 : (j == 3 ? B64 : B65)));
    default: return j == 0 ? C1 : (j == 1 ? C3 : (j == 2 ? C4 : (j == 3 ? C5 : C6)));
    }
}
__device__ __forceinline__ double dp_node(int q)
{
    return q == 0 ? A2 : (q == 1 ? A3 : (q == 2 ? A4 : (q == 3 ? A5 : 1.0)));
}
// derivative index (1..7) of input term j of stage q: k1 .. k_{q+1} for
// q < 5, then k1, k3, k4, k5, k6 for the seventh stage
__device__ __forceinline__ int dp_kidx(int q, int j) { return j == 0 ? 1 : (q < 5 ? j + 1 : j + 2); }

// E(t) into E: not inlined (the trigonometric argument reduction would
// otherwise be inlined at every call site and raise the kernel's register
// count); lane q evaluates stage q's field
__device__ __noinline__ void field_at(const SubPulse *ps, int np, double t, double *E) { pulse_E(ps, np, t, E); }

struct ResidentState {
    double *x, *xn, *k1, *k7;
    __device__ double *kd(const ResidentArgs &r, int i) const { return i == 1 ? k1 : (i == 7 ? k7 : r.k[i]); }
};

// the stage input x + dt sum_j b_qj k_j of stage q (products as the host
// controller forms them: dt * b rounded once, no contraction)
__device__ __forceinline__ StageInput stage_input(const ResidentArgs &r, const ResidentState &s, int q, double dt)
{
#pragma clang fp contract(off)
    StageInput in{};
    in.x = s.x;
    in.nt = q < 5 ? q + 1 : 5;
#pragma unroll
    for (int j = 0; j < 6; j++)
        if (j < in.nt) {
            in.k[j] = s.kd(r, dp_kidx(q, j));
            in.c[j] = dt * dp_coef(q, j);
        }
    return in;
}

__device__ __forceinline__ double *box_of(const ResidentArgs &r, long blk, int set)
{
    return r.Yb + (blk * SET_MAX + set) * r.Y_stride;
}

// off_q = (kGlobalShift + A) . a_0 / 2 pi of a stage input, A formed as the
// twiddle kernel forms it (the stage combination of the special slots, the
// derivatives this attempt forms taken from ao)
__device__ __forceinline__ double set_offset(const ResidentArgs &r, const StageInput &in, const AOverride *ao)
{
    double k_au[3];
    for (int e = 0; e < 3; e++) {
        double A = 0.0;
        if (e < r.g.dim) {
            A = in.x[2 * e];
#pragma unroll
            for (int j = 0; j < 6; j++)
                if (j < in.nt) A += in.c[j] * ((ao && ((ao->mask >> j) & 1)) ? ao->v[j][e] : in.k[j][2 * e]);
        }
        k_au[e] = r.g.kshift_au[e] + A;
    }
    const double *ad = r.g.lattice;
    return (k_au[0] * ad[0] + k_au[1] * ad[1] + k_au[2] * ad[2]) / (2 * M_PI);
}

__device__ __forceinline__ StageArgs stage_args(const ResidentArgs &r, const StageInput &in, double *dxdt,
                                                const double E[3], int set, bool warm)
{
    StageArgs a{};
    a.in = in;
    a.dxdt = dxdt;
    a.ph = PhaseLast{box_of(r, blockIdx.x, set), r.tw6 + set * r.tw_stride + 2 * r.twL_off, r.kphase, r.WL};
    a.nk = r.nk;
    a.n = r.n;
    a.P = r.P;
    a.dim = r.dim;
    a.F2 = r.F2;
    a.w = r.w;
    for (int d = 0; d < 3; d++) a.E[d] = E[d];
    a.dephase = r.F2 > 0;
    a.ucache = r.ucache;
    a.warm = warm && !r.cold_always;
    a.sweeps = a.warm ? r.sweeps : r.cold_sweeps;
    return a;
}

// Per-k current j_k (the observer's jk_body) for the k-points [k0, kend) of a
// block, 64 / (N N) k-points per pass: lane (kk, a, b) loads H0(k)_ab and
// rho_ab into LDS and adds its entry's terms; a segmented sum over the
// k-point's N N lanes gives j_k.
template <int N>
__device__ __forceinline__ void jk_pass(const ObsArgs &o, long k0, long kend, cd *Hs, cd *Rs)
{
    constexpr int NN = N * N, PK = 64 / NN;
    const int n = o.n, nn = n * n, M = (1 + 2 * o.dim) * nn;
    const int kk = threadIdx.x / NN, e = threadIdx.x % NN, a = e / N, b = e % N;
    for (long kb = k0; kb < kend; kb += PK) {
        const long k = kb + kk;
        const bool ok = k < kend && a < n && b < n;
        const int ent = ok ? a * n + b : 0;
        cd h = mk(0, 0), rh = mk(0, 0);
        if (ok) {
            h = phase_last(o.ph, k, M, ent);
            const double *rho = o.x + 2 * (8 + k * o.P);
            rh = mk(rho[2 * ent], rho[2 * ent + 1]);
        }
        Hs[kk * NN + a * N + b] = h;
        Rs[kk * NN + a * N + b] = rh;
        __syncthreads();
        double jl[3] = {0, 0, 0};
        if (ok) {
            const cd *H = Hs + kk * NN, *R = Rs + kk * NN;
            // (H rho - rho H)[b][a]
            cd t = mk(0, 0);
            for (int l = 0; l < n; l++) {
                t = cfma(H[b * N + l], R[l * N + a], t);
                t = csub(t, cmul(R[b * N + l], H[l * N + a]));
            }
            const cd rba = R[b * N + a];
            for (int d = 0; d < o.dim; d++) {
                const cd Dv = phase_last(o.ph, k, M, (1 + d) * nn + ent);
                const cd dHv = phase_last(o.ph, k, M, (1 + o.dim + d) * nn + ent);
                const cd dt = cmul(Dv, t);
                const cd hr = cmul(dHv, rba);
                jl[d] += -dt.y - hr.x;
            }
        }
        for (int d = 0; d < 3; d++) jl[d] = seg_sum<NN>(jl[d]);
        if (e == 0 && k < kend)
            for (int d = 0; d < 3; d++) o.partial[3 * k + d] = jl[d];
        __syncthreads();
    }
}

// one stage on the block's k-points, its arguments in LDS; not inlined, so
// that the stage body gets the registers the persistent loop's state would
// otherwise share with it
template <int N, int LR>
__device__ __noinline__ void stage_call(const StageArgs *sa, long blk)
{
    // H_E formed separately: the paired form's registers count against this
    // call's frame inside the resident kernel
    rhs_stage_body<N, LR, false>(*sa, blk);
}

template <int N, int LR>
__global__ __launch_bounds__(64) void resident_kernel(ResidentArgs r)
{
    using G = Geo<N, LR>;
    cg::grid_group grid = cg::this_grid();
    __shared__ double Es[SET_MAX][3];
    __shared__ AOverride Ao[SET_MAX];
    // the stage's arguments, written by lane 0: the stage body reads them from
    // LDS, so they occupy no registers across the persistent loop
    __shared__ StageArgs Sa;
    __shared__ StageInput Si[SET_MAX];
    __shared__ cd Hs[64], Rs[64];
    // e^{i 2 pi i (r + rmin) / N} of the block's k-points, e^{i 2 pi off_q (r + rmin)} of the sets
    __shared__ cd T0[G::KPW][RES_WMAX], F[SET_MAX][RES_WMAX];
    __shared__ double Off[SET_MAX];

    const long blk = blockIdx.x, nblk = gridDim.x;
    const bool dephase = r.F2 > 0;
    ResidentState s{r.x, r.xn, r.k[1], r.k[7]};
    bool ucache_valid = false;
    long steps = 0, rhs = 0, attempts = 0;
    int code = 0, parity = 0;
    const int W = r.g.W[0], rmin = r.g.rmin[0], Nk0 = r.g.Nk[0];
    const long k0 = blk * G::KPW, kend = k0 + G::KPW < r.nk ? k0 + G::KPW : r.nk;
    const long nbox2 = r.nbox * 2 * r.nn;
    const int2 *kph = reinterpret_cast<const int2 *>(r.kphase);
    for (int t = threadIdx.x; t < G::KPW * W; t += 64) {
        const int kk = t / W, rr = t % W;
        const long k = k0 + kk;
        const long i = k < r.nk ? kph[k].y / W : 0;
        long m = (i * (rr + rmin)) % Nk0;
        if (m < 0) m += Nk0;
        double sn, cs;
        sincospi(2.0 * (double)m / Nk0, &sn, &cs);
        T0[kk][rr] = mk(cs, sn);
    }
    __syncthreads();

    // the block's phase data of sets [q0, q0 + m): twiddle rows of its
    // k-points and its merged boxes (Si, Ao, Es set; ao: with overrides)
    auto block_prep = [&](int q0, int m, bool ao) {
        if (threadIdx.x < m) Off[q0 + threadIdx.x] = set_offset(r, Si[q0 + threadIdx.x], ao ? &Ao[q0 + threadIdx.x] : nullptr);
        __syncthreads();
        for (int t = threadIdx.x; t < m * W; t += 64) {
            const int q = q0 + t / W, rr = t % W;
            double sn, cs;
            sincospi(2.0 * Off[q] * (double)(rr + rmin), &sn, &cs);
            F[q][rr] = mk(cs, sn);
        }
        __syncthreads();
        for (int t = threadIdx.x; t < m * G::KPW * W; t += 64) {
            const int q = q0 + t / (G::KPW * W), rem = t % (G::KPW * W), kk = rem / W, rr = rem % W;
            const long k = k0 + kk;
            if (k < r.nk) {
                const cd v = cmul(T0[kk][rr], F[q][rr]);
                reinterpret_cast<double2 *>(r.tw6 + q * r.tw_stride)[kph[k].y + rr] = make_double2(v.x, v.y);
            }
        }
        // the merged boxes of the m sets differ only in E: each model entry
        // is read once and merged for every set (the sums of the model merge, merge_body)
        for (long t = threadIdx.x; t < nbox2; t += 64) {
            const long M = 2L * r.nn;
            const long b = t / M;
            const int mi = (int)(t - b * M), op = mi >= r.nn, e = mi - op * r.nn;
            const double2 h = reinterpret_cast<const double2 *>(r.Hb)[b * r.nn + e];
            double2 dv[3] = {};
            if (op)
                for (int d = 0; d < r.dim; d++)
                    dv[d] = reinterpret_cast<const double2 *>(r.Db)[((long)d * r.nbox + b) * r.nn + e];
            for (int q = q0; q < q0 + m; q++) {
                double re = h.x, im = h.y;
                if (op)
                    for (int d = 0; d < r.dim; d++) {
                        re += Es[q][d] * dv[d].x;
                        im += Es[q][d] * dv[d].y;
                    }
                reinterpret_cast<double2 *>(box_of(r, blk, q))[t] = make_double2(re, im);
            }
        }
        __threadfence_block();
        __syncthreads();
    };

    auto run_stage = [&](const StageInput &in, double *dxdt, const double E[3], int set, double *xout,
                         double dt) {
        if (threadIdx.x == 0) {
#pragma clang fp contract(off)
            StageArgs a = stage_args(r, in, dxdt, E, set, ucache_valid);
            if (xout) {
                // the attempt's last stage: x_new and the error terms
                a.xout = xout;
                a.err_partial = r.err_partial;
                const double de[6] = {DC1, DC3, DC4, DC5, DC6, DC7};
                for (int j = 0; j < 6; j++) a.err_e[j] = dt * de[j];
                a.err_ea = r.eps_abs;
                a.err_er = r.eps_rel;
                a.err_dt = dt;
            }
            Sa = a;
        }
        __syncthreads();
        stage_call<N, LR>(&Sa, blk);
        if (dephase) ucache_valid = true;
        // the next stage loads, from other lanes, what this one stored
        __threadfence_block();
        wsync();
    };

    // j(t) and the fresh derivative k1 = f(x, t) at an output time (set 0 holds
    // the phase data of x itself)
    auto interval_start = [&](double time, int out, bool fresh) {
        if (threadIdx.x == 0) {
            field_at(r.pulses, r.npulse, time, Es[0]);
            StageInput in0{};
            in0.x = s.x;
            Si[0] = in0;
        }
        __syncthreads();
        block_prep(0, 1, false);
        if (out < r.output_count) {
            ObsArgs o{};
            o.x = s.x;
            o.ph = PhaseLast{r.Yobs, r.tw6 + 2 * r.twL_off, r.kphase, r.WL};
            o.nk = r.nk;
            o.n = r.n;
            o.P = r.P;
            o.dim = r.dim;
            o.partial = r.jpartial;
            o.pstride = 3;
            jk_pass<N>(o, k0, kend, Hs, Rs);
        }
        if (fresh) {
            run_stage(Si[0], s.k1, Es[0], 0, nullptr, 0.0);
            rhs++;
        }
        grid_sync(grid);
        if (out < r.output_count && blk == 0) {
            // fixed summation order: lane t sums k = t, t + 64, ..., then the lanes
            double acc[3] = {0, 0, 0};
            for (long k = threadIdx.x; k < r.nk; k += 64)
                for (int d = 0; d < 3; d++) acc[d] += r.jpartial[3 * k + d];
            for (int d = 0; d < 3; d++) acc[d] = seg_sum<64>(acc[d]);
            if (threadIdx.x == 0) {
                for (int d = 0; d < 3; d++) r.j_out[3 * out + d] = acc[d] / (double)r.nk;
                r.t_out[out] = time;
            }
        }
    };

    // one Dormand-Prince attempt from (x, k1, t, dt): xn, k7 and the error norm
    auto attempt = [&](double t, double dt) -> double {
        if (threadIdx.x < SET_MAX) {
#pragma clang fp contract(off)
            field_at(r.pulses, r.npulse, t + dt * dp_node(threadIdx.x), Es[threadIdx.x]);
            Si[threadIdx.x] = stage_input(r, s, threadIdx.x, dt);
        }
        __syncthreads();
        if (threadIdx.x < SET_MAX) {
            // A slots of derivatives this attempt forms before they are read: -E at their stage
            const int q = threadIdx.x, nt = q < 5 ? q + 1 : 5;
            int mask = 0;
            for (int j = 0; j < nt; j++) {
                const int ki = dp_kidx(q, j);
                if (ki >= 2) {
                    mask |= 1 << j;
                    for (int e = 0; e < 3; e++) Ao[q].v[j][e] = e < r.dim ? -Es[ki - 2][e] : 0.0;
                }
            }
            Ao[q].mask = mask;
        }
        __syncthreads();
        block_prep(0, SET_MAX, true);
        for (int q = 0; q < SET_MAX; q++)
            run_stage(Si[q], q < 5 ? r.k[q + 2] : s.k7, Es[q], q, q == 5 ? s.xn : nullptr, dt);
        rhs += 6;
        grid_sync(grid);
        double err = 0;
        for (long b = 0; b < nblk; b++) err = fmax(err, r.err_partial[b]);
        return err;
    };

    {
#pragma clang fp contract(off)
        const double dt_out = (r.t1 - r.t0) / (r.output_count - 1);
        const double t_end = r.t1 + dt_out / 2;
        double time = r.t0, dt = dt_out;
        long step = 0;
        int out = 0;
        while (code == 0 && less_eq_with_sign(time + dt_out, t_end)) {
            double start = time, end = time + dt_out;
            const int obs_out = out;
            out++;
            int fails = 0;
            bool observed = false;
            while (code == 0 && less_with_sign(start, end)) {
                if (less_with_sign(end, start + dt)) dt = end - start;
                if (!observed) {
                    interval_start(time, obs_out, true);
                    observed = true;
                }
                for (;;) {
                    if (attempts >= r.max_attempts) {
                        code = 3;
                        break;
                    }
                    const double err = attempt(start, dt);
                    attempts++;
                    if (!(err <= 1.0)) {
                        if (isnan(err)) {
                            code = 1;
                            break;
                        }
                        const double f = 0.9 * pow(err, -1.0 / 3.0);
                        dt *= f > 0.2 ? f : 0.2;
                        if (++fails >= 500) {
                            code = 2;
                            break;
                        }
                        continue;
                    }
                    start += dt;
                    if (err < 0.5) {
                        const double e = fmax(pow(5.0, -5.0), err);
                        dt *= 0.9 * pow(e, -1.0 / 5.0);
                    }
                    break;
                }
                if (code) break;
                fails = 0;
                double *tx = s.x;
                s.x = s.xn;
                s.xn = tx;
                double *tk = s.k1;
                s.k1 = s.k7;
                s.k7 = tk;
                parity ^= 1;
                steps++;
            }
            // an interval without a step (start == end) still records its output
            if (!observed && code == 0) interval_start(time, obs_out, false);
            step++;
            time = r.t0 + (double)step * dt_out;
        }
        if (code == 0 && out < r.output_count) interval_start(time, out, false);
    }
    if (blk == 0 && threadIdx.x == 0) {
        r.status[0] = steps;
        r.status[1] = rhs;
        r.status[2] = attempts;
        r.status[3] = code;
        r.status[4] = parity;
    }
}

template <int N, int LR>
static hipError_t launch_res_NL(const ResidentArgs &r, hipStream_t s)
{
    using G = Geo<N, LR>;
    const long blocks = (r.nk + G::KPW - 1) / G::KPW;
    if (blocks <= 0) return hipErrorInvalidValue;
    ResidentArgs a = r;
    void *args[] = {&a};
    return hipLaunchCooperativeKernel(reinterpret_cast<const void *>(&resident_kernel<N, LR>), dim3((unsigned)blocks),
                                      dim3(64), args, 0, s);
}

template <int N, int LR>
static int max_blocks_NL()
{
    int per_cu = 0, dev = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void *>(&resident_kernel<N, LR>),
                                                     64, 0) != hipSuccess)
        return 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    return per_cu * cus;
}

hipError_t launch_resident(int N, int LR, const ResidentArgs &r, hipStream_t s)
{
    switch (N * 100 + LR) {
    case 201: return launch_res_NL<2, 1>(r, s);
    case 202: return launch_res_NL<2, 2>(r, s);
    case 401: return launch_res_NL<4, 1>(r, s);
    case 402: return launch_res_NL<4, 2>(r, s);
    case 404: return launch_res_NL<4, 4>(r, s);
    case 801: return launch_res_NL<8, 1>(r, s);
    case 802: return launch_res_NL<8, 2>(r, s);
    case 804: return launch_res_NL<8, 4>(r, s);
    }
    return hipErrorInvalidValue;
}

int resident_max_blocks(int N, int LR)
{
    switch (N * 100 + LR) {
    case 201: return max_blocks_NL<2, 1>();
    case 202: return max_blocks_NL<2, 2>();
    case 401: return max_blocks_NL<4, 1>();
    case 402: return max_blocks_NL<4, 2>();
    case 404: return max_blocks_NL<4, 4>();
    case 801: return max_blocks_NL<8, 1>();
    case 802: return max_blocks_NL<8, 2>();
    case 804: return max_blocks_NL<8, 4>();
    }
    return 0;
}

}  // namespace sbe
