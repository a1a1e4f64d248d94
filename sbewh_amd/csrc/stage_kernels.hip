// Launchers of the fused SBE right-hand-side stage kernels (device code in
// stage_device.hpp).
#include "stage_device.hpp"

namespace sbe {

template <int N, int LR>
static hipError_t launch_occ_NL(const OccArgs &a, hipStream_t s)
{
    using G = Geo<N, LR>;
    long blocks = (a.nk + G::KPW - 1) / G::KPW;
    if (blocks <= 0) return hipSuccess;
    hipLaunchKernelGGL((occupation_kernel<N, LR>), dim3((unsigned)blocks), dim3(64), 0, s, a);
    return hipGetLastError();
}

template <int N, int LR>
static hipError_t launch_rhs_NL(const StageArgs &a, hipStream_t s)
{
    using G = Geo<N, LR>;
    long blocks = (a.nk + G::KPW - 1) / G::KPW;
    if (blocks <= 0) return hipSuccess;
    hipLaunchKernelGGL((rhs_stage_kernel<N, LR>), dim3((unsigned)blocks), dim3(64), 0, s, a);
    return hipGetLastError();
}

template <int N, int LR>
static hipError_t launch_rhs_set_NL(const StageSet &ss, hipStream_t s)
{
    using G = Geo<N, LR>;
    if (ss.m <= 0) return hipSuccess;
    if (ss.m > SET_MAX) return hipErrorInvalidValue;
    for (int q = 1; q < ss.m; q++)
        if (ss.st[q].nk != ss.st[0].nk) return hipErrorInvalidValue;
    if (ss.ticket && (!ss.err_out || !ss.st[ss.m - 1].err_partial)) return hipErrorInvalidValue;
    long blocks = (ss.st[0].nk + G::KPW - 1) / G::KPW;
    if (blocks <= 0) return hipSuccess;
    hipLaunchKernelGGL((rhs_stage_set_kernel<N, LR>), dim3((unsigned)blocks), dim3(64), 0, s, ss);
    return hipGetLastError();
}

template <int N, int LR>
static hipError_t launch_rhs_batched_NL(const StageArgs *args, const long *block0, int m, long blocks, hipStream_t s)
{
    if (blocks <= 0 || m <= 0) return hipSuccess;
    hipLaunchKernelGGL((rhs_stage_batched_kernel<N, LR>), dim3((unsigned)blocks), dim3(64), 0, s, args, block0, m);
    return hipGetLastError();
}

// This file is compiled twice.  SBE_STAGE_PART 2 holds the one-lane-per-row
// stage kernels (N = 2, 4, 8 at LR = 1), built with the default machine
// scheduler: the ILP-oriented one takes them past the register count of
// three waves per SIMD.  SBE_STAGE_PART 1 holds everything else and the
// dispatchers, built with the ILP-oriented scheduler.  0 (the default)
// builds both in one object.
#ifndef SBE_STAGE_PART
#define SBE_STAGE_PART 0
#endif

hipError_t launch_rhs_stage_r1(int N, const StageArgs &a, hipStream_t s);
hipError_t launch_rhs_stage_set_r1(int N, const StageSet &ss, hipStream_t s);
hipError_t launch_rhs_stage_batched_r1(int N, const StageArgs *args, const long *block0, int m, long blocks,
                                       hipStream_t s);

#if SBE_STAGE_PART != 1
hipError_t launch_rhs_stage_r1(int N, const StageArgs &a, hipStream_t s)
{
    switch (N) {
    case 2: return launch_rhs_NL<2, 1>(a, s);
    case 4: return launch_rhs_NL<4, 1>(a, s);
    case 8: return launch_rhs_NL<8, 1>(a, s);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_rhs_stage_set_r1(int N, const StageSet &ss, hipStream_t s)
{
    switch (N) {
    case 2: return launch_rhs_set_NL<2, 1>(ss, s);
    case 4: return launch_rhs_set_NL<4, 1>(ss, s);
    case 8: return launch_rhs_set_NL<8, 1>(ss, s);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_rhs_stage_batched_r1(int N, const StageArgs *args, const long *block0, int m, long blocks,
                                       hipStream_t s)
{
    switch (N) {
    case 2: return launch_rhs_batched_NL<2, 1>(args, block0, m, blocks, s);
    case 4: return launch_rhs_batched_NL<4, 1>(args, block0, m, blocks, s);
    case 8: return launch_rhs_batched_NL<8, 1>(args, block0, m, blocks, s);
    }
    return hipErrorInvalidValue;
}
#endif

#if SBE_STAGE_PART != 2
int stage_k_per_block(int N, int LR) { return 64 / (N * LR); }

template <int N, int LR>
static hipError_t launch_init_NL(const InitArgs &a, hipStream_t s)
{
    using G = Geo<N, LR>;
    long blocks = (a.nk + G::KPW - 1) / G::KPW;
    if (blocks <= 0) return hipSuccess;
    hipLaunchKernelGGL((init_state_kernel<N, LR>), dim3((unsigned)blocks), dim3(64), 0, s, a);
    return hipGetLastError();
}

// (N, LR) variants compiled for the tuner; index order is the tuning order
static const int kVariants[][2] = {{2, 1}, {2, 2}, {4, 1}, {4, 2}, {4, 4}, {8, 1}, {8, 2}, {8, 4}, {16, 2}, {16, 4}};

int stage_variants(int N, int *lr_out, int cap)
{
    int k = 0;
    for (auto &v : kVariants)
        if (v[0] == N && k < cap) lr_out[k++] = v[1];
    return k;
}

hipError_t launch_rhs_stage(int N, int LR, const StageArgs &a, hipStream_t s)
{
    switch (N * 100 + LR) {
    case 201: return launch_rhs_stage_r1(2, a, s);
    case 202: return launch_rhs_NL<2, 2>(a, s);
    case 401: return launch_rhs_stage_r1(4, a, s);
    case 402: return launch_rhs_NL<4, 2>(a, s);
    case 404: return launch_rhs_NL<4, 4>(a, s);
    case 801: return launch_rhs_stage_r1(8, a, s);
    case 802: return launch_rhs_NL<8, 2>(a, s);
    case 804: return launch_rhs_NL<8, 4>(a, s);
    case 1602: return launch_rhs_NL<16, 2>(a, s);
    case 1604: return launch_rhs_NL<16, 4>(a, s);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_rhs_stage_set(int N, int LR, const StageSet &ss, hipStream_t s)
{
    switch (N * 100 + LR) {
    case 201: return launch_rhs_stage_set_r1(2, ss, s);
    case 202: return launch_rhs_set_NL<2, 2>(ss, s);
    case 401: return launch_rhs_stage_set_r1(4, ss, s);
    case 402: return launch_rhs_set_NL<4, 2>(ss, s);
    case 404: return launch_rhs_set_NL<4, 4>(ss, s);
    case 801: return launch_rhs_stage_set_r1(8, ss, s);
    case 802: return launch_rhs_set_NL<8, 2>(ss, s);
    case 804: return launch_rhs_set_NL<8, 4>(ss, s);
    case 1602: return launch_rhs_set_NL<16, 2>(ss, s);
    case 1604: return launch_rhs_set_NL<16, 4>(ss, s);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_rhs_stage_batched(int N, int LR, const StageArgs *args, const long *block0, int m, long blocks,
                                    hipStream_t s)
{
    switch (N * 100 + LR) {
    case 201: return launch_rhs_stage_batched_r1(2, args, block0, m, blocks, s);
    case 202: return launch_rhs_batched_NL<2, 2>(args, block0, m, blocks, s);
    case 401: return launch_rhs_stage_batched_r1(4, args, block0, m, blocks, s);
    case 402: return launch_rhs_batched_NL<4, 2>(args, block0, m, blocks, s);
    case 404: return launch_rhs_batched_NL<4, 4>(args, block0, m, blocks, s);
    case 801: return launch_rhs_stage_batched_r1(8, args, block0, m, blocks, s);
    case 802: return launch_rhs_batched_NL<8, 2>(args, block0, m, blocks, s);
    case 804: return launch_rhs_batched_NL<8, 4>(args, block0, m, blocks, s);
    case 1602: return launch_rhs_batched_NL<16, 2>(args, block0, m, blocks, s);
    case 1604: return launch_rhs_batched_NL<16, 4>(args, block0, m, blocks, s);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_occupations(int N, int LR, const OccArgs &a, hipStream_t s)
{
    if (a.Q < 3 + 2 * a.n) return hipErrorInvalidValue;
    switch (N * 100 + LR) {
    case 201: return launch_occ_NL<2, 1>(a, s);
    case 202: return launch_occ_NL<2, 2>(a, s);
    case 401: return launch_occ_NL<4, 1>(a, s);
    case 402: return launch_occ_NL<4, 2>(a, s);
    case 404: return launch_occ_NL<4, 4>(a, s);
    case 801: return launch_occ_NL<8, 1>(a, s);
    case 802: return launch_occ_NL<8, 2>(a, s);
    case 804: return launch_occ_NL<8, 4>(a, s);
    case 1602: return launch_occ_NL<16, 2>(a, s);
    case 1604: return launch_occ_NL<16, 4>(a, s);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_init_state(int N, int LR, const InitArgs &a, hipStream_t s)
{
    switch (N * 100 + LR) {
    case 201: return launch_init_NL<2, 1>(a, s);
    case 202: return launch_init_NL<2, 2>(a, s);
    case 401: return launch_init_NL<4, 1>(a, s);
    case 402: return launch_init_NL<4, 2>(a, s);
    case 404: return launch_init_NL<4, 4>(a, s);
    case 801: return launch_init_NL<8, 1>(a, s);
    case 802: return launch_init_NL<8, 2>(a, s);
    case 804: return launch_init_NL<8, 4>(a, s);
    case 1602: return launch_init_NL<16, 2>(a, s);
    case 1604: return launch_init_NL<16, 4>(a, s);
    }
    return hipErrorInvalidValue;
}

#endif

}  // namespace sbe
