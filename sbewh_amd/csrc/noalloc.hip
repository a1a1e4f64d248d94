// sbe_observe_device_noalloc: the current j(t) of a device state for callers
// that cannot allocate device memory (code under HIP graph capture, tight
// serving loops).  Same kernels and summation order as sbe_observe; the
// per-k partial currents live in a file-scope device array instead of the
// context's buffer, and the result goes to a caller-owned device buffer.
#include <mutex>

#include "context.hpp"

namespace {

// per-k partial currents: 3 doubles per k-point
constexpr long SCRATCH_DOUBLES = 3L * 262144;
static __device__ __align__(256) double scratch[SCRATCH_DOUBLES];

// device address of `scratch` on each GPU, taken once per device with that
// device current (the array's host-side name is not a device address)
constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
double *g_scratch[MAX_DEVICES] = {};

double *scratch_on(int device)
{
    if (device < 0 || device >= MAX_DEVICES) fail(SBE_ERR_ARG, "device %d out of range", device);
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[device]) {
        int cur = 0;
        HIPCHK(hipGetDevice(&cur));
        if (cur != device) HIPCHK(hipSetDevice(device));
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device) HIPCHK(hipSetDevice(cur));
        HIPCHK(e);
        if (!p) fail(SBE_ERR_HIP, "no device address for the observer scratch array");
        g_scratch[device] = static_cast<double *>(p);
    }
    return g_scratch[device];
}

}  // namespace

extern "C" int sbe_observe_device_noalloc(sbe_ctx *ctx, const double *d_state, double *d_j, void *stream)
{
    if (!ctx || !d_state || !d_j) return SBE_ERR_ARG;
    return guarded(ctx, [&] {
        sbe_ctx &c = *ctx;
        if (3L * c.nk > SCRATCH_DOUBLES)
            fail(SBE_ERR_ARG, "sbe_observe_device_noalloc supports at most %ld k-points (context has %ld)",
                 SCRATCH_DOUBLES / 3, c.nk);
        double *part = scratch_on(c.device);
        hipStream_t s = stream ? static_cast<hipStream_t>(stream) : c.stream;
        StageInput in{};
        in.x = d_state;
        HIPCHK(launch_twiddles(c.geom, &in, 1, c.tw, s));
        HIPCHK(launch_phase_partial(c.geom, c.pbufs(), c.obs_ops, 0, 0, 0, c.Yobs, nullptr, s));
        ObsArgs o{};
        o.x = d_state;
        o.ph = c.last(c.Yobs);
        o.nk = c.nk;
        o.n = c.n;
        o.P = c.P;
        o.dim = c.dim;
        o.partial = part;
        o.j_out = d_j;
        o.pstride = 3;
        HIPCHK(launch_observe_j(o, s));
    });
}
