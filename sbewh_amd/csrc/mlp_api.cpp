// Surrogate MLP ensemble (include/sbewh_surrogate.h): parameters, training
// (Adam, one captured HIP graph per training step replayed for every batch)
// and inference, all members at once, on the kernels of mlp_kernels.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/sbewh_surrogate.h"
#include "mlp.hpp"

using namespace sbe;

namespace {

thread_local std::string g_mlp_err;

struct MlpFail {
    int code;
    std::string msg;
};

[[noreturn]] void mfail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    throw MlpFail{code, buf};
}

#define MCHK(x)                                                                                   \
    do {                                                                                          \
        hipError_t e_ = (x);                                                                      \
        if (e_ != hipSuccess) mfail(SBE_MLP_ERR_HIP, "HIP error %d (%s) in %s", (int)e_,          \
                                    hipGetErrorString(e_), #x);                                   \
    } while (0)

template <class T>
T *dmalloc(size_t count)
{
    void *p = nullptr;
    MCHK(hipMalloc(&p, (count ? count : 1) * sizeof(T)));
    return static_cast<T *>(p);
}

}  // namespace

struct sbe_mlp {
    std::string err;
    int E = 0, L = 0;             // members, weight layers
    std::vector<int> d;           // widths d[0..L]
    std::vector<long> woff, boff; // per layer offsets inside one member's parameters
    long P = 0;                   // parameters per member
    hipStream_t stream = nullptr;
    float *p = nullptr, *m = nullptr, *v = nullptr, *g = nullptr;  // [E][P]
    bf16 *pb = nullptr;
    int *counter = nullptr;
    // batch buffers (capacity cap rows per member)
    int cap = 0, wmax = 0;
    bf16 *X = nullptr, *dY = nullptr, *dZa = nullptr, *dZb = nullptr;
    std::vector<float *> Z;       // [L-1] pre-activations
    std::vector<bf16 *> H;        // [L-1] activations
    float *Y = nullptr, *T = nullptr;

    ~sbe_mlp()
    {
        free_batch();
        for (void *q : {(void *)p, (void *)m, (void *)v, (void *)g, (void *)pb, (void *)counter})
            if (q) (void)hipFree(q);
        if (stream) (void)hipStreamDestroy(stream);
    }
    void free_batch()
    {
        for (void *q : {(void *)X, (void *)dY, (void *)dZa, (void *)dZb, (void *)Y, (void *)T})
            if (q) (void)hipFree(q);
        for (auto *q : Z) (void)hipFree(q);
        for (auto *q : H) (void)hipFree(q);
        X = dY = dZa = dZb = nullptr;
        Y = T = nullptr;
        Z.clear();
        H.clear();
        cap = 0;
    }
    void ensure_batch(int rows)
    {
        if (rows <= cap) return;
        free_batch();
        cap = rows;
        X = dmalloc<bf16>((size_t)E * rows * d[0]);
        for (int l = 0; l + 1 < L; l++) {
            Z.push_back(dmalloc<float>((size_t)E * rows * d[l + 1]));
            H.push_back(dmalloc<bf16>((size_t)E * rows * d[l + 1]));
        }
        Y = dmalloc<float>((size_t)E * rows * d[L]);
        T = dmalloc<float>((size_t)E * rows * d[L]);
        dY = dmalloc<bf16>((size_t)E * rows * d[L]);
        dZa = dmalloc<bf16>((size_t)E * rows * wmax);
        dZb = dmalloc<bf16>((size_t)E * rows * wmax);
    }

    // forward of all members on the `rows` rows staged in X; Y receives the output
    void forward(int rows)
    {
        for (int l = 0; l < L; l++) {
            GemmArgs a{};
            a.M = rows;
            a.N = d[l + 1];
            a.K = d[l];
            a.A = l == 0 ? X : H[l - 1];
            a.lda = d[l];
            a.strideA = (long)cap * d[l];
            a.B = pb + woff[l];
            a.ldb = d[l];
            a.strideB = P;
            a.bias = p + boff[l];
            a.strideBias = P;
            if (l + 1 < L) {
                a.mode = 1;
                a.C = Z[l];
                a.ldc = d[l + 1];
                a.strideC = (long)cap * d[l + 1];
                a.H = H[l];
                a.ldh = d[l + 1];
                a.strideH = (long)cap * d[l + 1];
            } else {
                a.mode = 2;
                a.C = Y;
                a.ldc = d[L];
                a.strideC = (long)cap * d[L];
            }
            MCHK(launch_gemm(a, 0, 1, E, stream));
        }
    }

    // backward from dY (bf16) into g (all members)
    void backward(int rows)
    {
        const bf16 *dcur = dY;
        int wcur = d[L];
        bf16 *bufs[2] = {dZa, dZb};
        int nb = 0;
        for (int l = L - 1; l >= 0; l--) {
            // dW_l = dcur^T h_{l-1}  (N_out x N_in, sum over the batch)
            GemmArgs w{};
            w.M = d[l + 1];
            w.N = d[l];
            w.K = rows;
            w.A = dcur;
            w.lda = wcur;
            w.strideA = (long)cap * (l == L - 1 ? d[L] : wmax);
            w.B = l == 0 ? X : H[l - 1];
            w.ldb = d[l];
            w.strideB = (long)cap * d[l];
            w.mode = 0;
            w.C = g + woff[l];
            w.ldc = d[l];
            w.strideC = P;
            MCHK(launch_gemm(w, 1, 0, E, stream));
            MCHK(launch_colsum(dcur, rows, d[l + 1], w.strideA, E, g + boff[l], P, stream));
            if (l == 0) break;
            // dZ_{l-1} = (dcur W_l) * silu'(Z_{l-1})
            GemmArgs h{};
            h.M = rows;
            h.N = d[l];
            h.K = d[l + 1];
            h.A = dcur;
            h.lda = wcur;
            h.strideA = w.strideA;
            h.B = pb + woff[l];
            h.ldb = d[l];
            h.strideB = P;
            h.mode = 3;
            h.H = bufs[nb];
            h.ldh = d[l];
            h.strideH = (long)cap * wmax;
            h.Zin = Z[l - 1];
            h.ldz = d[l];
            h.strideZ = (long)cap * d[l];
            MCHK(launch_gemm(h, 0, 0, E, stream));
            dcur = bufs[nb];
            wcur = d[l];
            nb ^= 1;
        }
    }
};

namespace {

template <class F>
int mguard(sbe_mlp *n, F &&f)
{
    try {
        f();
        return SBE_MLP_OK;
    } catch (const MlpFail &e) {
        if (n) n->err = e.msg;
        else g_mlp_err = e.msg;
        return e.code;
    } catch (const std::exception &e) {
        if (n) n->err = e.what();
        else g_mlp_err = e.what();
        return SBE_MLP_ERR_ARG;
    }
}

}  // namespace

extern "C" {

int sbe_mlp_create(int members, int layers, const int *dims, unsigned long long seed, sbe_mlp **out)
{
    if (!out) return SBE_MLP_ERR_ARG;
    *out = nullptr;
    sbe_mlp *n = new sbe_mlp();
    int rc = mguard(nullptr, [&] {
        if (members < 1 || members > 64 || layers < 1 || layers > 8 || !dims)
            mfail(SBE_MLP_ERR_ARG, "invalid ensemble shape");
        n->E = members;
        n->L = layers;
        n->d.assign(dims, dims + layers + 1);
        for (int w : n->d)
            if (w < 1 || w > 4096) mfail(SBE_MLP_ERR_ARG, "layer widths must be in [1, 4096]");
        long off = 0;
        for (int l = 0; l < layers; l++) {
            n->woff.push_back(off);
            off += (long)n->d[l + 1] * n->d[l];
            n->boff.push_back(off);
            off += n->d[l + 1];
            if (l + 1 < layers) n->wmax = std::max(n->wmax, n->d[l + 1]);
        }
        n->wmax = std::max(n->wmax, 1);
        n->P = off;
        MCHK(hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking));
        const size_t tot = (size_t)members * n->P;
        n->p = dmalloc<float>(tot);
        n->m = dmalloc<float>(tot);
        n->v = dmalloc<float>(tot);
        n->g = dmalloc<float>(tot);
        n->pb = dmalloc<bf16>(tot);
        n->counter = dmalloc<int>(1);
        // He-uniform weights, zero biases, one generator stream per member
        std::vector<float> h(tot, 0.0f);
        for (int e = 0; e < members; e++) {
            std::mt19937_64 rng(seed + 7919ULL * (unsigned long long)e);
            for (int l = 0; l < layers; l++) {
                const float lim = std::sqrt(6.0f / (float)n->d[l]) * (l + 1 < layers ? 1.0f : 0.5f);
                std::uniform_real_distribution<float> U(-lim, lim);
                for (long i = 0; i < (long)n->d[l + 1] * n->d[l]; i++) h[e * n->P + n->woff[l] + i] = U(rng);
            }
        }
        MCHK(hipMemcpy(n->p, h.data(), tot * sizeof(float), hipMemcpyHostToDevice));
        MCHK(hipMemset(n->m, 0, tot * sizeof(float)));
        MCHK(hipMemset(n->v, 0, tot * sizeof(float)));
        MCHK(hipMemset(n->counter, 0, sizeof(int)));
        MCHK(launch_to_bf16(n->p, n->pb, (long)tot, n->stream));
        MCHK(hipStreamSynchronize(n->stream));
    });
    if (rc) {
        delete n;
        return rc;
    }
    *out = n;
    return SBE_MLP_OK;
}

void sbe_mlp_destroy(sbe_mlp *mlp) { delete mlp; }

const char *sbe_mlp_last_error(const sbe_mlp *mlp) { return mlp ? mlp->err.c_str() : g_mlp_err.c_str(); }

long sbe_mlp_param_count(const sbe_mlp *mlp) { return mlp ? mlp->P : SBE_MLP_ERR_ARG; }

int sbe_mlp_get_params(const sbe_mlp *mlp, int member, float *params)
{
    if (!mlp || !params || member < 0 || member >= mlp->E) return SBE_MLP_ERR_ARG;
    sbe_mlp *n = const_cast<sbe_mlp *>(mlp);
    return mguard(n, [&] {
        MCHK(hipMemcpyAsync(params, n->p + (size_t)member * n->P, n->P * sizeof(float), hipMemcpyDeviceToHost,
                            n->stream));
        MCHK(hipStreamSynchronize(n->stream));
    });
}

int sbe_mlp_set_params(sbe_mlp *mlp, int member, const float *params)
{
    if (!mlp || !params || member < 0 || member >= mlp->E) return SBE_MLP_ERR_ARG;
    return mguard(mlp, [&] {
        sbe_mlp &n = *mlp;
        MCHK(hipMemcpyAsync(n.p + (size_t)member * n.P, params, n.P * sizeof(float), hipMemcpyHostToDevice,
                            n.stream));
        MCHK(launch_to_bf16(n.p + (size_t)member * n.P, n.pb + (size_t)member * n.P, n.P, n.stream));
        MCHK(hipStreamSynchronize(n.stream));
    });
}

int sbe_mlp_forward(sbe_mlp *mlp, int count, const float *x, float *y)
{
    if (!mlp || !x || !y || count < 1) return SBE_MLP_ERR_ARG;
    return mguard(mlp, [&] {
        sbe_mlp &n = *mlp;
        const int chunk = 4096;
        n.ensure_batch(std::min(count, chunk));
        const int din = n.d[0], dout = n.d[n.L];
        std::vector<int> perm;
        int *dperm = dmalloc<int>((size_t)n.E * n.cap);
        int *dctr = dmalloc<int>(1);
        float *dx = dmalloc<float>((size_t)count * din);
        std::vector<float> yb((size_t)n.E * n.cap * dout);
        try {
            MCHK(hipMemcpyAsync(dx, x, (size_t)count * din * sizeof(float), hipMemcpyHostToDevice, n.stream));
            MCHK(hipMemsetAsync(dctr, 0, sizeof(int), n.stream));
            for (int s0 = 0; s0 < count; s0 += n.cap) {
                const int rows = std::min(n.cap, count - s0);
                perm.assign((size_t)n.E * n.cap, 0);
                for (int e = 0; e < n.E; e++)
                    for (int b = 0; b < rows; b++) perm[(size_t)e * n.cap + b] = s0 + b;
                MCHK(hipMemcpyAsync(dperm, perm.data(), perm.size() * sizeof(int), hipMemcpyHostToDevice, n.stream));
                MCHK(launch_gather(dx, dx, din, 0, dperm, n.cap, dctr, 1, rows, n.cap, n.E, n.X, n.T, n.stream));
                n.forward(rows);
                MCHK(hipMemcpyAsync(yb.data(), n.Y, yb.size() * sizeof(float), hipMemcpyDeviceToHost, n.stream));
                MCHK(hipStreamSynchronize(n.stream));
                for (int e = 0; e < n.E; e++)
                    std::memcpy(y + ((size_t)e * count + s0) * dout, yb.data() + (size_t)e * n.cap * dout,
                                (size_t)rows * dout * sizeof(float));
            }
        } catch (...) {
            (void)hipFree(dperm);
            (void)hipFree(dctr);
            (void)hipFree(dx);
            throw;
        }
        MCHK(hipFree(dperm));
        MCHK(hipFree(dctr));
        MCHK(hipFree(dx));
    });
}

int sbe_mlp_gradient(sbe_mlp *mlp, int count, const float *x, const float *t, float *loss, float *grad)
{
    if (!mlp || !x || !t || !loss || !grad || count < 1 || count > 4096) return SBE_MLP_ERR_ARG;
    return mguard(mlp, [&] {
        sbe_mlp &n = *mlp;
        n.ensure_batch(count);
        const int din = n.d[0], dout = n.d[n.L];
        std::vector<int> perm((size_t)n.E * count);
        for (int e = 0; e < n.E; e++)
            for (int b = 0; b < count; b++) perm[(size_t)e * count + b] = b;
        int *dperm = dmalloc<int>(perm.size());
        int *dctr = dmalloc<int>(1);
        float *dx = dmalloc<float>((size_t)count * din), *dt = dmalloc<float>((size_t)count * dout);
        float *dl = dmalloc<float>(n.E);
        try {
            MCHK(hipMemcpyAsync(dperm, perm.data(), perm.size() * sizeof(int), hipMemcpyHostToDevice, n.stream));
            MCHK(hipMemcpyAsync(dx, x, (size_t)count * din * sizeof(float), hipMemcpyHostToDevice, n.stream));
            MCHK(hipMemcpyAsync(dt, t, (size_t)count * dout * sizeof(float), hipMemcpyHostToDevice, n.stream));
            MCHK(hipMemsetAsync(dctr, 0, sizeof(int), n.stream));
            MCHK(launch_gather(dx, dt, din, dout, dperm, count, dctr, 1, count, n.cap, n.E, n.X, n.T, n.stream));
            n.forward(count);
            MCHK(launch_loss_grad(n.Y, n.T, count, n.cap, dout, n.E, n.dY, dl, 1, dctr, n.stream));
            n.backward(count);
            MCHK(hipMemcpyAsync(loss, dl, n.E * sizeof(float), hipMemcpyDeviceToHost, n.stream));
            MCHK(hipMemcpyAsync(grad, n.g, (size_t)n.E * n.P * sizeof(float), hipMemcpyDeviceToHost, n.stream));
            MCHK(hipStreamSynchronize(n.stream));
        } catch (...) {
            for (void *q : {(void *)dperm, (void *)dctr, (void *)dx, (void *)dt, (void *)dl}) (void)hipFree(q);
            throw;
        }
        for (void *q : {(void *)dperm, (void *)dctr, (void *)dx, (void *)dt, (void *)dl}) MCHK(hipFree(q));
    });
}

int sbe_mlp_train(sbe_mlp *mlp, int count, const float *x, const float *t, int batch, int epochs, float lr,
                  unsigned long long seed, float *loss_out, double *seconds)
{
    if (!mlp || !x || !t || count < 1 || batch < 1 || batch > count || batch > 4096 || epochs < 1 || !(lr > 0))
        return SBE_MLP_ERR_ARG;
    return mguard(mlp, [&] {
        sbe_mlp &n = *mlp;
        auto tic = std::chrono::steady_clock::now();
        n.ensure_batch(batch);
        const int din = n.d[0], dout = n.d[n.L];
        const int spe = count / batch;   // steps per epoch (the remainder of a shuffle is dropped)
        const long total = (long)spe * epochs;
        int *dperm = dmalloc<int>((size_t)n.E * count);
        float *dx = dmalloc<float>((size_t)count * din), *dt = dmalloc<float>((size_t)count * dout);
        float *dl = dmalloc<float>((size_t)n.E * total);
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        auto cleanup = [&] {
            if (exec) (void)hipGraphExecDestroy(exec);
            if (graph) (void)hipGraphDestroy(graph);
            for (void *q : {(void *)dperm, (void *)dx, (void *)dt, (void *)dl}) (void)hipFree(q);
        };
        try {
            MCHK(hipMemcpyAsync(dx, x, (size_t)count * din * sizeof(float), hipMemcpyHostToDevice, n.stream));
            MCHK(hipMemcpyAsync(dt, t, (size_t)count * dout * sizeof(float), hipMemcpyHostToDevice, n.stream));
            MCHK(hipMemsetAsync(n.counter, 0, sizeof(int), n.stream));
            MCHK(hipMemsetAsync(n.m, 0, (size_t)n.E * n.P * sizeof(float), n.stream));
            MCHK(hipMemsetAsync(n.v, 0, (size_t)n.E * n.P * sizeof(float), n.stream));
            MCHK(hipStreamSynchronize(n.stream));
            // one training step of every member, captured once and replayed
            MCHK(hipStreamBeginCapture(n.stream, hipStreamCaptureModeThreadLocal));
            MCHK(launch_gather(dx, dt, din, dout, dperm, count, n.counter, spe, batch, n.cap, n.E, n.X, n.T, n.stream));
            n.forward(batch);
            MCHK(launch_loss_grad(n.Y, n.T, batch, n.cap, dout, n.E, n.dY, dl, (int)total, n.counter, n.stream));
            n.backward(batch);
            MCHK(launch_adam(n.p, n.pb, n.m, n.v, n.g, (long)n.E * n.P, lr, 0.9f, 0.999f, 1e-8f, total, n.counter,
                             n.stream));
            MCHK(launch_increment(n.counter, n.stream));
            MCHK(hipStreamEndCapture(n.stream, &graph));
            MCHK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
            std::vector<int> perm((size_t)n.E * count);
            std::vector<std::mt19937_64> rng;
            for (int e = 0; e < n.E; e++) rng.emplace_back(seed + 104729ULL * (unsigned long long)e);
            for (int ep = 0; ep < epochs; ep++) {
                for (int e = 0; e < n.E; e++) {
                    int *q = perm.data() + (size_t)e * count;
                    for (int i = 0; i < count; i++) q[i] = i;
                    std::shuffle(q, q + count, rng[e]);
                }
                // the previous epoch's replays read the permutation: wait for them
                MCHK(hipStreamSynchronize(n.stream));
                MCHK(hipMemcpyAsync(dperm, perm.data(), perm.size() * sizeof(int), hipMemcpyHostToDevice, n.stream));
                for (int s = 0; s < spe; s++) MCHK(hipGraphLaunch(exec, n.stream));
            }
            MCHK(hipStreamSynchronize(n.stream));
            if (loss_out) {
                std::vector<float> lh((size_t)n.E * total);
                MCHK(hipMemcpy(lh.data(), dl, lh.size() * sizeof(float), hipMemcpyDeviceToHost));
                for (int ep = 0; ep < epochs; ep++)
                    for (int e = 0; e < n.E; e++) {
                        double acc = 0;
                        for (int s = 0; s < spe; s++) acc += lh[(size_t)e * total + (long)ep * spe + s];
                        loss_out[(size_t)ep * n.E + e] = (float)(acc / spe);
                    }
            }
        } catch (...) {
            cleanup();
            throw;
        }
        cleanup();
        auto toc = std::chrono::steady_clock::now();
        if (seconds) *seconds = std::chrono::duration<double>(toc - tic).count();
    });
}

}  // extern "C"
