// High-harmonic spectrum of a current j(t) and per-harmonic yields.
//
// Definitions follow the reference's post-processing scripts:
//   fadeout        scripts/util/hhg.py:9-14   (cos^2 ramp over the last N = int(rel * n)
//                                              samples r[-N-1:-1], then r[-1] = 0)
//   fftSpectrum    scripts/util/hhg.py:16-25  (zero-pad to N = int(max(n, 2 pi/(dt d_omega))),
//                                              S = |FFT|^2 / N * omega^(2 power), omega_k = 2 pi k/(N dt), k < N/2)
//   calcSpectrum   scripts/util/common.py:187-201 (omega_min < omega < omega_max mask)
//   showHHG        scripts/showExpectationValues.py:36-58 (j / max|j|, per-direction
//                                              spectra summed, directions below 1e-10 max dropped)
//   extractProfiles scripts/ellipticityDependency.py:77-83 (sum over [h w0 - w, h w0 + w))
//   directOmega    scripts/util/hhg.py:27-38  (omega grid limited by the sampling theorem)
//   directSpectrum scripts/util/hhg.py:40-42  (S = |omega^power sum_t exp(i omega t) j(t)|^2)
//   smearSpectrum  scripts/util/hhg.py:45-49  (Gaussian filter of FWHM smear_width, zero outside,
//                                              scipy.ndimage.gaussian_filter1d with truncate 4)
// The discrete Fourier transform of arbitrary length uses Bluestein's chirp-z
// algorithm over an iterative radix-2 FFT (no external FFT library).
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

#include "../../include/sbewh_hip.h"

namespace {

using cplx = std::complex<double>;

void fft_pow2(std::vector<cplx> &a, bool inverse)
{
    const size_t n = a.size();
    for (size_t i = 1, j = 0; i < n; i++) {
        size_t bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(a[i], a[j]);
    }
    for (size_t len = 2; len <= n; len <<= 1) {
        const double ang = 2 * M_PI / (double)len * (inverse ? 1 : -1);
        const size_t half = len / 2;
        // twiddles of this level computed directly (no recurrence drift)
        std::vector<cplx> w(half);
        for (size_t k = 0; k < half; k++) w[k] = cplx(std::cos(ang * (double)k), std::sin(ang * (double)k));
        for (size_t i = 0; i < n; i += len)
            for (size_t k = 0; k < half; k++) {
                cplx u = a[i + k], v = a[i + k + half] * w[k];
                a[i + k] = u + v;
                a[i + k + half] = u - v;
            }
    }
    if (inverse)
        for (auto &x : a) x /= (double)n;
}

// forward DFT X_k = sum_n x_n exp(-2 pi i n k / N), any N (Bluestein)
std::vector<cplx> dft(const std::vector<cplx> &x)
{
    const size_t N = x.size();
    if (N == 0) return {};
    if ((N & (N - 1)) == 0) {
        std::vector<cplx> a = x;
        fft_pow2(a, false);
        return a;
    }
    size_t M = 1;
    while (M < 2 * N - 1) M <<= 1;
    // chirp w_n = exp(i pi n^2 / N), n^2 reduced mod 2N for accuracy
    std::vector<cplx> w(N);
    for (size_t n = 0; n < N; n++) {
        unsigned long long q = (unsigned long long)n * n % (2ull * N);
        double ang = M_PI * (double)q / (double)N;
        w[n] = cplx(std::cos(ang), std::sin(ang));
    }
    std::vector<cplx> a(M), b(M);
    for (size_t n = 0; n < N; n++) a[n] = x[n] * std::conj(w[n]);
    b[0] = w[0];
    for (size_t n = 1; n < N; n++) b[n] = b[M - n] = w[n];
    fft_pow2(a, false);
    fft_pow2(b, false);
    for (size_t i = 0; i < M; i++) a[i] *= b[i];
    fft_pow2(a, true);
    std::vector<cplx> X(N);
    for (size_t k = 0; k < N; k++) X[k] = a[k] * std::conj(w[k]);
    return X;
}

void fadeout(std::vector<double> &r, double rel)
{
    const long n = (long)r.size();
    const long N = (long)(rel * (double)n);
    // np.linspace(0, pi/2, N, endpoint=True); r[-N-1:-1] *= cos^2
    for (long i = 0; i < N; i++) {
        double x = N == 1 ? 0.0 : (M_PI / 2) * (double)i / (double)(N - 1);
        long idx = n - N - 1 + i;
        if (idx >= 0) r[idx] *= std::cos(x) * std::cos(x);
    }
    if (n > 0) r[n - 1] = 0;
}

}  // namespace

extern "C" {

int sbe_hhg_spectrum(const double *t, const double *j, int nt, int ncomp, double d_omega, double omega_min,
                     double omega_max, double fadeout_rel, int power, double *omega_out, double *S_out,
                     int capacity, int *count)
{
    if (!t || !j || nt < 2 || ncomp < 1 || ncomp > 3 || !(d_omega > 0) || !count) return SBE_ERR_ARG;
    const double dt = t[1] - t[0];
    if (!(dt > 0)) return SBE_ERR_ARG;
    const long N = (long)std::max((double)nt, 2 * M_PI / dt / d_omega);
    const long cut = N / 2;
    // normalisation by the global maximum of |j| (showHHG)
    double gmax = 0;
    std::vector<double> cmax(ncomp, 0.0);
    for (int c = 0; c < ncomp; c++)
        for (int i = 0; i < nt; i++) cmax[c] = std::max(cmax[c], std::fabs(j[(size_t)i * ncomp + c]));
    for (double v : cmax) gmax = std::max(gmax, v);
    std::vector<double> total(cut, 0.0);
    std::vector<long> sel;
    for (long k = 0; k < cut; k++) {
        double w = 2 * M_PI / dt / (double)N * (double)k;
        if (omega_min < w && w < omega_max) sel.push_back(k);
    }
    if ((long)sel.size() > capacity) {
        *count = (int)sel.size();
        return SBE_ERR_ARG;
    }
    if (gmax > 0) {
        for (int c = 0; c < ncomp; c++) {
            if (cmax[c] < 1e-10 * gmax) continue;
            std::vector<double> r(nt);
            for (int i = 0; i < nt; i++) r[i] = j[(size_t)i * ncomp + c] / gmax;
            fadeout(r, fadeout_rel);
            std::vector<cplx> x(N, 0.0);
            for (int i = 0; i < nt; i++) x[i] = r[i];
            std::vector<cplx> X = dft(x);
            for (long k = 0; k < cut; k++) {
                double w = 2 * M_PI / dt / (double)N * (double)k;
                total[k] += std::norm(X[k]) / (double)N * std::pow(w, 2 * power);
            }
        }
    }
    for (size_t q = 0; q < sel.size(); q++) {
        long k = sel[q];
        if (omega_out) omega_out[q] = 2 * M_PI / dt / (double)N * (double)k;
        if (S_out) S_out[q] = total[k];
    }
    *count = (int)sel.size();
    return SBE_OK;
}

int sbe_direct_omega(const double *t, int nt, double omega_min, double omega_max, double d_omega, double *omega_out,
                     int capacity, int *count)
{
    if (!t || nt < 2 || !(d_omega > 0) || !count) return SBE_ERR_ARG;
    const double nyq_max = M_PI / (t[1] - t[0]), nyq_d = M_PI / (t[nt - 1] - t[0]);
    if (!(nyq_max > 0) || !(nyq_d > 0)) return SBE_ERR_ARG;
    if (nyq_max < omega_max) omega_max = nyq_max;
    if (nyq_d > d_omega) d_omega = nyq_d;
    // numpy.arange(omega_min, omega_max, d_omega)
    const double len = std::ceil((omega_max - omega_min) / d_omega);
    const long n = len > 0 ? (long)len : 0;
    *count = (int)n;
    if (n > capacity) return SBE_ERR_ARG;
    if (omega_out)
        for (long i = 0; i < n; i++) omega_out[i] = omega_min + (double)i * d_omega;
    return SBE_OK;
}

int sbe_direct_spectrum(const double *t, int nt, const double *omega, int nw, const double *j, int power,
                        double *S_out)
{
    if (!t || nt < 1 || !omega || nw < 0 || !j || !S_out) return SBE_ERR_ARG;
    for (int q = 0; q < nw; q++) {
        cplx acc = 0;
        for (int i = 0; i < nt; i++) acc += std::exp(cplx(0, omega[q] * t[i])) * j[i];
        S_out[q] = std::norm(std::pow(omega[q], power) * acc);
    }
    return SBE_OK;
}

int sbe_hhg_spectrum_direct(const double *t, const double *j, int nt, int ncomp, double d_omega, double omega_min,
                            double omega_max, int power, double *omega_out, double *S_out, int capacity, int *count)
{
    if (!t || !j || nt < 2 || ncomp < 1 || ncomp > 3 || !count) return SBE_ERR_ARG;
    int nw = 0;
    int rc = sbe_direct_omega(t, nt, omega_min, omega_max, d_omega, nullptr, 0, &nw);
    if (rc && nw == 0) return rc;
    *count = nw;
    if (nw > capacity) return SBE_ERR_ARG;
    std::vector<double> om(nw);
    int c2 = 0;
    sbe_direct_omega(t, nt, omega_min, omega_max, d_omega, om.data(), nw, &c2);
    double gmax = 0;
    std::vector<double> cmax(ncomp, 0.0);
    for (int c = 0; c < ncomp; c++)
        for (int i = 0; i < nt; i++) cmax[c] = std::max(cmax[c], std::fabs(j[(size_t)i * ncomp + c]));
    for (double v : cmax) gmax = std::max(gmax, v);
    std::vector<double> total(nw, 0.0), S(nw), r(nt);
    if (gmax > 0)
        for (int c = 0; c < ncomp; c++) {
            if (cmax[c] < 1e-10 * gmax) continue;
            for (int i = 0; i < nt; i++) r[i] = j[(size_t)i * ncomp + c] / gmax;
            sbe_direct_spectrum(t, nt, om.data(), nw, r.data(), power, S.data());
            for (int q = 0; q < nw; q++) total[q] += S[q];
        }
    for (int q = 0; q < nw; q++) {
        if (omega_out) omega_out[q] = om[q];
        if (S_out) S_out[q] = total[q];
    }
    return SBE_OK;
}

int sbe_smear_spectrum(const double *omega, const double *S, int count, double smear_width, double *out)
{
    if (!omega || !S || !out || count < 0) return SBE_ERR_ARG;
    if (smear_width < 0 || count < 2) {
        std::memmove(out, S, sizeof(double) * (size_t)count);
        return SBE_OK;
    }
    const double sigma = smear_width / (omega[1] - omega[0]) / (2 * std::sqrt(2 * std::log(2.0)));
    if (!(sigma > 0)) return SBE_ERR_ARG;
    const long rad = (long)(4.0 * sigma + 0.5);
    std::vector<double> w(2 * rad + 1);
    double sum = 0;
    for (long k = -rad; k <= rad; k++) {
        w[k + rad] = std::exp(-0.5 * (double)k * (double)k / (sigma * sigma));
        sum += w[k + rad];
    }
    for (auto &x : w) x /= sum;
    std::vector<double> res(count, 0.0);
    for (long i = 0; i < count; i++) {
        double acc = 0;
        for (long k = -rad; k <= rad; k++) {
            const long p = i + k;
            if (p >= 0 && p < count) acc += w[k + rad] * S[p];
        }
        res[i] = acc;
    }
    std::memcpy(out, res.data(), sizeof(double) * (size_t)count);
    return SBE_OK;
}

int sbe_harmonic_yields(const double *omega, const double *S, int count, double omega0, double width, int hmax,
                        double *yields)
{
    if (!omega || !S || count < 0 || !(omega0 > 0) || hmax < 1 || !yields) return SBE_ERR_ARG;
    for (int h = 1; h <= hmax; h++) {
        const double lo = h * omega0 - width * omega0, hi = h * omega0 + width * omega0;
        // searchsorted (left) bounds: omega in [lo, hi)
        const double *b = std::lower_bound(omega, omega + count, lo);
        const double *e = std::lower_bound(omega, omega + count, hi);
        double acc = 0;
        for (const double *p = b; p < e; p++) acc += S[p - omega];
        yields[h - 1] = acc;
    }
    return SBE_OK;
}

}  // extern "C"
