"""numpy restatement of the reference's spectrum post-processing (test
infrastructure only): scripts/util/hhg.py:9-25 (fadeout, fftSpectrum),
scripts/util/common.py:187-201 (mask), scripts/showExpectationValues.py:36-58
(normalisation and direction sum), scripts/ellipticityDependency.py:77-83
(harmonic windows)."""
import numpy as np


def fadeout(r, rel):
    N = int(rel * r.size)
    fct = np.cos(np.linspace(0, np.pi / 2, N, endpoint=True)) ** 2
    r[-N - 1:-1] *= fct
    r[-1] = 0
    return r


def fft_spectrum(t, d_omega, j, power=1, rel_fadeout=0.1):
    N = int(max(t.size, 2 * np.pi / (t[1] - t[0]) / d_omega))
    cut = N // 2
    omega = 2 * np.pi / (t[1] - t[0]) / N * np.arange(cut)
    j = fadeout(j.copy(), rel_fadeout)
    j = np.pad(j, (0, N - j.size), "constant")
    S = np.abs(np.fft.fft(j)) ** 2 / j.size
    return omega, S[:omega.size] * omega ** (2 * power)


def hhg_spectrum(t, j, d_omega, omega_min, omega_max, rel_fadeout=0.1, power=1):
    j = np.atleast_2d(j.T).T
    maxJ = np.abs(j).max(axis=0)
    g = maxJ.max()
    tot = None
    for d in range(j.shape[1]):
        if maxJ[d] < 1e-10 * g:
            continue
        om, S = fft_spectrum(t, d_omega, j[:, d] / g, power, rel_fadeout)
        tot = S if tot is None else tot + S
    mask = (omega_min < om) & (om < omega_max)
    return om[mask], tot[mask]


def harmonic_yields(omega, S, omega0, width, hmax):
    out = np.empty(hmax)
    for h in range(1, hmax + 1):
        lo = np.searchsorted(omega, h * omega0 - width * omega0)
        hi = np.searchsorted(omega, h * omega0 + width * omega0)
        out[h - 1] = S[lo:hi].sum()
    return out


def direct_omega(t, omega_min, omega_max, d_omega):
    """scripts/util/hhg.py:27-38"""
    nyq_max = np.pi / (t[1] - t[0])
    if nyq_max < omega_max:
        omega_max = nyq_max
    nyq_d = np.pi / (t[-1] - t[0])
    if nyq_d > d_omega:
        d_omega = nyq_d
    return np.arange(omega_min, omega_max, d_omega)


def direct_spectrum(t, omega, j, power=1):
    """scripts/util/hhg.py:40-42"""
    phase = np.exp(1j * np.outer(omega, t))
    return np.abs(omega ** power * (phase @ j)) ** 2


def smear_spectrum(omega, S, smear_width):
    """scripts/util/hhg.py:45-49"""
    from scipy.ndimage import gaussian_filter1d
    if smear_width < 0:
        return S
    sigma = smear_width / (omega[1] - omega[0]) / (2 * np.sqrt(2 * np.log(2)))
    return gaussian_filter1d(S, sigma, mode="constant")


def hhg_spectrum_direct(t, j, d_omega, omega_min, omega_max, power=1):
    """showHHG normalisation and component sum over the direct spectrum"""
    j = np.atleast_2d(j.T).T
    maxJ = np.abs(j).max(axis=0)
    g = maxJ.max()
    om = direct_omega(t, omega_min, omega_max, d_omega)
    tot = np.zeros(om.size)
    for d in range(j.shape[1]):
        if maxJ[d] < 1e-10 * g:
            continue
        tot = tot + direct_spectrum(t, om, j[:, d] / g, power)
    return om, tot
