"""HHG spectrum and harmonic yields (host code in libsbewh_hip.so) against the
numpy restatement of the reference's scripts (oracle/spectrum_ref.py), which
is itself pinned by a closed form: a pure cosine at harmonic h of w0 puts its
energy into the h-th harmonic window, and the spectrum of a zero-padded
signal equals the textbook DFT sum."""
import math

import numpy as np
import pytest

from sbewh_amd import units
from sbewh_amd.spectrum import hhg_spectrum, harmonic_yields
from oracle import spectrum_ref as ref


def _signal(nt=1024, seed=0, ncomp=3):
    rng = np.random.default_rng(seed)
    t = np.linspace(-1765.0, 1765.0, nt)
    w0 = units.omega_from_lambda_nm(3200.0)
    env = np.cos(np.pi * t / (t[-1] - t[0])) ** 2
    j = np.zeros((nt, ncomp))
    for c in range(ncomp):
        for h in range(1, 30, 2):
            j[:, c] += rng.normal() * 10.0 ** (-h / 6) * env * np.cos(h * w0 * t + rng.uniform(0, 6))
    return t, j, w0


def test_reference_restatement_is_a_dft():
    t = np.linspace(0, 99, 100)
    x = np.random.default_rng(1).normal(size=100)
    dw = 2 * np.pi / (t[1] - t[0]) / 150          # N = 150 with zero padding
    om, S = ref.fft_spectrum(t, dw, x.copy(), power=0, rel_fadeout=0.0)
    xx = x.copy()
    xx[-1] = 0                                     # fadeout(rel=0) still zeroes the last sample
    n = np.arange(150)
    X = np.array([np.sum(np.r_[xx, np.zeros(50)] * np.exp(-2j * np.pi * n * k / 150)) for k in range(75)])
    assert np.allclose(S, np.abs(X) ** 2 / 150, rtol=1e-12, atol=1e-12)


def test_cosine_lands_in_its_harmonic_window():
    t = np.linspace(-3000.0, 3000.0, 4001)
    w0 = 0.02
    j = np.cos(7 * w0 * t) * np.exp(-(t / 1500.0) ** 2)
    om, S = ref.hhg_spectrum(t, j[:, None], units.from_eV(0.01), 0.0, 1.0)
    y = ref.harmonic_yields(om, S, w0, 0.5, 10)
    assert np.argmax(y) == 6
    assert y[6] > 1e3 * np.delete(y, 6).max()


@pytest.mark.parametrize("ncomp,nt", [(1, 1024), (3, 1024), (2, 801), (3, 641)])
def test_library_matches_reference_restatement(ncomp, nt):
    t, j, w0 = _signal(nt, seed=ncomp, ncomp=ncomp)
    om_r, S_r = ref.hhg_spectrum(t, j, units.from_eV(0.01), 0.0, units.from_eV(15.0))
    om, S = hhg_spectrum(t, j)
    assert om.shape == om_r.shape and np.allclose(om, om_r, rtol=1e-15, atol=0)
    above = S_r > 1e-10 * S_r.max()                # above the FFT round-off floor
    assert np.abs(np.log10(S[above]) - np.log10(S_r[above])).max() < 1e-8
    y_r = ref.harmonic_yields(om_r, S_r, w0, 0.5, 40)
    y = harmonic_yields(om, S, w0, 0.5, 40)
    ok = y_r > 0
    assert np.abs(np.log10(y[ok]) - np.log10(y_r[ok])).max() < 1e-8


def test_direct_spectrum_restatement_is_the_dft_at_fft_frequencies():
    """pin of the oracle's directSpectrum: at omega_k = 2 pi k / (N dt) the
    direct sum over a zero-padded real signal is N times the FFT spectrum
    (|sum_n e^{+i w_k t_n} j_n| = |DFT_k| for real j, any time origin)"""
    t = np.linspace(-50.0, 49.0, 100)
    x = np.random.default_rng(3).normal(size=100)
    dw = 2 * np.pi / (t[1] - t[0]) / 160                 # N = 160
    om, S_fft = ref.fft_spectrum(t, dw, x.copy(), power=1, rel_fadeout=0.0)
    xx = x.copy()
    xx[-1] = 0                                            # fadeout(rel=0) zeroes the last sample
    S_dir = ref.direct_spectrum(t, om, xx, power=1)
    assert np.allclose(S_dir, 160 * S_fft, rtol=1e-10, atol=1e-12 * S_dir.max())
    # the grid: Nyquist-limited maximum, window-limited spacing
    g = ref.direct_omega(t, 0.0, 10.0, 1e-6)
    assert np.isclose(g[1] - g[0], np.pi / (t[-1] - t[0])) and g[-1] < np.pi / (t[1] - t[0])


def test_smear_restatement_limits():
    """pin of the oracle's smearSpectrum: it conserves the sum away from the
    edges, leaves a constant unchanged in the interior, and turns a delta
    into a Gaussian of the requested FWHM"""
    om = np.arange(400) * 0.01
    S = np.zeros(400)
    S[200] = 1.0
    out = ref.smear_spectrum(om, S, 0.2)
    assert abs(out.sum() - 1.0) < 1e-12
    half = out.max() / 2
    above = om[out >= half]
    fwhm = above[-1] - above[0]
    assert abs(fwhm - 0.2) <= 0.011                       # one grid step
    flat = ref.smear_spectrum(om, np.ones(400), 0.2)
    assert np.allclose(flat[50:350], 1.0, rtol=1e-12)
    assert np.array_equal(ref.smear_spectrum(om, S, -1.0), S)


@pytest.mark.parametrize("ncomp,nt", [(1, 801), (3, 641)])
def test_library_direct_and_smear_match_restatement(ncomp, nt):
    from sbewh_amd.spectrum import direct_omega, direct_spectrum, hhg_spectrum_direct, smear_spectrum
    t, j, w0 = _signal(nt, seed=10 + ncomp, ncomp=ncomp)
    dw, wmax = units.from_eV(0.02), units.from_eV(15.0)
    assert np.array_equal(direct_omega(t, 0.0, wmax, dw), ref.direct_omega(t, 0.0, wmax, dw))
    om = ref.direct_omega(t, 0.0, wmax, dw)
    S1 = direct_spectrum(t, om, j[:, 0])
    S1r = ref.direct_spectrum(t, om, j[:, 0])
    assert np.abs(S1 - S1r).max() <= 1e-11 * S1r.max()
    om_l, S_l = hhg_spectrum_direct(t, j, 0.02, 0.0, 15.0)
    om_r, S_r = ref.hhg_spectrum_direct(t, j, dw, 0.0, wmax)
    assert np.array_equal(om_l, om_r)
    above = S_r > 1e-10 * S_r.max()
    assert np.abs(np.log10(S_l[above]) - np.log10(S_r[above])).max() < 1e-8
    for width in (0.5 * w0, 0.05 * w0, -1.0):
        sm_l = smear_spectrum(om_r, S_r, width)
        sm_r = ref.smear_spectrum(om_r, S_r, width)
        assert np.abs(sm_l - sm_r).max() <= 1e-12 * sm_r.max()
