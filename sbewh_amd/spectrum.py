"""HHG spectra and per-harmonic yields from j(t) (computed in libsbewh_hip.so).

Defaults follow the reference's spectrum options (scripts/util/common.py:166-186):
omega in (0, 15 eV), d_omega = 0.01 eV, FFT with fadeout 0.1, power 1.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import units
from ._lib import load, check
from .ctypes_types import dptr


def hhg_spectrum(t: np.ndarray, j: np.ndarray, d_omega_eV: float = 0.01, omega_min_eV: float = 0.0,
                 omega_max_eV: float = 15.0, fadeout_rel: float = 0.1, power: int = 1):
    """(omega [a.u.], S(omega)) of the current j (shape (nt, ncomp))."""
    t = np.ascontiguousarray(t, float)
    j = np.ascontiguousarray(np.atleast_2d(np.asarray(j, float).T).T)
    nt, nc = j.shape
    dw = units.from_eV(d_omega_eV)
    cap = int(max(nt, 2 * np.pi / (t[1] - t[0]) / dw)) // 2 + 2
    om = np.zeros(cap)
    S = np.zeros(cap)
    cnt = C.c_int()
    check(load().sbe_hhg_spectrum(dptr(t), dptr(j.reshape(-1)), nt, nc, dw, units.from_eV(omega_min_eV),
                                  units.from_eV(omega_max_eV), fadeout_rel, power, dptr(om), dptr(S), cap,
                                  C.byref(cnt)))
    return om[:cnt.value].copy(), S[:cnt.value].copy()


def harmonic_yields(omega: np.ndarray, S: np.ndarray, omega0: float, width: float = 0.5, hmax: int = 40):
    """Y_h = sum S over [h w0 - width w0, h w0 + width w0), h = 1..hmax."""
    omega = np.ascontiguousarray(omega, float)
    S = np.ascontiguousarray(S, float)
    y = np.zeros(hmax)
    check(load().sbe_harmonic_yields(dptr(omega), dptr(S), len(omega), omega0, width, hmax, dptr(y)))
    return y


def direct_omega(t: np.ndarray, omega_min: float, omega_max: float, d_omega: float) -> np.ndarray:
    """directOmega: the grid of the direct spectrum (a.u.)"""
    t = np.ascontiguousarray(t, float)
    cnt = C.c_int()
    load().sbe_direct_omega(dptr(t), len(t), omega_min, omega_max, d_omega, None, 0, C.byref(cnt))
    om = np.zeros(max(1, cnt.value))
    check(load().sbe_direct_omega(dptr(t), len(t), omega_min, omega_max, d_omega, dptr(om), cnt.value, C.byref(cnt)))
    return om[:cnt.value]


def direct_spectrum(t: np.ndarray, omega: np.ndarray, j: np.ndarray, power: int = 1) -> np.ndarray:
    """directSpectrum of one real signal: |omega^power sum_t exp(i omega t) j(t)|^2"""
    t = np.ascontiguousarray(t, float)
    omega = np.ascontiguousarray(omega, float)
    j = np.ascontiguousarray(j, float)
    S = np.zeros(len(omega))
    check(load().sbe_direct_spectrum(dptr(t), len(t), dptr(omega), len(omega), dptr(j), power, dptr(S)))
    return S


def hhg_spectrum_direct(t: np.ndarray, j: np.ndarray, d_omega_eV: float = 0.01, omega_min_eV: float = 0.0,
                        omega_max_eV: float = 15.0, power: int = 1):
    """(omega [a.u.], S) of the current j (nt, ncomp) by the direct sum (useFFT = False)"""
    t = np.ascontiguousarray(t, float)
    j = np.ascontiguousarray(np.atleast_2d(np.asarray(j, float).T).T)
    nt, nc = j.shape
    args = (dptr(t), dptr(j.reshape(-1)), nt, nc, units.from_eV(d_omega_eV), units.from_eV(omega_min_eV),
            units.from_eV(omega_max_eV), power)
    cnt = C.c_int()
    load().sbe_hhg_spectrum_direct(*args, None, None, 0, C.byref(cnt))
    om, S = np.zeros(max(1, cnt.value)), np.zeros(max(1, cnt.value))
    check(load().sbe_hhg_spectrum_direct(*args, dptr(om), dptr(S), cnt.value, C.byref(cnt)))
    return om[:cnt.value], S[:cnt.value]


def smear_spectrum(omega: np.ndarray, S: np.ndarray, smear_width: float) -> np.ndarray:
    """smearSpectrum: Gaussian filter of FWHM smear_width (a.u.); < 0 returns S"""
    omega = np.ascontiguousarray(omega, float)
    S = np.ascontiguousarray(S, float)
    out = np.zeros_like(S)
    check(load().sbe_smear_spectrum(dptr(omega), dptr(S), len(S), smear_width, dptr(out)))
    return out
