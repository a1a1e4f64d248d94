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
