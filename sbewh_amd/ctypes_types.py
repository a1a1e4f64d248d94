"""ctypes mirrors of include/sbewh_types.h and builders from a ``model.Config``.

The ``Inputs`` object owns the numpy buffers the C structs point into; keep it
alive for as long as a library context built from it exists.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import units
from .model import Config, Pulse, PULSE_TYPES


class SbeModel(C.Structure):
    _fields_ = [("num_wann", C.c_int), ("num_cells", C.c_int),
                ("cells", C.POINTER(C.c_int)), ("H", C.POINTER(C.c_double)),
                ("D", C.POINTER(C.c_double)), ("lattice", C.c_double * 9)]


class SbeGrid(C.Structure):
    _fields_ = [("dim", C.c_int), ("Nk", C.c_int * 3), ("Nf", C.c_int * 3),
                ("k_global_shift", C.c_double * 3)]


class SbeParams(C.Structure):
    _fields_ = [("relaxation_F2", C.c_double), ("soothing_width", C.c_double),
                ("gs_temp", C.c_double), ("eps_abs", C.c_double), ("eps_rel", C.c_double),
                ("sweep_count", C.c_int), ("flags", C.c_int)]


class SbeSubpulse(C.Structure):
    _fields_ = [("type", C.c_int), ("cycles_on", C.c_int), ("E0", C.c_double),
                ("omega", C.c_double), ("cep", C.c_double), ("fwhm", C.c_double),
                ("t_start_fwhm", C.c_double), ("t_end_fwhm", C.c_double),
                ("center", C.c_double), ("rising_cycles", C.c_double), ("pol", C.c_double * 3)]


class SbeInstance(C.Structure):
    _fields_ = [("Nk", C.c_int * 3), ("pulse_offset", C.c_int), ("num_pulses", C.c_int),
                ("relaxation_F2", C.c_double)]


class SbeRegion(C.Structure):
    _fields_ = [("k", C.c_double * 3), ("fwhm", C.c_double), ("moving_frame", C.c_int)]


def region_array(regions) -> "C.Array":
    """regions: sequence of (k[3], fwhm, moving_frame) or objects with those
    attributes; fwhm = inf is the whole zone"""
    arr = (SbeRegion * max(1, len(regions)))()
    for i, r in enumerate(regions):
        k, fwhm, mv = (r.k, r.fwhm, r.moving_frame) if hasattr(r, "fwhm") else r
        for d in range(3):
            arr[i].k[d] = float(k[d])
        arr[i].fwhm = float(fwhm)
        arr[i].moving_frame = int(bool(mv))
    return arr


class SbeObservations(C.Structure):
    _fields_ = [(name, C.POINTER(C.c_double)) for name in
                ("t", "E", "A", "weight_sum", "j", "occ_ham", "occ_wan", "occ_ham_min", "occ_ham_max", "wan")] + \
               [("density_index", C.POINTER(C.c_int))]


class SbeSteplog(C.Structure):
    _fields_ = [("capacity", C.c_int), ("count", C.c_int), ("t", C.POINTER(C.c_double)),
                ("dt", C.POINTER(C.c_double)), ("err", C.POINTER(C.c_double)),
                ("accepted", C.POINTER(C.c_int))]


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def iptr(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_int))


def subpulse(p: Pulse) -> SbeSubpulse:
    """reference unit conversion of one sub-pulse (code/src/io/PulseParameter.cpp:133-138)."""
    s = SbeSubpulse()
    s.type = PULSE_TYPES[p.type]
    s.cycles_on = int(p.cycles_on)
    s.E0 = units.from_V_nm(p.Emax_V_nm)
    s.omega = units.omega_from_lambda_nm(p.lambda_nm)
    s.cep = p.cep
    s.fwhm = units.from_fs(p.fwhm_fs)
    s.t_start_fwhm = p.t_start
    s.t_end_fwhm = p.t_end
    s.center = units.from_fs(p.t_central_fs)
    s.rising_cycles = p.rising_cycles
    for d in range(3):
        s.pol[d] = p.pol[d]
    return s


def pulse_array(pulses: list):
    arr = (SbeSubpulse * len(pulses))()
    for i, p in enumerate(pulses):
        arr[i] = subpulse(p)
    return arr


class Inputs:
    """C views of a Config: model, grid, params, pulses."""

    def __init__(self, cfg: Config):
        m = cfg.model
        self.cfg = cfg
        self.cells = np.ascontiguousarray(m.cells, np.int32)
        self.H = np.ascontiguousarray(m.H.view(np.float64).reshape(-1))
        self.D = np.ascontiguousarray(m.D.view(np.float64).reshape(-1))
        self.model = SbeModel()
        self.model.num_wann = m.num_wann
        self.model.num_cells = m.num_cells
        self.model.cells = iptr(self.cells)
        self.model.H = dptr(self.H)
        self.model.D = dptr(self.D)
        for i, v in enumerate(np.asarray(m.lattice, float).reshape(-1)):
            self.model.lattice[i] = v
        self.grid = SbeGrid()
        self.grid.dim = cfg.dim
        for d in range(3):
            self.grid.Nk[d] = int(cfg.Nk[d]) if d < cfg.dim else 1
            self.grid.Nf[d] = int(cfg.Nf[d]) if d < cfg.dim else 1
            self.grid.k_global_shift[d] = cfg.k_global_shift[d]
        self.params = SbeParams()
        self.params.relaxation_F2 = 1 / units.from_fs(cfg.T2_fs) if cfg.T2_fs > 0 else -1.0
        self.params.soothing_width = units.from_eV(cfg.soothing_width_eV)
        self.params.gs_temp = units.from_K(cfg.temperature_K)
        self.params.eps_abs = cfg.eps_abs
        self.params.eps_rel = cfg.eps_rel
        self.params.sweep_count = cfg.sweeps
        self.params.flags = (1 if getattr(cfg, "cold_jacobi", False) else 0) | \
            (2 if getattr(cfg, "gpu_layout", False) else 0) | \
            (4 if getattr(cfg, "device_controller", False) else 0)
        self.pulses = pulse_array(cfg.pulses)
        self.num_pulses = len(cfg.pulses)


class StepLog:
    """numpy-backed sbe_steplog."""

    def __init__(self, capacity: int = 200000):
        self.t = np.zeros(capacity)
        self.dt = np.zeros(capacity)
        self.err = np.zeros(capacity)
        self.accepted = np.zeros(capacity, np.int32)
        self.c = SbeSteplog(capacity, 0, dptr(self.t), dptr(self.dt), dptr(self.err), iptr(self.accepted))

    @property
    def count(self) -> int:
        return min(self.c.count, self.c.capacity)

    def accepted_dts(self) -> np.ndarray:
        n = self.count
        return self.dt[:n][self.accepted[:n] == 1]

    def as_dict(self) -> dict:
        n = self.count
        return {"t": self.t[:n].copy(), "dt": self.dt[:n].copy(), "err": self.err[:n].copy(),
                "accepted": self.accepted[:n].copy()}


def time_window(cfg: Config) -> tuple:
    """[t0, t1] of the propagation: earliest start / latest end of the sub-pulses
    (code/src/pulse/Pulse3D.cpp:42-62, code/src/io/PropagationParameter.cpp:99-100)."""
    t0, t1 = math.inf, -math.inf
    for p in cfg.pulses:
        s = subpulse(p)
        if p.type == "sin2Ramp":
            half = 2 * math.pi / s.omega * (s.rising_cycles + 0.5 * s.cycles_on)
            a, b = -half, half
        else:
            a, b = s.t_start_fwhm * s.fwhm, s.t_end_fwhm * s.fwhm
        t0 = min(t0, a + s.center)
        t1 = max(t1, b + s.center)
    return t0, t1
