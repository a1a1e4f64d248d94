"""Batched scan of many pulses over one model (sbe_scan_* in include/sbewh_hip.h).

The reference runs the configurations of a multi-run one after the other
(code/src/main/runSBE.cpp:205-216); ``Scan`` advances all of them together on
the GPU, each with its own k-grid, pulse, T2 and adaptive step size, and
returns per instance what ``SBEContext.integrate`` returns for the same
inputs.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import math

import numpy as np

from . import model, units
from ._lib import SbeError, load
from .ctypes_types import Inputs, SbeInstance, SbeSubpulse, dptr, subpulse
from .model import Config


def _check(rc, h):
    if rc < 0:
        raise SbeError(f"sbewh scan error {rc}: {load().sbe_scan_last_error(h).decode()}")
    return rc


_SHARED = ("dim", "Nf", "soothing_width_eV", "temperature_K", "eps_abs", "eps_rel", "sweeps", "occupied_bands",
           "output_count", "k_global_shift", "cold_jacobi")


class Scan:
    """Owns one sbe_scan over configurations that share the model and every
    setting except Nk, pulses and T2_fs."""

    def __init__(self, configs: list):
        if not configs:
            raise ValueError("a scan needs at least one configuration")
        base = configs[0]
        for c in configs[1:]:
            if c.model is not base.model:
                raise ValueError("all configurations of a scan must share one model object")
            for f in _SHARED:
                if getattr(c, f) != getattr(base, f):
                    raise ValueError(f"configurations differ in {f}")
        L = load()
        self.configs = list(configs)
        self.inp = Inputs(base)
        pulses = []
        inst = (SbeInstance * len(configs))()
        for i, c in enumerate(configs):
            inst[i].pulse_offset = len(pulses)
            inst[i].num_pulses = len(c.pulses)
            for d in range(3):
                inst[i].Nk[d] = int(c.Nk[d]) if d < c.dim else 1
            inst[i].relaxation_F2 = 1 / units.from_fs(c.T2_fs) if c.T2_fs > 0 else -1.0
            pulses.extend(c.pulses)
        self._pulses = (SbeSubpulse * len(pulses))(*[subpulse(p) for p in pulses])
        self._inst = inst
        h = C.c_void_p()
        rc = L.sbe_scan_create(C.byref(self.inp.model), C.byref(self.inp.grid), C.byref(self.inp.params), inst,
                               len(configs), self._pulses, len(pulses), C.byref(h))
        if rc < 0:
            raise SbeError(f"sbe_scan_create failed ({rc}): {L.sbe_scan_last_error(None).decode()}")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            load().sbe_scan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, output_count: int | None = None) -> dict:
        """Propagate every instance; returns t [B, O], j [B, O, 3], steps [B]
        (negative: that instance's step control failed), rhs_calls [B] and the
        wall time in seconds."""
        B = len(self.configs)
        O = int(output_count or self.configs[0].output_count)
        t = np.zeros((B, O))
        j = np.zeros((B, O, 3))
        steps = np.zeros(B, np.int64)
        rhs = np.zeros(B, np.int64)
        secs = C.c_double()
        _check(load().sbe_scan_run(self.h, self.configs[0].occupied_bands, O, dptr(t), dptr(j),
                                   steps.ctypes.data_as(C.POINTER(C.c_long)), rhs.ctypes.data_as(C.POINTER(C.c_long)),
                                   C.byref(secs)), self.h)
        return {"t": t, "j": j, "steps": steps, "rhs_calls": rhs, "seconds": secs.value}

    def final_state(self, index: int) -> np.ndarray:
        L = load()
        n = _check(L.sbe_scan_state_size(self.h, index), self.h)
        out = np.zeros(2 * n)
        _check(L.sbe_scan_final_state(self.h, index, dptr(out)), self.h)
        return out.view(complex)

    def tuning(self) -> dict:
        lr = C.c_int()
        cnt = C.c_int()
        var = (C.c_int * 8)()
        ms = (C.c_double * 8)()
        _check(load().sbe_scan_tuning(self.h, C.byref(lr), C.byref(cnt), var, ms), self.h)
        return {"lanes_per_row": lr.value, "timed": {int(var[i]): float(ms[i]) for i in range(cnt.value)}}


def ellipticity_scan(count: int = 64, seed: int = 1004, grids=(60, 90, 120, 150), base: Config | None = None,
                     cycles: float = 8.0, lambda_nm: float = 3000.0) -> list:
    """The 64-pulse scan over the 2-D 8-band model: ellipticity ~ U[0, 1],
    E0 ~ U[3, 15] MV/cm (0.3 - 1.5 V/nm), CEP ~ U[0, 2 pi) and square k-grids
    from `grids`, each grid used equally often in a seeded random order (16 x
    {60^2, 90^2, 120^2, 150^2}: 777 600 k-points in total).  The pulse
    convention is that of the reference's ellipticity-scan example
    (examples/08_ellipticityScan/generateEllipticitiyRuns.py:18-23)."""
    base = base or model.config("square2d")
    rng = np.random.default_rng(seed)
    eps = rng.uniform(0.0, 1.0, count)
    E0 = rng.uniform(0.3, 1.5, count)
    cep = rng.uniform(0.0, 2 * math.pi, count)
    sizes = np.array([grids[i % len(grids)] for i in range(count)])
    rng.shuffle(sizes)
    out = []
    for i in range(count):
        g = int(sizes[i])
        pul = model.elliptic_sin2(float(E0[i]), float(eps[i]), lambda_nm, cycles, cep=float(cep[i]))
        out.append(dataclasses.replace(base, name=f"scan{i}", Nk=(g, g, 1), pulses=pul))
    return out
