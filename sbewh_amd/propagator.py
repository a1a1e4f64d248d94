"""Python mirror of the reference's propagation interface over the C ABI.

Reference shapes kept:
  * ``Propagator.calcInitialState(fermiLevel)`` and ``Propagator(state, dsdt, t)``
    mirror PropagatorCPU (code/src/PropagatorCPU.h:52-55);
  * ``Observer(state, t)`` / ``collectExpectationValues()`` mirror ObserverCPU
    (code/src/ObserverCPU.h:43, code/src/ObserverBase.h:56) for the current j;
  * ``integrate_const`` mirrors the odeint call of propagateCPUintertwined
    (code/src/main/runSBE.cpp:170-183) and returns the accepted-step count.
All arithmetic runs in libsbewh_hip.so on the GPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import load, check, SbeError
from .ctypes_types import Inputs, SbeObservations, StepLog, dptr, pulse_array, region_array
from .model import Config


def _v3(x):
    return np.ascontiguousarray(np.asarray(x, float).reshape(3))


class SBEContext:
    """Owns one sbe_ctx (device buffers + stream) for a configuration."""

    def __init__(self, cfg: Config):
        L = load()
        self.cfg = cfg
        self.inp = Inputs(cfg)
        h = C.c_void_p()
        rc = L.sbe_create(C.byref(self.inp.model), C.byref(self.inp.grid), C.byref(self.inp.params),
                          self.inp.pulses, self.inp.num_pulses, C.byref(h))
        if rc < 0:
            raise SbeError(f"sbe_create failed ({rc}): {L.sbe_last_error(None).decode()}")
        self.h = h
        self.n = cfg.model.num_wann
        self.state_size = L.sbe_state_size(h)
        self.num_k = L.sbe_num_k(h)
        self.P = L.sbe_matrix_stride(h)

    def close(self):
        if getattr(self, "h", None):
            load().sbe_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_pulses(self, pulses):
        arr = pulse_array(pulses)
        check(load().sbe_set_pulses(self.h, arr, len(pulses)), self.h)

    @property
    def last_axis_cells(self) -> int:
        """extent W_L of the model's R box along the last grid axis (the axis
        the stage kernel contracts while loading H(k))."""
        r = np.asarray(self.cfg.model.cells)[:, self.cfg.dim - 1]
        return int(r.max() - r.min() + 1)

    @property
    def nf(self):
        a = (C.c_int * 3)()
        check(load().sbe_get_nf(self.h, a), self.h)
        return tuple(a)

    @property
    def time_window(self):
        a, b = C.c_double(), C.c_double()
        check(load().sbe_time_window(self.h, C.byref(a), C.byref(b)), self.h)
        return a.value, b.value

    def E(self, t: float) -> np.ndarray:
        e = np.zeros(3)
        check(load().sbe_pulse_E(self.h, t, dptr(e)), self.h)
        return e

    def new_state(self) -> np.ndarray:
        return np.zeros(self.state_size, complex)

    def fermi_level(self, occupied_bands: int = 0, below_gap: bool = False) -> float:
        mu = C.c_double()
        check(load().sbe_fermi_level(self.h, occupied_bands, int(below_gap), C.byref(mu)), self.h)
        return mu.value

    def initial_state(self, mu: float) -> np.ndarray:
        st = self.new_state()
        check(load().sbe_initial_state(self.h, mu, dptr(st.view(float))), self.h)
        return st

    def rhs(self, state: np.ndarray, t: float) -> np.ndarray:
        state = np.ascontiguousarray(state, complex)
        out = np.empty_like(state)
        check(load().sbe_rhs(self.h, dptr(state.view(float)), dptr(out.view(float)), t), self.h)
        return out

    def rhs_device(self, d_state_ptr: int, d_dsdt_ptr: int, t: float, stream: int = 0):
        check(load().sbe_rhs_device(self.h, C.c_void_p(d_state_ptr), C.c_void_p(d_dsdt_ptr), t,
                                    C.c_void_p(stream) if stream else None), self.h)

    def observe(self, state: np.ndarray) -> np.ndarray:
        state = np.ascontiguousarray(state, complex)
        j = np.zeros(3)
        check(load().sbe_observe(self.h, dptr(state.view(float)), dptr(j)), self.h)
        return j

    def integrate(self, state: np.ndarray, output_count: int, replay: StepLog | None = None,
                  log_capacity: int = 400000, observer=None) -> dict:
        """device-resident integrate_const; `state` is advanced in place;
        observer: an optional FullObserver recording every output time;
        log_capacity 0: no step log (with Config.device_controller, 1-D and
        n <= 8: the resident integrator, step control on the GPU)."""
        assert state.dtype == np.complex128 and state.flags.c_contiguous
        t = np.zeros(output_count)
        j = np.zeros((output_count, 3))
        log = StepLog(log_capacity) if log_capacity else None
        lp = C.byref(log.c) if log is not None else None
        rhs = C.c_long()
        rp = C.byref(replay.c) if replay is not None else None
        if observer is not None:
            steps = load().sbe_integrate_observed(self.h, dptr(state.view(float)), output_count, observer.h, dptr(t),
                                                  dptr(j.reshape(-1)), lp, rp, C.byref(rhs))
        else:
            steps = load().sbe_integrate(self.h, dptr(state.view(float)), output_count, dptr(t),
                                         dptr(j.reshape(-1)), lp, rp, C.byref(rhs))
        check(steps, self.h)
        dev = C.c_int()
        check(load().sbe_integrate_info(self.h, C.byref(dev)), self.h)
        return {"t": t, "j": j, "steps": int(steps), "rhs_calls": int(rhs.value), "log": log,
                "device_controller": bool(dev.value)}

    def hk_mesh(self, A=(0, 0, 0), E=(0, 0, 0)) -> np.ndarray:
        out = np.zeros((self.num_k, 2, self.n, self.n), complex)
        check(load().sbe_hk_mesh(self.h, dptr(_v3(A)), dptr(_v3(E)), dptr(out.view(float).reshape(-1))), self.h)
        return out

    def bench_steps(self, state: np.ndarray, t: float, dt: float, warmup: int, steps: int,
                    kernel_timing: bool = True) -> dict:
        """time `steps` adaptive controller attempts (after `warmup`) advancing a
        device copy of `state` from time t; see sbe_bench_steps."""
        ms, sms, pms = C.c_double(), C.c_double(), C.c_double()
        rhs = C.c_long()
        check(load().sbe_bench_steps(self.h, dptr(state.view(float)), t, dt, warmup, steps, C.byref(ms),
                                     C.byref(rhs), C.byref(sms) if kernel_timing else None,
                                     C.byref(pms) if kernel_timing else None), self.h)
        return {"ms": ms.value, "rhs": rhs.value, "stage_ms": sms.value, "phase_ms": pms.value}

    def k_index_map(self) -> np.ndarray:
        """(N_k, 3) integer grid indices of the state's density matrices"""
        g = np.zeros((self.num_k, 3), np.int32)
        check(load().sbe_k_index_map(self.h, g.ctypes.data_as(C.POINTER(C.c_int))), self.h)
        return g

    def stage_tuning(self) -> dict:
        lr, cnt = C.c_int(), C.c_int()
        var = (C.c_int * 8)()
        ms = (C.c_double * 8)()
        check(load().sbe_stage_tuning(self.h, C.byref(lr), C.byref(cnt), var, ms), self.h)
        return {"lanes_per_row": lr.value, "timed": {int(var[i]): float(ms[i]) for i in range(cnt.value)}}

    def rho(self, state: np.ndarray) -> np.ndarray:
        n = self.n
        return state[8:].reshape(self.num_k, self.P)[:, :n * n].reshape(self.num_k, n, n)


def jacobi_batch(H: np.ndarray, sweeps: int):
    """wavefront Jacobi on a batch (count, n, n) of Hermitian matrices -> (U, eig)."""
    H = np.ascontiguousarray(H, complex)
    cnt, n, _ = H.shape
    U = np.zeros_like(H)
    e = np.zeros((cnt, n))
    check(load().sbe_jacobi_batch(n, sweeps, cnt, dptr(H.view(float).reshape(-1)), dptr(U.view(float).reshape(-1)),
                                  dptr(e.reshape(-1))))
    return U, e


class Propagator:
    """RHS functor with the reference signature (code/src/PropagatorCPU.h:52-55)."""

    def __init__(self, ctx: SBEContext):
        self.ctx = ctx

    def calcInitialState(self, fermiLevel: float) -> np.ndarray:
        return self.ctx.initial_state(fermiLevel)

    def __call__(self, state: np.ndarray, dsdt: np.ndarray, t: float) -> None:
        dsdt[:] = self.ctx.rhs(state, t)

    def reset(self):
        pass


class Observer:
    """Current observer (code/src/ObserverCPU.h:43): records t, A and j per call."""

    def __init__(self, ctx: SBEContext):
        self.ctx = ctx
        self.reset()

    def reset(self):
        self.t, self.A, self.E, self.j = [], [], [], []

    def __call__(self, state: np.ndarray, t: float) -> None:
        self.t.append(t)
        self.A.append(state[:3].real.copy())
        self.E.append(self.ctx.E(t))
        self.j.append(self.ctx.observe(state))

    def collectExpectationValues(self) -> dict:
        return {"t": np.array(self.t), "A": np.array(self.A), "E": np.array(self.E),
                "expectationValues/j": np.array(self.j)}


class Region:
    """One k-space region sample (ObserverRegion_t, code/src/ObserverBase.h:33-37):
    centre k (fractional), fwhm (fractional; inf = whole zone), moving frame."""

    def __init__(self, k, fwhm=float("inf"), moving_frame=True, name="expectationValues"):
        self.k = tuple(float(x) for x in k)
        self.fwhm = float(fwhm)
        self.moving_frame = bool(moving_frame)
        self.name = name


WHOLE_ZONE = Region((0.0, 0.0, 0.0))


class FullObserver:
    """ObserverCPU with k-space regions, occupations and density matrices
    (code/src/ObserverCPU.cpp:137-308) on the GPU.  ``record(state, t)`` is
    ObserverCPU::operator(); ``collectExpectationValues()`` returns the
    records like ExpectationValues_t (code/src/ExpectationValues.h:9-28):
    t, E, A, occupationHamMin / Max [t][n], and per region sample
    kPointWeightSum [t], j [t][3], occupationHam / occupationWan [t][n],
    wan [tDens][n][n] with tDensIndices."""

    def __init__(self, ctx: SBEContext, regions=(WHOLE_ZONE,), smearing_width: float = -1.0,
                 density_stride: int = 0, capacity: int = 4096):
        self.ctx = ctx
        self.regions = list(regions)
        h = C.c_void_p()
        arr = region_array(self.regions)
        check(load().sbe_observer_create(ctx.h, arr, len(self.regions), smearing_width, density_stride, capacity,
                                         C.byref(h)), ctx.h)
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            load().sbe_observer_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def reset(self):
        check(load().sbe_observer_reset(self.h), self.ctx.h)

    def __call__(self, state: np.ndarray, t: float) -> None:
        self.record(state, t)

    def record(self, state: np.ndarray, t: float) -> None:
        state = np.ascontiguousarray(state, complex)
        check(load().sbe_observer_record(self.h, dptr(state.view(float)), t), self.ctx.h)

    def count(self):
        dc = C.c_int()
        c = load().sbe_observer_count(self.h, C.byref(dc))
        check(c, self.ctx.h)
        return c, dc.value

    def collectExpectationValues(self) -> dict:
        cnt, dcnt = self.count()
        R, n = len(self.regions), self.ctx.n
        out = {"t": np.zeros(cnt), "E": np.zeros((cnt, 3)), "A": np.zeros((cnt, 3)),
               "kPointWeightSum": np.zeros((cnt, R)), "j": np.zeros((cnt, R, 3)),
               "occupationHam": np.zeros((cnt, R, n)), "occupationWan": np.zeros((cnt, R, n)),
               "occupationHamMin": np.zeros((cnt, n)), "occupationHamMax": np.zeros((cnt, n)),
               "wan": np.zeros((dcnt, R, n, n), complex), "tDensIndices": np.zeros(dcnt, np.int32)}
        ob = SbeObservations()
        for key, name in (("t", "t"), ("E", "E"), ("A", "A"), ("kPointWeightSum", "weight_sum"), ("j", "j"),
                          ("occupationHam", "occ_ham"), ("occupationWan", "occ_wan"),
                          ("occupationHamMin", "occ_ham_min"), ("occupationHamMax", "occ_ham_max")):
            setattr(ob, name, dptr(out[key].reshape(-1)))
        ob.wan = dptr(out["wan"].view(float).reshape(-1)) if dcnt else None
        ob.density_index = out["tDensIndices"].ctypes.data_as(C.POINTER(C.c_int)) if dcnt else None
        check(load().sbe_observer_results(self.h, C.byref(ob)), self.ctx.h)
        return out


def integrate_const(ctx: SBEContext, state: np.ndarray, output_count: int, observer=None,
                    replay: StepLog | None = None) -> int:
    """integrate_const(controlled dopri5, propagator, state, t0, t1 + dt/2, dt, observer),
    device resident; returns the number of accepted steps.  observer: an
    Observer (current only) or a FullObserver (recorded on the device at every
    output time)."""
    if isinstance(observer, FullObserver):
        res = ctx.integrate(state, output_count, replay=replay, observer=observer)
        return res["steps"]
    res = ctx.integrate(state, output_count, replay=replay)
    if observer is not None:
        observer.reset()
        for t, j in zip(res["t"], res["j"]):
            observer.t.append(t)
            observer.j.append(j)
            observer.E.append(ctx.E(t))
    return res["steps"]


__all__ = ["SBEContext", "Propagator", "Observer", "FullObserver", "Region", "WHOLE_ZONE", "integrate_const",
           "jacobi_batch", "SbeError"]
