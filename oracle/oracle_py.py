"""ctypes binding of the CPU oracle (test infrastructure only).

Imported by tests/, by __graft_entry__.smoke() (as the checker) and by
bench.py's cpu_baseline leg.  The product package ``sbewh_amd`` never imports
this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from sbewh_amd.ctypes_types import (Inputs, SbeModel, SbeGrid, SbeParams, SbeRegion, SbeSubpulse, SbeSteplog,
                                     StepLog, dptr, region_array)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "libsbe_oracle.so")
# the same source built at -O1, the reference's own optimisation level
LIB_PATH_O1 = os.path.join(HERE, "build", "libsbe_oracle_O1.so")


def build(force: bool = False) -> str:
    """compile the oracle with its Makefile (gcc + OpenMP)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE] + (["-B"] if force else []), check=True)
    return LIB_PATH


_libs = {}


def lib(path: str = LIB_PATH):
    if path not in _libs:
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        P = C.POINTER
        vp = C.c_void_p
        L.oracle_create.restype = vp
        L.oracle_create.argtypes = [P(SbeModel), P(SbeGrid), P(SbeParams), P(SbeSubpulse), C.c_int]
        L.oracle_destroy.argtypes = [vp]
        L.oracle_last_error.restype = C.c_char_p
        L.oracle_set_pulses.argtypes = [vp, P(SbeSubpulse), C.c_int]
        L.oracle_set_phase_method.argtypes = [vp, C.c_int]
        L.oracle_state_size.restype = C.c_long
        L.oracle_state_size.argtypes = [vp]
        L.oracle_num_k.restype = C.c_long
        L.oracle_num_k.argtypes = [vp]
        L.oracle_padded_mat_size.argtypes = [vp]
        L.oracle_get_nf.argtypes = [vp, P(C.c_int)]
        L.oracle_k_frac.argtypes = [vp, C.c_long, P(C.c_double)]
        L.oracle_pulse_E.argtypes = [vp, C.c_double, P(C.c_double)]
        L.oracle_pulse_A.argtypes = [vp, C.c_double, P(C.c_double)]
        L.oracle_start_time.restype = C.c_double
        L.oracle_start_time.argtypes = [vp]
        L.oracle_end_time.restype = C.c_double
        L.oracle_end_time.argtypes = [vp]
        L.oracle_hk_point.argtypes = [vp, P(C.c_double), P(C.c_double), P(C.c_double), P(C.c_double)]
        L.oracle_hk_full_point.argtypes = [vp, P(C.c_double), P(C.c_double), P(C.c_double), P(C.c_double)]
        L.oracle_hk_mesh.argtypes = [vp, P(C.c_double), P(C.c_double), P(C.c_double)]
        L.oracle_jacobi.argtypes = [C.c_int, C.c_int, P(C.c_double), P(C.c_double)]
        L.oracle_rhs.argtypes = [vp, P(C.c_double), P(C.c_double), C.c_double]
        L.oracle_initial_state.argtypes = [vp, C.c_double, P(C.c_double)]
        L.oracle_fermi_level.argtypes = [vp, C.c_int, C.c_int, P(C.c_double)]
        L.oracle_observe_j.argtypes = [vp, P(C.c_double), P(C.c_double)]
        L.oracle_observe_full.argtypes = [vp, P(C.c_double), P(SbeRegion), C.c_int, C.c_double] + [P(C.c_double)] * 7
        L.oracle_integrate.restype = C.c_long
        L.oracle_integrate.argtypes = [vp, P(C.c_double), C.c_int, P(C.c_double), P(C.c_double),
                                       P(SbeSteplog), P(SbeSteplog), P(C.c_long)]
        L.oracle_bench_attempts.restype = C.c_double
        L.oracle_bench_attempts.argtypes = [vp, P(C.c_double), C.c_double, C.c_double, C.c_int]
        L.oracle_num_threads.restype = C.c_int
        L.oracle_num_threads.argtypes = []
        _libs[path] = L
    return _libs[path]


def _v3(x):
    return np.ascontiguousarray(np.asarray(x, float).reshape(3))


def jacobi(A: np.ndarray, sweeps: int):
    """(U, eig) of the reference's cyclic Jacobi; U A U^dagger = diag(eig)."""
    n = A.shape[0]
    a = np.ascontiguousarray(A, np.complex128).copy()
    e = np.zeros(n)
    lib().oracle_jacobi(n, sweeps, dptr(a.view(np.float64).reshape(-1)), dptr(e))
    return a, e


class Oracle:
    """CPU restatement of PropagatorCPU + ObserverCPU + the dopri5 driver."""

    def __init__(self, cfg, phase_method: str = "separable", lib_path: str = LIB_PATH):
        self.inp = Inputs(cfg)
        self.L = L = lib(lib_path)
        self.h = L.oracle_create(C.byref(self.inp.model), C.byref(self.inp.grid), C.byref(self.inp.params),
                                 self.inp.pulses, self.inp.num_pulses)
        if not self.h:
            raise RuntimeError("oracle_create: " + L.oracle_last_error().decode())
        L.oracle_set_phase_method(self.h, 0 if phase_method == "direct" else 1)
        self.n = cfg.model.num_wann
        self.state_size = L.oracle_state_size(self.h)
        self.num_k = L.oracle_num_k(self.h)
        self.P = L.oracle_padded_mat_size(self.h)

    def close(self):
        if self.h:
            self.L.oracle_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_pulses(self, pulses):
        from sbewh_amd.ctypes_types import pulse_array
        arr = pulse_array(pulses)
        if self.L.oracle_set_pulses(self.h, arr, len(pulses)):
            raise RuntimeError(self.L.oracle_last_error().decode())

    def set_phase_method(self, method: str):
        self.L.oracle_set_phase_method(self.h, 0 if method == "direct" else 1)

    @property
    def nf(self):
        a = (C.c_int * 3)()
        self.L.oracle_get_nf(self.h, a)
        return tuple(a)

    def kfrac(self, k: int) -> np.ndarray:
        a = np.zeros(3)
        self.L.oracle_k_frac(self.h, k, dptr(a))
        return a

    def E(self, t):
        a = np.zeros(3)
        self.L.oracle_pulse_E(self.h, t, dptr(a))
        return a

    def A(self, t):
        a = np.zeros(3)
        self.L.oracle_pulse_A(self.h, t, dptr(a))
        return a

    @property
    def t_start(self):
        return self.L.oracle_start_time(self.h)

    @property
    def t_end(self):
        return self.L.oracle_end_time(self.h)

    def hk_point(self, kfrac, E=(0, 0, 0)):
        n = self.n
        H0 = np.zeros((n, n), complex)
        HE = np.zeros((n, n), complex)
        self.L.oracle_hk_point(self.h, dptr(_v3(kfrac)), dptr(_v3(E)), dptr(H0.view(float).reshape(-1)),
                              dptr(HE.view(float).reshape(-1)))
        return H0, HE

    def hk_full_point(self, kfrac):
        n = self.n
        H0 = np.zeros((n, n), complex)
        D = np.zeros((3, n, n), complex)
        dH = np.zeros((3, n, n), complex)
        self.L.oracle_hk_full_point(self.h, dptr(_v3(kfrac)), dptr(H0.view(float).reshape(-1)),
                                   dptr(D.view(float).reshape(-1)), dptr(dH.view(float).reshape(-1)))
        return H0, D, dH

    def hk_mesh(self, A=(0, 0, 0), E=(0, 0, 0)):
        out = np.zeros((self.num_k, 2, self.n, self.n), complex)
        self.L.oracle_hk_mesh(self.h, dptr(_v3(A)), dptr(_v3(E)), dptr(out.view(float).reshape(-1)))
        return out

    def new_state(self):
        return np.zeros(self.state_size, complex)

    def rhs(self, state: np.ndarray, t: float) -> np.ndarray:
        out = np.zeros_like(state)
        self.L.oracle_rhs(self.h, dptr(state.view(float)), dptr(out.view(float)), t)
        return out

    def initial_state(self, mu: float) -> np.ndarray:
        st = self.new_state()
        self.L.oracle_initial_state(self.h, mu, dptr(st.view(float)))
        return st

    def fermi_level(self, occupied_bands: int = 0, below_gap: bool = False) -> float:
        mu = C.c_double()
        if self.L.oracle_fermi_level(self.h, occupied_bands, int(below_gap), C.byref(mu)):
            raise RuntimeError(self.L.oracle_last_error().decode())
        return mu.value

    def observe_j(self, state: np.ndarray) -> np.ndarray:
        j = np.zeros(3)
        self.L.oracle_observe_j(self.h, dptr(state.view(float)), dptr(j))
        return j

    def observe_full(self, state: np.ndarray, regions, smearing: float, density: bool = False) -> dict:
        """ObserverCPU at one output time for the region samples `regions`
        ((k[3], fwhm, moving) each): dict of weight_sum [R], j [R,3],
        occ_ham [R,n], occ_wan [R,n], occ_min [n], occ_max [n], wan [R,n,n]"""
        R, n = len(regions), self.n
        out = {"weight_sum": np.zeros(R), "j": np.zeros((R, 3)), "occ_ham": np.zeros((R, n)),
               "occ_wan": np.zeros((R, n)), "occ_min": np.zeros(n), "occ_max": np.zeros(n)}
        wan = np.zeros((R, n, n), complex) if density else None
        self.L.oracle_observe_full(self.h, dptr(state.view(float)), region_array(regions), R, smearing,
                                   dptr(out["weight_sum"]), dptr(out["j"].reshape(-1)), dptr(out["occ_ham"].reshape(-1)),
                                   dptr(out["occ_wan"].reshape(-1)), dptr(out["occ_min"]), dptr(out["occ_max"]),
                                   dptr(wan.view(float).reshape(-1)) if density else None)
        if density:
            out["wan"] = wan
        return out

    def integrate(self, state: np.ndarray, output_count: int, replay: StepLog | None = None,
                  log_capacity: int = 400000):
        """returns dict(t, j, steps, rhs_calls, log); state is advanced in place."""
        t = np.zeros(output_count)
        j = np.zeros((output_count, 3))
        log = StepLog(log_capacity)
        rhs = C.c_long()
        steps = self.L.oracle_integrate(self.h, dptr(state.view(float)), output_count, dptr(t),
                                       dptr(j.reshape(-1)), C.byref(log.c),
                                       C.byref(replay.c) if replay is not None else None, C.byref(rhs))
        if steps < 0:
            raise RuntimeError("oracle_integrate: " + self.L.oracle_last_error().decode())
        return {"t": t, "j": j, "steps": int(steps), "rhs_calls": int(rhs.value), "log": log}

    @property
    def num_threads(self) -> int:
        """OpenMP threads of the oracle's parallel loops"""
        return self.L.oracle_num_threads()

    def bench_attempts(self, state: np.ndarray, t: float, dt: float, count: int) -> float:
        """wall seconds of `count` dopri5 attempts (6 RHS each) at fixed (t, dt)."""
        return self.L.oracle_bench_attempts(self.h, dptr(state.view(float)), t, dt, count)

    # reference layout helpers
    def rho(self, state: np.ndarray) -> np.ndarray:
        """(N_k, n, n) view of the density matrices inside a state vector."""
        n = self.n
        body = state[8:].reshape(self.num_k, self.P)[:, :n * n]
        return body.reshape(self.num_k, n, n)
