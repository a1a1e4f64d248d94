"""ctypes loader for libsbewh_hip.so (the C ABI in include/sbewh_hip.h).

There is no CPU fallback: if the library is missing, loading raises.
"""
from __future__ import annotations

import ctypes as C
import os

from .ctypes_types import (SbeModel, SbeGrid, SbeParams, SbeSubpulse, SbeSteplog, SbeInstance, SbeRegion,
                           SbeObservations)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libsbewh_hip.so")

# every symbol include/sbewh_hip.h declares
EXPORTS = ("sbe_version", "sbe_last_error", "sbe_create", "sbe_destroy", "sbe_set_pulses", "sbe_state_size",
           "sbe_num_k", "sbe_get_nf", "sbe_time_window", "sbe_pulse_E", "sbe_pulse_field", "sbe_pulse_window",
           "sbe_fermi_level", "sbe_initial_state",
           "sbe_matrix_stride", "sbe_k_index_map", "sbe_rhs", "sbe_rhs_device", "sbe_observe", "sbe_integrate", "sbe_hk_mesh", "sbe_jacobi_batch",
           "sbe_bench_steps", "sbe_stage_tuning", "sbe_hhg_spectrum", "sbe_harmonic_yields", "sbe_scan_create",
           "sbe_scan_destroy", "sbe_scan_last_error", "sbe_scan_run", "sbe_scan_state_size", "sbe_scan_final_state",
           "sbe_scan_tuning", "sbe_observer_create", "sbe_observer_destroy", "sbe_observer_reset",
           "sbe_observer_record", "sbe_observer_record_device", "sbe_observer_count", "sbe_observer_results",
           "sbe_integrate_observed", "sbe_integrate_info", "sbe_observe_device_noalloc", "sbe_direct_omega", "sbe_direct_spectrum",
           "sbe_hhg_spectrum_direct", "sbe_smear_spectrum")
# every symbol include/sbewh_surrogate.h declares
SURROGATE_EXPORTS = ("sbe_mlp_create", "sbe_mlp_destroy", "sbe_mlp_last_error", "sbe_mlp_param_count",
                     "sbe_mlp_get_params", "sbe_mlp_set_params", "sbe_mlp_forward", "sbe_mlp_gradient",
                     "sbe_mlp_train")

_lib = None


class SbeError(RuntimeError):
    pass


def load():
    """Load the HIP library (built by ``python -m sbewh_amd.build``)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise SbeError(f"{LIB_PATH} is missing: build it with `python -m sbewh_amd.build` "
                       "(there is no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    P, vp, d, i, l = C.POINTER, C.c_void_p, C.c_double, C.c_int, C.c_long
    dp = P(C.c_double)
    L.sbe_version.restype = C.c_char_p
    L.sbe_last_error.restype = C.c_char_p
    L.sbe_last_error.argtypes = [vp]
    L.sbe_create.argtypes = [P(SbeModel), P(SbeGrid), P(SbeParams), P(SbeSubpulse), i, P(vp)]
    L.sbe_destroy.argtypes = [vp]
    L.sbe_set_pulses.argtypes = [vp, P(SbeSubpulse), i]
    L.sbe_state_size.restype = l
    L.sbe_state_size.argtypes = [vp]
    L.sbe_num_k.restype = l
    L.sbe_num_k.argtypes = [vp]
    L.sbe_get_nf.argtypes = [vp, P(C.c_int)]
    L.sbe_matrix_stride.argtypes = [vp]
    L.sbe_k_index_map.argtypes = [vp, P(C.c_int)]
    L.sbe_time_window.argtypes = [vp, dp, dp]
    L.sbe_pulse_E.argtypes = [vp, d, dp]
    L.sbe_pulse_field.argtypes = [P(SbeSubpulse), i, d, dp]
    L.sbe_pulse_window.argtypes = [P(SbeSubpulse), i, dp, dp]
    L.sbe_fermi_level.argtypes = [vp, i, i, dp]
    L.sbe_initial_state.argtypes = [vp, d, dp]
    L.sbe_rhs.argtypes = [vp, dp, dp, d]
    L.sbe_rhs_device.argtypes = [vp, vp, vp, d, vp]
    L.sbe_observe.argtypes = [vp, dp, dp]
    L.sbe_integrate.restype = l
    L.sbe_integrate.argtypes = [vp, dp, i, dp, dp, P(SbeSteplog), P(SbeSteplog), P(C.c_long)]
    L.sbe_hk_mesh.argtypes = [vp, dp, dp, dp]
    L.sbe_jacobi_batch.argtypes = [i, i, l, dp, dp, dp]
    L.sbe_bench_steps.argtypes = [vp, dp, d, d, i, i, dp, P(C.c_long), dp, dp]
    L.sbe_stage_tuning.argtypes = [vp, P(C.c_int), P(C.c_int), P(C.c_int), dp]
    L.sbe_hhg_spectrum.argtypes = [dp, dp, i, i, d, d, d, d, i, dp, dp, i, P(C.c_int)]
    L.sbe_harmonic_yields.argtypes = [dp, dp, i, d, d, i, dp]
    L.sbe_direct_omega.argtypes = [dp, i, d, d, d, dp, i, P(C.c_int)]
    L.sbe_direct_spectrum.argtypes = [dp, i, dp, i, dp, i, dp]
    L.sbe_hhg_spectrum_direct.argtypes = [dp, dp, i, i, d, d, d, i, dp, dp, i, P(C.c_int)]
    L.sbe_smear_spectrum.argtypes = [dp, dp, i, d, dp]
    L.sbe_scan_create.argtypes = [P(SbeModel), P(SbeGrid), P(SbeParams), P(SbeInstance), i, P(SbeSubpulse), i,
                                  P(vp)]
    L.sbe_scan_destroy.argtypes = [vp]
    L.sbe_scan_last_error.restype = C.c_char_p
    L.sbe_scan_last_error.argtypes = [vp]
    L.sbe_scan_run.argtypes = [vp, i, i, dp, dp, P(C.c_long), P(C.c_long), dp]
    L.sbe_scan_state_size.restype = l
    L.sbe_scan_state_size.argtypes = [vp, i]
    L.sbe_scan_final_state.argtypes = [vp, i, dp]
    L.sbe_scan_tuning.argtypes = [vp, P(C.c_int), P(C.c_int), P(C.c_int), dp]
    L.sbe_observe_device_noalloc.argtypes = [vp, vp, vp, vp]
    L.sbe_observer_create.argtypes = [vp, P(SbeRegion), i, d, i, i, P(vp)]
    L.sbe_observer_destroy.argtypes = [vp]
    L.sbe_observer_reset.argtypes = [vp]
    L.sbe_observer_record.argtypes = [vp, dp, d]
    L.sbe_observer_record_device.argtypes = [vp, vp, d, vp]
    L.sbe_observer_count.argtypes = [vp, P(C.c_int)]
    L.sbe_observer_results.argtypes = [vp, P(SbeObservations)]
    L.sbe_integrate_info.argtypes = [vp, P(i)]
    L.sbe_integrate_observed.restype = l
    L.sbe_integrate_observed.argtypes = [vp, dp, i, vp, dp, dp, P(SbeSteplog), P(SbeSteplog), P(C.c_long)]
    fp = P(C.c_float)
    L.sbe_mlp_create.argtypes = [i, i, P(C.c_int), C.c_ulonglong, P(vp)]
    L.sbe_mlp_destroy.argtypes = [vp]
    L.sbe_mlp_last_error.restype = C.c_char_p
    L.sbe_mlp_last_error.argtypes = [vp]
    L.sbe_mlp_param_count.restype = l
    L.sbe_mlp_param_count.argtypes = [vp]
    L.sbe_mlp_get_params.argtypes = [vp, i, fp]
    L.sbe_mlp_set_params.argtypes = [vp, i, fp]
    L.sbe_mlp_forward.argtypes = [vp, i, fp, fp]
    L.sbe_mlp_gradient.argtypes = [vp, i, fp, fp, fp, fp]
    L.sbe_mlp_train.argtypes = [vp, i, fp, fp, i, i, C.c_float, C.c_ulonglong, fp, P(C.c_double)]
    _lib = L
    return L


def check(rc: int, ctx=None):
    if rc < 0:
        msg = load().sbe_last_error(ctx).decode()
        raise SbeError(f"sbewh error {rc}: {msg}")
    return rc
