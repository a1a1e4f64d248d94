"""The product's pulse E(t) (libsbewh_hip.so, host code; no GPU needed) on the
reference's own pulse test cases (code/tests/pulse.cpp:12-99): the same
properties the reference asserts (E(0) = E0, zero field at the window edges,
zero net field, A0), and agreement with the oracle's closed forms, for every
pulse type the product supports (sin^2, Gaussian, pure Gaussian, sin^2 ramp)
and the Pulse3D composition."""
import ctypes as C
import math

import numpy as np
import pytest
from scipy.integrate import trapezoid

from sbewh_amd import _lib, units
from sbewh_amd.ctypes_types import pulse_array
from sbewh_amd.model import Pulse
from tests.test_units_pulses import _oracle_with


def product_E(pulses, ts):
    L = _lib.load()
    arr = pulse_array(pulses)
    out = np.zeros((len(ts), 3))
    e = (C.c_double * 3)()
    for i, t in enumerate(ts):
        assert L.sbe_pulse_field(arr, len(pulses), float(t), e) == 0
        out[i] = e[:]
    return out


def product_window(pulses):
    L = _lib.load()
    arr = pulse_array(pulses)
    t0, t1 = C.c_double(), C.c_double()
    assert L.sbe_pulse_window(arr, len(pulses), C.byref(t0), C.byref(t1)) == 0
    return t0.value, t1.value


CASES = {
    # code/tests/pulse.cpp:12-24
    "pureGauss": [Pulse("pureGauss", 1.0, 5000.0, 0.0, 20.0, t_start=-5, t_end=5)],
    # code/tests/pulse.cpp:26-40
    "gauss": [Pulse("gauss", 1.0, 5000.0, 0.0, 20.0, t_start=-5, t_end=5)],
    # code/tests/pulse.cpp:42-57
    "sin2": [Pulse("sin2", 1.0, 5000.0, 0.0, 20.0, t_start=-2.5, t_end=2.5)],
    # code/tests/pulse.cpp:59-79
    "sin2Ramp4": [Pulse("sin2Ramp", 1.0, 5000.0, 0.0, 20.0, rising_cycles=2.3, cycles_on=4)],
    "sin2Ramp5": [Pulse("sin2Ramp", 1.0, 5000.0, 0.0, 20.0, rising_cycles=2.3, cycles_on=5)],
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_product_pulse_matches_oracle_and_reference_properties(name):
    pulses = CASES[name]
    o = _oracle_with(pulses)
    t0, t1 = product_window(pulses)
    assert t0 == o.t_start and t1 == o.t_end
    ts = np.linspace(t0, t1, 200001)
    E = product_E(pulses, ts)
    Eo = np.array([o.E(t) for t in ts[::97]])
    E0 = units.from_V_nm(1.0)
    assert np.abs(E[::97] - Eo).max() <= 1e-14 * E0
    assert abs(E[0, 0] / E0) < 1e-10 and abs(E[-1, 0] / E0) < 1e-10
    if name != "sin2Ramp5" and name != "pureGauss":
        assert abs(product_E(pulses, [0.0])[0, 0] / E0 - 1) < 1e-8
    # zero net field over the window (A returns to 0), except the pure Gaussian
    # whose A0 = E0 FWHM sqrt(pi / 4 ln 2) (code/tests/pulse.cpp:23)
    net = trapezoid(E[:, 0], ts)
    if name == "pureGauss":
        assert math.isclose(net, E0 * units.from_fs(20.0) * math.sqrt(math.pi / 4 / math.log(2)), rel_tol=1e-8)
    else:
        assert abs(net / E0 / (t1 - t0)) < 1e-6


def test_product_pulse3d_composition():
    # code/tests/pulse.cpp:81-99
    ts = 100.0
    p1 = Pulse("sin2", 1.0, 5000.0, 0.0, 20.0, t_start=-2.5, t_end=2.5, pol=(0, 0, 2))
    p2 = Pulse("sin2", 1.0, 2500.0, 0.0, 40.0, t_start=-2.5, t_end=2.5, pol=(0, 1, 0))
    p3 = Pulse("sin2", 1.0, 2500.0, 0.0, 60.0, t_start=-2.5, t_end=2.5, pol=(1, 0, 0), t_central_fs=-units.to_fs(ts))
    p4 = Pulse("sin2", 1.0, 5000.0, 0.0, 20.0, t_start=-2.5, t_end=2.5, pol=(1, 0, 0), t_central_fs=units.to_fs(ts))
    t = 122.0
    E = product_E([p1, p2, p3, p4], [t])[0]
    s = [product_E([p], [t])[0] for p in (p1, p2, p3, p4)]
    assert E[0] == s[2][0] + s[3][0]
    assert E[1] == s[1][1]
    assert E[2] == s[0][2]
    w = [product_window([p]) for p in (p1, p2, p3, p4)]
    assert product_window([p1, p2, p3, p4]) == (min(a for a, _ in w), max(b for _, b in w))


def test_product_pulse_rejects_invalid_input():
    L = _lib.load()
    e = (C.c_double * 3)()
    bad = pulse_array([Pulse("sin2", 1.0, 5000.0, 0.0, 20.0, pol=(0, 0, 0))])
    assert L.sbe_pulse_field(bad, 1, 0.0, e) == -1
    assert "polarisation" in L.sbe_last_error(None).decode()
    assert L.sbe_pulse_field(bad, 0, 0.0, e) == -1
