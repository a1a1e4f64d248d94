"""Units and pulses.

Units are pinned by the reference's own unit-conversion test
(code/tests/unitConversion.cpp:9-18).  The oracle's pulse functions are pinned
by the reference's pulse tests (code/tests/pulse.cpp:12-99): E(0) = E0,
vanishing field at the window edges, A = -integral of E, A0 = E0/omega, and the
Pulse3D composition rule.
"""
import math

import numpy as np
import pytest
from scipy.integrate import cumulative_trapezoid, trapezoid

from sbewh_amd import units
from sbewh_amd.model import Pulse
from oracle.oracle_py import Oracle
from tests import cases


def test_reference_unit_expectations():
    # code/tests/unitConversion.cpp:10-17
    assert abs(1 - 24.2 / (units.to_fs(1) * 1e3)) < 1e-3
    assert abs(1 - 0.52917e-10 / (units.to_A(1) * 1e-10)) < 1e-3
    assert abs(1 - 3.51e16 / units.W_CM2) < 1e-3
    assert abs(1 - 5.14e11 / (units.to_V_nm(1) * 1e9)) < 1e-3
    assert abs(1 - 27.2 / units.to_eV(1)) < 1e-3
    assert abs(1 - 137 / units.SPEED_OF_LIGHT) < 1e-3
    assert abs(units.from_K(315775.13) - 1) < 1e-3
    v = 42.0
    for f, g in ((units.from_fs, units.to_fs), (units.from_A, units.to_A), (units.from_nm, units.to_nm),
                 (units.from_V_nm, units.to_V_nm), (units.from_eV, units.to_eV), (units.from_K, units.to_K)):
        assert math.isclose(f(g(v)), v, rel_tol=1e-15)
    assert math.isclose(units.to_nm(units.from_A(v)), 0.1 * v, rel_tol=1e-14)


def _oracle_with(pulses, dim=3):
    import dataclasses
    cfg = dataclasses.replace(cases.chain(), pulses=pulses)
    return Oracle(cfg)


def _e0():
    return units.from_V_nm(1.0)


def _check_pulse(o, E0, omega, tol_int, has_center=True):
    t0, t1 = o.t_start, o.t_end
    ts = np.linspace(t0, t1, 400001)
    E = np.array([o.E(t)[0] for t in ts])
    A = np.array([o.A(t)[0] for t in ts])
    intE = cumulative_trapezoid(E, ts, initial=0)
    err = math.sqrt(trapezoid((intE + A) ** 2, ts)) / (E0 * (t1 - t0))
    assert err < tol_int
    if has_center:
        assert abs(o.E(0.0)[0] / E0 - 1) < 1e-8
    assert abs(o.E(t0)[0] / E0) < 1e-10
    assert abs(o.E(t1)[0] / E0) < 1e-10
    assert abs(trapezoid(E, ts) / E0 / (t1 - t0)) < 1e-6


def test_sin2_pulse_reference_case():
    # code/tests/pulse.cpp:42-57: E0 = 1 V/nm, FWHM = 20 fs, 5 um, window -2.5..2.5 FWHM
    o = _oracle_with([Pulse("sin2", 1.0, 5000.0, 0.0, 20.0, t_start=-2.5, t_end=2.5)])
    omega = units.omega_from_lambda_nm(5000.0)
    _check_pulse(o, _e0(), omega, 1e-5)
    # A0 = E0 / omega: with cep = pi/2 the envelope peak gives A(0) = -E0/omega exactly
    o2 = _oracle_with([Pulse("sin2", 1.0, 5000.0, math.pi / 2, 20.0, t_start=-2.5, t_end=2.5)])
    assert math.isclose(o2.A(0.0)[0], -_e0() / omega, rel_tol=1e-14)


def test_gaussian_pulse_reference_case():
    # code/tests/pulse.cpp:26-40
    o = _oracle_with([Pulse("gauss", 1.0, 5000.0, 0.0, 20.0, t_start=-5, t_end=5)])
    _check_pulse(o, _e0(), units.omega_from_lambda_nm(5000.0), 1e-5)


def test_pure_gauss_reference_case():
    # code/tests/pulse.cpp:12-24: A0 = E0 FWHM sqrt(pi / 4 ln 2)
    o = _oracle_with([Pulse("pureGauss", 1.0, 5000.0, 0.0, 20.0, t_start=-5, t_end=5)])
    E0, fw = _e0(), units.from_fs(20.0)
    t0, t1 = o.t_start, o.t_end
    ts = np.linspace(t0, t1, 400001)
    E = np.array([o.E(t)[0] for t in ts])
    A = np.array([o.A(t)[0] for t in ts])
    intE = cumulative_trapezoid(E, ts, initial=0)
    assert math.sqrt(trapezoid((intE + A) ** 2, ts)) / (E0 * (t1 - t0)) < 1e-5
    assert abs(o.E(t0)[0] / E0) < 1e-10 and abs(o.E(t1)[0] / E0) < 1e-10
    assert math.isclose(-o.A(t1)[0], E0 * fw * math.sqrt(math.pi / 4 / math.log(2)), rel_tol=1e-8)


@pytest.mark.parametrize("cycles_on", [4, 5])
def test_sin2_ramp_reference_case(cycles_on):
    # code/tests/pulse.cpp:59-79: rising cycles 2.3, 4 or 5 cycles on
    p = Pulse("sin2Ramp", 1.0, 5000.0, 0.0, 20.0, rising_cycles=2.3, cycles_on=cycles_on)
    o = _oracle_with([p])
    _check_pulse(o, _e0(), units.omega_from_lambda_nm(5000.0), 1e-4, has_center=(cycles_on == 4))


def test_pulse3d_composition():
    # code/tests/pulse.cpp:81-99: sum of sub-pulses, centres, normalised polarisations
    ts = units.from_fs(100.0 * units.FS)  # 100 a.u.
    p1 = Pulse("sin2", 1.0, 5000.0, 0.0, 20.0, t_start=-2.5, t_end=2.5, pol=(0, 0, 2))
    p2 = Pulse("sin2", 1.0, 2500.0, 0.0, 40.0, t_start=-2.5, t_end=2.5, pol=(0, 1, 0))
    p3 = Pulse("sin2", 1.0, 2500.0, 0.0, 60.0, t_start=-2.5, t_end=2.5, pol=(1, 0, 0), t_central_fs=-units.to_fs(ts))
    p4 = Pulse("sin2", 1.0, 5000.0, 0.0, 20.0, t_start=-2.5, t_end=2.5, pol=(1, 0, 0), t_central_fs=units.to_fs(ts))
    o = _oracle_with([p1, p2, p3, p4])
    singles = [_oracle_with([p]) for p in (p1, p2, p3, p4)]
    t = 122.0
    E = o.E(t)
    assert math.isclose(E[0], singles[2].E(t)[0] + singles[3].E(t)[0], rel_tol=1e-13, abs_tol=1e-300)
    assert E[1] == singles[1].E(t)[1]
    assert E[2] == singles[0].E(t)[2]
    assert o.t_start == min(s.t_start for s in singles)
    assert o.t_end == max(s.t_end for s in singles)
