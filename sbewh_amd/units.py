"""Hartree atomic-unit conversions.

The constants are the ones the reference uses (code/src/util/unitConversion.h:4-77);
tests/test_units.py checks them against the reference's own unit-conversion
test values (code/tests/unitConversion.cpp).
"""

FS = 2.4188843266049513e-2        # 1 a.u. of time in fs
ANGSTROM = 5.291772109060855e-1   # 1 bohr in angstrom
NM = 5.291772109060855e-2         # 1 bohr in nm
V_NM = 514.22067475617267         # 1 a.u. of field in V/nm
EV = 27.211386245771678           # 1 hartree in eV
KELVIN = 315775.0248015561        # 1 hartree in K
W_CM2 = 3.50944758e16             # intensity of a unit field in W/cm^2
SPEED_OF_LIGHT = 137.035999177    # c in a.u.


def from_fs(v):
    return v / FS


def to_fs(v):
    return v * FS


def from_A(v):
    return v / ANGSTROM


def to_A(v):
    return v * ANGSTROM


def from_nm(v):
    return v / NM


def to_nm(v):
    return v * NM


def from_um(v):
    return v / 5.291772109060855e-5


def from_V_nm(v):
    return v / V_NM


def to_V_nm(v):
    return v * V_NM


def from_MV_cm(v):
    """1 MV/cm = 0.1 V/nm."""
    return from_V_nm(0.1 * v)


def from_eV(v):
    return v / EV


def to_eV(v):
    return v * EV


def from_K(v):
    return v / KELVIN


def to_K(v):
    return v * KELVIN


def omega_from_lambda_nm(lam_nm):
    """Carrier frequency of a wavelength in nm (code/src/io/PulseParameter.cpp:135)."""
    import math
    return 2 * math.pi * SPEED_OF_LIGHT / from_nm(lam_nm)
