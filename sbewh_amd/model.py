"""Seeded synthetic tight-binding models and the benchmark configurations.

A model is what the reference reads from a Wannier90 ``_tb.dat`` file after
degeneracy division, symmetrisation and dimensional projection
(code/src/io/TightBindingParameter.cpp:287-329): a list of lattice cells R with
Hamiltonian blocks H(R) and position blocks r_d(R), plus the lattice vectors,
all in atomic units.  ``TBModel`` holds exactly that; ``sbewh_amd.w90`` reads
it from Wannier90 ``_tb.dat`` / ``_wsvec.dat`` files and writes it in the
``_tb.dat`` layout.

The configurations mirror the workloads the benchmark is quoted on:

* ``chain1d``  1-D two-band chain, R = -4..4, N_k = 256, sin^2 3.2 um pulse
* ``square2d`` 2-D 8-band model on a square lattice, |R_i| <= 6, 150 x 150 k
* ``cubic3d``  3-D 16-band model on a cubic lattice, R in [-4,4]^3, 48^3 k

Every random draw comes from ``numpy.random.default_rng(seed)`` with the seed
recorded in the config, so a config name plus its seed fixes the model
bit-for-bit.
"""
from __future__ import annotations

import dataclasses
import functools
import math

import numpy as np

from . import units


@dataclasses.dataclass
class TBModel:
    cells: np.ndarray      # (N_R, 3) int32 lattice-vector indices
    H: np.ndarray          # (N_R, n, n) complex128, hartree
    D: np.ndarray          # (N_R, 3, n, n) complex128, bohr
    lattice: np.ndarray    # (3, 3) float64, row i = a_i in bohr

    @property
    def num_wann(self) -> int:
        return self.H.shape[1]

    @property
    def num_cells(self) -> int:
        return self.H.shape[0]

    def hermiticity_error(self) -> float:
        """max |X(-R) - X(R)^dagger| over H and D (0 for a valid model)."""
        idx = {tuple(c): i for i, c in enumerate(self.cells.tolist())}
        err = 0.0
        for i, c in enumerate(self.cells.tolist()):
            j = idx.get(tuple(-x for x in c))
            if j is None:
                return math.inf
            err = max(err, np.abs(self.H[j] - self.H[i].conj().T).max())
            for d in range(3):
                err = max(err, np.abs(self.D[j, d] - self.D[i, d].conj().T).max())
        return float(err)

    def hk(self, kfrac: np.ndarray) -> np.ndarray:
        """H(k) = sum_R exp(i 2 pi k.R) H(R) for fractional k of shape (..., 3)."""
        ph = np.exp(2j * np.pi * (np.asarray(kfrac) @ self.cells.T.astype(float)))
        return np.tensordot(ph, self.H, axes=([-1], [0]))


@dataclasses.dataclass
class Pulse:
    """One sub-pulse in reference input units (code/src/io/PulseParameter.cpp:121-162)."""
    type: str = "sin2"          # sin2 | gauss | pureGauss | sin2Ramp
    Emax_V_nm: float = 1.0
    lambda_nm: float = 3000.0
    cep: float = 0.0
    fwhm_fs: float = 10.0
    t_central_fs: float = 0.0
    t_start: float = -1.0       # in units of FWHM
    t_end: float = 1.0
    pol: tuple = (1.0, 0.0, 0.0)
    cycles_on: int = 3
    rising_cycles: float = 1.0


PULSE_TYPES = {"sin2": 0, "gauss": 1, "pureGauss": 2, "sin2Ramp": 3}


def sin2_fwhm_for_cycles(lambda_nm: float, cycles: float) -> float:
    """FWHM parameter (fs) of a sin^2 pulse whose window +-FWHM holds `cycles`
    optical cycles: 2 FWHM = cycles * T."""
    period_fs = units.to_fs(2 * math.pi / units.omega_from_lambda_nm(lambda_nm))
    return 0.5 * cycles * period_fs


def elliptic_sin2(E0_V_nm: float, ellipticity: float, lambda_nm: float, cycles: float,
                  cep: float = 0.0) -> list:
    """Two orthogonal sin^2 sub-pulses with field ratio `ellipticity` and a quarter
    period phase lag, the convention of
    examples/08_ellipticityScan/generateEllipticitiyRuns.py:18-23
    (E1 = E0/sqrt(1+eps^2), E2 = eps E1, cep2 = cep1 + pi/2)."""
    fw = sin2_fwhm_for_cycles(lambda_nm, cycles)
    e1 = E0_V_nm / math.sqrt(1 + ellipticity ** 2)
    out = [Pulse("sin2", e1, lambda_nm, cep, fw, pol=(1.0, 0.0, 0.0))]
    if ellipticity > 0:
        out.append(Pulse("sin2", ellipticity * e1, lambda_nm, cep + math.pi / 2, fw, pol=(0.0, 1.0, 0.0)))
    return out


@dataclasses.dataclass
class Config:
    name: str
    model: TBModel
    dim: int
    Nk: tuple
    Nf: tuple
    pulses: list
    T2_fs: float
    soothing_width_eV: float = 0.025
    temperature_K: float = 300.0
    eps_abs: float = 1e-12
    eps_rel: float = 1e-8
    sweeps: int = 5
    occupied_bands: int = 1
    output_count: int = 1024
    k_global_shift: tuple = (0.0, 0.0, 0.0)
    cold_jacobi: bool = False   # start each RHS's Jacobi from the identity
    gpu_layout: bool = False    # state vectors in the reference's GPU layout (rho_k unpadded)
    device_controller: bool = False  # 1-D, n <= 8: the whole integration in one launch (step control on the GPU)

    @property
    def num_k(self) -> int:
        return int(np.prod(self.Nk))

    def with_grid(self, Nk: tuple, Nf: tuple | None = None) -> "Config":
        return dataclasses.replace(self, Nk=tuple(Nk), Nf=tuple(Nf if Nf is not None else (0, 0, 0)))


def _herm_random(rng, shape, sigma):
    """complex Gaussian entries with E|z|^2 = sigma^2."""
    return (rng.normal(size=shape) + 1j * rng.normal(size=shape)) * (sigma / math.sqrt(2))


def _cells_box(ext: tuple) -> list:
    """integer cells with |R_d| <= ext_d, in lexicographic order."""
    r = [range(-e, e + 1) for e in ext]
    return [(a, b, c) for a in r[0] for b in r[1] for c in r[2]]


def _positive(c) -> bool:
    """the half-space whose blocks are drawn; the mirrored half is the adjoint
    (the symmetrisation convention of code/src/util/w90util.cpp:353-389)."""
    return c[0] > 0 or (c[0] == 0 and c[1] > 0) or (c[0] == 0 and c[1] == 0 and c[2] > 0)


def chain1d_model(seed: int = 1001) -> TBModel:
    """1-D two-band chain: onsite diag(0, 3.0 eV); hoppings of modulus
    0.5 eV exp(-|R|/1.5) with uniform random phases; position operator with
    zero diagonal and off-diagonal modulus 2 A exp(-|R|/1.5), random phases."""
    rng = np.random.default_rng(seed)
    a = units.from_A(3.0)
    cells = [(R, 0, 0) for R in range(-4, 5)]
    n = 2
    idx = {c: i for i, c in enumerate(cells)}
    H = np.zeros((len(cells), n, n), complex)
    D = np.zeros((len(cells), 3, n, n), complex)
    for c in cells:
        if not (_positive(c) or c == (0, 0, 0)):
            continue
        R = abs(c[0])
        amp_t = units.from_eV(0.5) * math.exp(-R / 1.5)
        amp_r = units.from_A(2.0) * math.exp(-R / 1.5)
        ph = rng.uniform(0, 2 * math.pi, size=(n, n))
        phr = rng.uniform(0, 2 * math.pi, size=(n, n))
        h = amp_t * np.exp(1j * ph)
        r = np.zeros((n, n), complex)
        r[0, 1] = amp_r * np.exp(1j * phr[0, 1])
        r[1, 0] = amp_r * np.exp(1j * phr[1, 0])
        if c == (0, 0, 0):
            h = np.diag([0.0, units.from_eV(3.0)]).astype(complex)
            h[0, 1] = amp_t * np.exp(1j * ph[0, 1])
            h[1, 0] = np.conj(h[0, 1])
            r[1, 0] = np.conj(r[0, 1])
            H[idx[c]] = h
            D[idx[c], 0] = r
        else:
            H[idx[c]] = h
            D[idx[c], 0] = r
            m = idx[(-c[0], -c[1], -c[2])]
            H[m] = h.conj().T
            D[m, 0] = r.conj().T
    lat = np.eye(3) * a
    return TBModel(np.array(cells, np.int32), H, D, lat)


def _gaussian_model(seed, n, ext, lattice_A, sigma_eV, decay, dim):
    rng = np.random.default_rng(seed)
    cells = _cells_box(ext)
    idx = {c: i for i, c in enumerate(cells)}
    H = np.zeros((len(cells), n, n), complex)
    D = np.zeros((len(cells), 3, n, n), complex)
    sig = units.from_eV(sigma_eV)
    sigr = units.from_A(1.0)
    for c in cells:
        if not (_positive(c) or c == (0, 0, 0)):
            continue
        Rn = math.sqrt(sum(x * x for x in c))
        f = math.exp(-Rn / decay)
        h = _herm_random(rng, (n, n), sig * f)
        r = _herm_random(rng, (3, n, n), sigr * f)
        r[dim:] = 0
        if c == (0, 0, 0):
            h = 0.5 * (h + h.conj().T)
            r = 0.5 * (r + np.conj(np.swapaxes(r, -1, -2)))
            H[idx[c]] = h
            D[idx[c]] = r
        else:
            H[idx[c]] = h
            D[idx[c]] = r
            m = idx[tuple(-x for x in c)]
            H[m] = h.conj().T
            D[m] = np.conj(np.swapaxes(r, -1, -2))
    lat = np.eye(3) * units.from_A(lattice_A)
    return TBModel(np.array(cells, np.int32), H, D, lat)


def _grid_kfrac(Nk, dim):
    axes = [np.arange(Nk[d]) / Nk[d] if d < dim else np.zeros(1) for d in range(3)]
    g = np.stack(np.meshgrid(*axes, indexing="ij"), -1).reshape(-1, 3)
    return g


def _set_gap(model: TBModel, n_occ: int, gap_eV: float, Nk, dim) -> TBModel:
    """Add an onsite split diag(-s/2 (occupied), +s/2 (empty)) with s chosen by
    bisection so that min_k e_{n_occ}(k) - max_k e_{n_occ-1}(k) = gap_eV on the
    k-grid Nk."""
    n = model.num_wann
    origin = [i for i, c in enumerate(model.cells.tolist()) if c == [0, 0, 0]][0]
    kf = _grid_kfrac(Nk, dim)
    Hk0 = model.hk(kf)
    split = np.array([-0.5] * n_occ + [0.5] * (n - n_occ))

    def gap(s):
        ev = np.linalg.eigvalsh(Hk0 + np.diag(split * s))
        return ev[:, n_occ].min() - ev[:, n_occ - 1].max()

    from scipy.optimize import brentq

    target = units.from_eV(gap_eV)
    lo, hi = 0.0, units.from_eV(1.0)
    while gap(hi) < target:
        lo, hi = hi, 2 * hi
    s = brentq(lambda v: gap(v) - target, lo, hi, xtol=1e-14, rtol=1e-13)
    H = model.H.copy()
    H[origin] += np.diag(split * s)
    return dataclasses.replace(model, H=H)


def square2d_model(seed: int = 1002, gap_Nk=(150, 150, 1)) -> TBModel:
    """2-D 8-band model, square lattice a = 3 A, |R_i| <= 6 (169 cells), complex
    Gaussian blocks with sigma = 0.8 eV exp(-|R|/2) (H(-R) = H(R)^dagger), position
    blocks with 1 A exp(-|R|/2); onsite split for a 1.5 eV gap between the lower 4
    and upper 4 bands on the 150 x 150 grid."""
    m = _gaussian_model(seed, 8, (6, 6, 0), 3.0, 0.8, 2.0, 2)
    return _set_gap(m, 4, 1.5, gap_Nk, 2)


def cubic3d_model(seed: int = 1003, gap_Nk=(24, 24, 24)) -> TBModel:
    """3-D 16-band model, cubic lattice a = 4 A, R in [-4,4]^3 (729 cells), blocks
    with sigma = 0.8 eV exp(-|R|/1.5); onsite split for a 1.5 eV gap between the
    lower 4 and upper 12 bands on a 24^3 grid."""
    m = _gaussian_model(seed, 16, (4, 4, 4), 4.0, 0.8, 1.5, 3)
    return _set_gap(m, 4, 1.5, gap_Nk, 3)


@functools.lru_cache(maxsize=None)
def config(name: str) -> Config:
    if name == "chain1d":
        lam = 3200.0
        pul = [Pulse("sin2", 0.5, lam, 0.0, sin2_fwhm_for_cycles(lam, 8))]
        return Config("chain1d", chain1d_model(), 1, (256, 1, 1), (16, 1, 1), pul, T2_fs=5.0,
                      occupied_bands=1)
    if name == "square2d":
        pul = elliptic_sin2(1.0, 0.3, 3000.0, 8)
        return Config("square2d", square2d_model(), 2, (150, 150, 1), (15, 15, 1), pul, T2_fs=10.0,
                      occupied_bands=4)
    if name == "cubic3d":
        lam = 2000.0
        pul = [Pulse("sin2", 0.8, lam, 0.0, sin2_fwhm_for_cycles(lam, 10))]
        return Config("cubic3d", cubic3d_model(), 3, (48, 48, 48), (12, 12, 12), pul, T2_fs=8.0,
                      occupied_bands=4, output_count=801)
    raise KeyError(name)


CONFIG_NAMES = ("chain1d", "square2d", "cubic3d")
