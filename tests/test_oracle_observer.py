"""Pins of the oracle's full observer (regions, occupations, density matrix;
code/src/ObserverCPU.cpp:137-260,293-308, code/src/ObserverBase.cpp:3-17) by
closed forms and invariants, not by restating its formulas:

* rho_k = f(H0(k + A)) (Fermi-Dirac of the Hamiltonian at k + A, built here
  with numpy.linalg.eigh): the unsmeared Hamiltonian occupations are f(e_b(k))
  in ascending band order, their min / max over k are those of f(e_b(k)),
  and the region sums of a narrow window are those of its centre point;
* smearing: occupations of degenerate clusters far apart average within the
  cluster; a huge width averages over all bands, a tiny one changes nothing;
* region weights: a Gaussian that falls to 1/4 at fwhm / 2 (the reference's
  8 ln 2 / fwhm^2 exponent), whose sum
  over a fine 1-D grid is the integral N fwhm sqrt(pi / (8 ln 2)), a window is
  periodic in k, the whole zone weighs N_k, and a fixed-frame window centred
  at c + A_frac is the moving-frame window centred at c;
* the whole-zone region's current is the oracle's j (pinned by the
  A-derivative identities), its density matrix the k-average of rho_k, its
  Wannier occupations the k-average of diag rho_k.
"""
import dataclasses
import math

import numpy as np
import pytest

from oracle.oracle_py import Oracle
from tests import cases

ZONE = ((0.0, 0.0, 0.0), math.inf, 1)


def fd(e, mu, kT):
    return 1.0 / (1.0 + np.exp((e - mu) / kT))


def thermal_state(o, cfg, A, mu, kT):
    """state whose rho_k = f(H0(k + A)) with the special slots holding A"""
    st = o.new_state()
    st[:3] = A
    lat = np.asarray(cfg.model.lattice, float)
    A_frac = lat @ np.asarray(A, float) / (2 * np.pi)
    n = o.n
    eps = np.zeros((o.num_k, n))
    for k in range(o.num_k):
        kf = o.kfrac(k) + A_frac
        H0, _ = o.hk_point(kf)
        e, v = np.linalg.eigh(H0)
        rho = (v * fd(e, mu, kT)) @ v.conj().T
        st[8 + k * o.P: 8 + k * o.P + n * n] = rho.reshape(-1)
        eps[k] = e
    return st, eps


@pytest.mark.parametrize("name", ["chain", "square"])
def test_thermal_occupations_are_fermi_dirac(name):
    cfg = {"chain": cases.chain(sweeps=10), "square": cases.small_square(sweeps=10)}[name]
    o = Oracle(cfg)
    A = np.array([0.13, -0.07, 0.0]) if cfg.dim == 2 else np.array([0.21, 0.0, 0.0])
    e_all = []
    for k in range(o.num_k):
        e_all.append(np.linalg.eigvalsh(o.hk_point(o.kfrac(k))[0]))
    e_all = np.sort(np.concatenate(e_all))
    mu = 0.5 * (e_all[o.num_k * cfg.occupied_bands - 1] + e_all[o.num_k * cfg.occupied_bands])
    kT = 0.05   # hartree: occupations well inside (0, 1)
    st, eps = thermal_state(o, cfg, A, mu, kT)
    f = fd(eps, mu, kT)
    r = o.observe_full(st, [ZONE], -1.0, density=True)
    assert np.abs(r["occ_ham"][0] - f.mean(0)).max() < 1e-12
    assert np.abs(r["occ_min"] - f.min(0)).max() < 1e-12
    assert np.abs(r["occ_max"] - f.max(0)).max() < 1e-12
    rho = o.rho(st)
    assert r["weight_sum"][0] == o.num_k
    assert np.abs(r["occ_wan"][0] - np.einsum("kii->i", rho).real / o.num_k).max() < 1e-14
    assert np.abs(r["wan"][0] - rho.mean(0)).max() < 1e-14
    xp = cases.perturbed_state(st, o.n, o.P, o.num_k, seed=1)
    jz, jo = o.observe_full(xp, [ZONE], -1.0)["j"][0], o.observe_j(xp)
    assert np.abs(jz - jo).max() < 1e-13 * np.abs(jo).max()
    # a narrow fixed-frame window centred on grid point k1 + A_frac sees only k1
    k1 = o.num_k // 3
    lat = np.asarray(cfg.model.lattice, float)
    A_frac = lat @ A / (2 * np.pi)
    c = o.kfrac(k1) + A_frac
    c = c - np.floor(c)
    narrow = ((c[0], c[1], c[2]), 1e-3, 0)
    rn = o.observe_full(st, [narrow], -1.0)
    assert abs(rn["weight_sum"][0] - 1.0) < 1e-12
    assert np.abs(rn["occ_ham"][0] - f[k1]).max() < 1e-12


def test_smearing_limits():
    cfg = cases.small_square(sweeps=10)
    o = Oracle(cfg)
    mu = 0.0
    kT = 0.08
    st, eps = thermal_state(o, cfg, np.zeros(3), mu, kT)
    f = fd(eps, mu, kT)
    raw = o.observe_full(st, [ZONE], -1.0)
    tiny = o.observe_full(st, [ZONE], 1e-9)
    assert np.abs(tiny["occ_ham"] - raw["occ_ham"]).max() < 1e-14
    huge = o.observe_full(st, [ZONE], 1e6)
    # every band: the mean over bands of the raw occupations, k-averaged
    assert np.abs(huge["occ_ham"][0] - f.mean()).max() < 1e-9


def test_smearing_two_clusters_closed_form():
    """a hand-made 4-band model with flat bands at -5, -5 + d, 5, 5 + d
    (d << width << 10): occupations average within each cluster"""
    base = cases.chain(sweeps=10)
    n = 4
    cells = np.array([[0, 0, 0]], np.int32)
    H = np.zeros((1, n, n), complex)
    rng = np.random.default_rng(7)
    q, _ = np.linalg.qr(rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)))
    levels = np.array([-5.0, -5.0 + 1e-4, 5.0, 5.0 + 1e-4]) * 0.01
    H[0] = (q * levels) @ q.conj().T
    D = np.zeros((1, 3, n, n), complex)
    model = dataclasses.replace(base.model, cells=cells, H=H, D=D)
    cfg = dataclasses.replace(base, model=model, Nk=(4, 1, 1), Nf=(1, 1, 1), occupied_bands=2)
    o = Oracle(cfg)
    st = o.new_state()
    occ = np.array([0.9, 0.3, 0.2, 0.6])
    rho = (q * occ) @ q.conj().T
    for k in range(o.num_k):
        st[8 + k * o.P: 8 + k * o.P + n * n] = rho.reshape(-1)
    r = o.observe_full(st, [ZONE], 0.01 * 0.3)
    assert np.abs(r["occ_ham"][0] - np.array([0.6, 0.6, 0.4, 0.4])).max() < 1e-6
    raw = o.observe_full(st, [ZONE], -1.0)
    assert np.abs(raw["occ_ham"][0] - occ).max() < 1e-12


def indicator_state(o, ks):
    """rho_k = diag(1, 0, ...) at the listed k-points, 0 elsewhere: occ_wan[0]
    times the weight sum is the sum of the region weights at those points"""
    st = o.new_state()
    for k in ks:
        st[8 + k * o.P] = 1.0
    return st


def test_region_weight_is_a_gaussian_of_given_fwhm():
    cfg = cases.chain(sweeps=5)
    o = Oracle(cfg)
    N = o.num_k
    kg = np.array([o.kfrac(k)[0] for k in range(N)])
    k0 = int(np.flatnonzero(np.isclose(kg, 40 / N))[0])
    kh = int(np.flatnonzero(np.isclose(kg, 48 / N))[0])
    fwhm = 16 / N
    reg = [((40 / N, 0, 0), fwhm, 1)]
    r0 = o.observe_full(indicator_state(o, [k0]), reg, -1.0)
    rh = o.observe_full(indicator_state(o, [kh]), reg, -1.0)
    w0 = r0["occ_wan"][0][0] * r0["weight_sum"][0]
    wh = rh["occ_wan"][0][0] * rh["weight_sum"][0]
    # the reference's width parameter enters as 8 ln 2 / fwhm^2
    # (code/src/ObserverBase.cpp:12): the weight falls to 1/4 at fwhm / 2
    assert abs(w0 - 1.0) < 1e-14
    assert abs(wh - 0.25) < 1e-14
    # sum over the grid = N * fwhm * sqrt(pi / (8 ln 2)) (Gaussian integral)
    ws = r0["weight_sum"][0]
    assert abs(ws - N * fwhm * math.sqrt(math.pi / (8 * math.log(2)))) < 1e-9 * ws
    # periodic in k
    r1 = o.observe_full(indicator_state(o, [k0]), [((40 / N + 1.0, 0, 0), fwhm, 1)], -1.0)
    assert abs(r1["weight_sum"][0] - ws) < 1e-12 * ws


def test_fixed_and_moving_frames():
    cfg = cases.small_square(sweeps=8)
    o = Oracle(cfg)
    st = cases.perturbed_state(o.initial_state(o.fermi_level(4)), o.n, o.P, o.num_k, seed=2)
    A = np.array([0.37, -0.21, 0.0])   # perturbed_state's special slots
    assert np.allclose(st[:3].real, A)
    lat = np.asarray(cfg.model.lattice, float)
    A_frac = lat @ A / (2 * np.pi)
    c = np.array([0.2, 0.7, 0.0])
    mv = o.observe_full(st, [(tuple(c), 0.15, 1)], 0.001)
    fx = o.observe_full(st, [(tuple(c + A_frac), 0.15, 0)], 0.001)
    for key in ("weight_sum", "j", "occ_ham", "occ_wan"):
        assert np.abs(mv[key] - fx[key]).max() < 1e-10 * max(1.0, np.abs(fx[key]).max()), key
    # with A != 0 the two frames differ for the same centre
    fx0 = o.observe_full(st, [(tuple(c), 0.15, 0)], 0.001)
    assert np.abs(fx0["occ_ham"][0] - mv["occ_ham"][0]).max() > 1e-2
