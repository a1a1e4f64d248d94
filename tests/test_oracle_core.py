"""Pins of the oracle's phase sum, Jacobi, RHS, initial state and observer.

Each check is chosen so that a plausible restatement error fails it: closed
forms (n = 2 eigenvalues, exact coherent and relaxation dynamics in the
eigenbasis), independent numpy implementations (numpy.linalg.eigh, direct
brute-force sums with finite-difference dH/dk), invariants (Hermiticity,
trace, periodicity, stationarity of the ground state).
"""
import dataclasses
import math

import numpy as np
import pytest

from sbewh_amd import model, units
from oracle.oracle_py import Oracle, jacobi
from tests import cases

RNG = np.random.default_rng(12345)


# ---------------------------------------------------------------- Jacobi ----
@pytest.mark.parametrize("n", [2, 3, 5, 8, 16])
def test_jacobi_reference_properties(n):
    """code/tests/diagonalizator.cpp:53-104: U unitary to 1e-14, U A U^+ diagonal."""
    for _ in range(5):
        H = cases.random_hermitian(RNG, 1, n)[0]
        U, e = jacobi(H, 12)
        assert np.abs(U @ U.conj().T - np.eye(n)).max() < 1e-14
        D = U @ H @ U.conj().T
        assert np.abs(D - np.diag(np.diag(D))).max() < 1e-13
        assert np.allclose(np.diag(D).real, e, atol=1e-13)
        assert np.allclose(np.sort(e), np.linalg.eigvalsh(H), atol=1e-13)


def test_jacobi_2x2_closed_form():
    for _ in range(20):
        a, d = RNG.normal(size=2)
        b = complex(*RNG.normal(size=2))
        H = np.array([[a, b], [np.conj(b), d]])
        U, e = jacobi(H, 1)
        lam = (a + d) / 2 + np.array([-1, 1]) * math.sqrt(((a - d) / 2) ** 2 + abs(b) ** 2)
        assert np.allclose(np.sort(e), lam, atol=1e-14)
        # one rotation annihilates the off-diagonal element exactly
        D = U @ H @ U.conj().T
        assert abs(D[0, 1]) < 1e-15


def test_jacobi_sweep_count_matters():
    """fixed sweep count: one sweep leaves an 8x8 matrix visibly off-diagonal,
    convergence is quadratic over the sweeps and complete by eight."""
    H = cases.random_hermitian(RNG, 1, 8)[0]
    off = lambda U: np.abs((U @ H @ U.conj().T)[~np.eye(8, dtype=bool)]).max()
    o1, o5, o8 = (off(jacobi(H, s)[0]) for s in (1, 5, 8))
    assert o1 > 1e-3
    assert o5 < 1e-6 * o1
    assert o8 < 1e-13


# ------------------------------------------------------------- phase sum ----
@pytest.mark.parametrize("name", ["chain", "square"])
def test_phase_sum_direct_separable_numpy(name):
    cfg = cases.chain() if name == "chain" else cases.small_square()
    od = Oracle(cfg, "direct")
    os_ = Oracle(cfg, "separable")
    A, E = (0.13, -0.04, 0.0), (0.002, -0.003, 0.0)
    md, ms = od.hk_mesh(A, E), os_.hk_mesh(A, E)
    assert np.abs(md - ms).max() < 1e-13 * np.abs(md).max()
    # brute force in numpy for a few k: k_frac = grid + frac(A)
    lat = cfg.model.lattice
    off = (lat @ np.array(A)) / (2 * np.pi)
    for k in (0, 1, 7, od.num_k - 1):
        kf = od.kfrac(k) + off
        H0 = cfg.model.hk(kf)
        HE = H0 + sum(E[d] * np.tensordot(np.exp(2j * np.pi * (cfg.model.cells @ kf)), cfg.model.D[:, d], 1)
                      for d in range(cfg.dim))
        assert np.abs(md[k, 0] - H0).max() < 1e-13 * np.abs(H0).max()
        assert np.abs(md[k, 1] - HE).max() < 1e-13 * np.abs(HE).max()
    # Hermiticity and periodicity
    assert np.abs(md - np.conj(np.swapaxes(md, -1, -2))).max() < 1e-14 * np.abs(md).max()
    H0a, _ = od.hk_point([0.1, 0.2, 0.0])
    H0b, _ = od.hk_point([1.1, -0.8, 0.0])
    assert np.abs(H0a - H0b).max() < 1e-12 * np.abs(H0a).max()


def test_k_order_is_fourier_block_order():
    """code/src/WhFourierTransform.cpp:137-149 and :303-312: k index
    c + Nf2 (b + Nf1 (a + Nf0 blk)) has fractional coordinate blk/Nk + a/Nf."""
    cfg = cases.chain()
    o = Oracle(cfg)
    Nk, Nf = 256, 16
    assert o.nf == (16, 1, 1)
    for k in range(Nk):
        blk, a = divmod(k, Nf)
        assert math.isclose(o.kfrac(k)[0], blk / Nk + a / Nf, rel_tol=0, abs_tol=1e-15)
    sq = Oracle(cases.small_square(Nk=(30, 30, 1)))
    assert sq.nf == (15, 15, 1)  # extent 13 -> 15 (5-smooth), code/src/Parameter.cpp:4-32


def test_full_operators_finite_difference():
    """dH0/dk_d = sum_R i R_cart,d H(R) e^{...} equals the Cartesian k-derivative of H0(k)."""
    cfg = cases.small_square()
    o = Oracle(cfg)
    lat = cfg.model.lattice
    kf = np.array([0.17, 0.41, 0.0])
    H0, D, dH = o.hk_full_point(kf)
    assert np.abs(H0 - cfg.model.hk(kf)).max() < 1e-13 * np.abs(H0).max()
    h = 1e-5
    for d in range(2):
        dk = np.zeros(3)
        dk[d] = h
        dfrac = lat @ dk / (2 * np.pi)
        fd = (cfg.model.hk(kf + dfrac) - cfg.model.hk(kf - dfrac)) / (2 * h)
        assert np.abs(dH[d] - fd).max() < 1e-7 * np.abs(dH[d]).max()
        Dk = np.tensordot(np.exp(2j * np.pi * (cfg.model.cells @ kf)), cfg.model.D[:, d], 1)
        assert np.abs(D[d] - Dk).max() < 1e-13 * np.abs(Dk).max()


# ------------------------------------------------------------------- RHS ----
def numpy_rhs(cfg, o, state, t):
    """independent restatement with numpy.linalg.eigh for the eigenbasis."""
    n, nk, P = o.n, o.num_k, o.P
    E = o.E(t)
    A = np.zeros(3)
    A[:cfg.dim] = state[:cfg.dim].real
    off = cfg.model.lattice @ (A) / (2 * np.pi)
    out = np.zeros_like(state)
    out[:cfg.dim] = -E[:cfg.dim]
    F2 = 1 / units.from_fs(cfg.T2_fs) if cfg.T2_fs > 0 else -1
    w = units.from_eV(cfg.soothing_width_eV)
    for k in range(nk):
        kf = o.kfrac(k) + off
        ph = np.exp(2j * np.pi * (cfg.model.cells @ kf))
        H0 = np.tensordot(ph, cfg.model.H, 1)
        HE = H0 + sum(E[d] * np.tensordot(ph, cfg.model.D[:, d], 1) for d in range(cfg.dim))
        rho = state[8 + P * k: 8 + P * k + n * n].reshape(n, n)
        d = -1j * (HE @ rho - rho @ HE)
        if F2 > 0:
            ev, V = np.linalg.eigh(H0)
            rH = V.conj().T @ rho @ V
            dE = (ev[:, None] - ev[None, :]) / w
            g = (1 - np.exp(-dE ** 2)) if w >= 0 else (1 - np.eye(n))
            d += V @ (-F2 * g * rH) @ V.conj().T
        out[8 + P * k: 8 + P * k + n * n] = d.reshape(-1)
    return out


@pytest.mark.parametrize("T2,w", [(5.0, 0.025), (5.0, -1.0), (-1.0, 0.025)])
def test_rhs_against_numpy(T2, w):
    cfg = dataclasses.replace(cases.chain(sweeps=10), T2_fs=T2, soothing_width_eV=w, Nk=(32, 1, 1),
                              Nf=(16, 1, 1))
    o = Oracle(cfg)
    st = o.initial_state(o.fermi_level(1))
    x = cases.perturbed_state(st, o.n, o.P, o.num_k, seed=2, amp=0.05)
    t = 0.4 * o.t_start
    ref = numpy_rhs(cfg, o, x, t)
    got = o.rhs(x, t)
    assert np.abs(got - ref).max() < 1e-12 * np.abs(ref).max()


def test_rhs_invariants_and_ground_state():
    cfg = cases.small_square(Nk=(15, 15, 1))
    o = Oracle(cfg)
    mu = o.fermi_level(cfg.occupied_bands)
    st = o.initial_state(mu)
    # stationary ground state at E = 0 (pulse edge), A = 0
    r0 = o.rhs(st, o.t_start)
    assert np.abs(r0).max() < 1e-13
    x = cases.perturbed_state(st, o.n, o.P, o.num_k, seed=9)
    r = o.rhs(x, 0.3 * o.t_start)
    d = o.rho(r)
    assert np.abs(d - np.conj(np.swapaxes(d, 1, 2))).max() < 1e-14 * np.abs(d).max()
    assert np.abs(np.trace(d, axis1=1, axis2=2)).max() < 1e-14 * np.abs(d).max()
    assert np.allclose(r[:2].real, -o.E(0.3 * o.t_start)[:2]) and np.all(r[2:8] == 0)
    # pads of the n*n = 64 matrix do not exist; for n = 2 they stay zero
    c = Oracle(cases.chain())
    xc = cases.perturbed_state(c.initial_state(c.fermi_level(1)), c.n, c.P, c.num_k)
    rc = c.rhs(xc, 0.0)
    assert np.all(rc[8:].reshape(c.num_k, c.P)[:, 4:] == 0)


def test_dephasing_closed_form_in_eigenbasis():
    """E = 0: in the eigenbasis of H0, d rho_ab/dt = (-i (e_a - e_b) - F2 g_ab) rho_ab."""
    cfg = dataclasses.replace(cases.chain(sweeps=10), Nk=(16, 1, 1), Nf=(16, 1, 1))
    o = Oracle(cfg)
    st = o.initial_state(o.fermi_level(1))
    x = cases.perturbed_state(st, o.n, o.P, o.num_k, seed=4, amp=0.1)
    x[0] = 0.0  # A = 0 so that H0 is the grid Hamiltonian
    t = o.t_start  # E = 0
    r = o.rhs(x, t)
    F2 = 1 / units.from_fs(cfg.T2_fs)
    w = units.from_eV(cfg.soothing_width_eV)
    for k in range(o.num_k):
        H0 = cfg.model.hk(o.kfrac(k))
        ev, V = np.linalg.eigh(H0)
        rho = o.rho(x)[k]
        drho = o.rho(r)[k]
        rh, drh = V.conj().T @ rho @ V, V.conj().T @ drho @ V
        g = 1 - np.exp(-((ev[:, None] - ev[None, :]) / w) ** 2)
        expect = (-1j * (ev[:, None] - ev[None, :]) - F2 * g) * rh
        assert np.abs(drh - expect).max() < 1e-12 * np.abs(expect).max()


def test_initial_state_and_fermi_level():
    cfg = cases.small_square(Nk=(15, 15, 1))
    o = Oracle(cfg)
    mu = o.fermi_level(4)
    ev = np.sort(np.linalg.eigvalsh(np.array([cfg.model.hk(o.kfrac(k)) for k in range(o.num_k)])).ravel())
    nk = o.num_k
    assert math.isclose(mu, 0.5 * (ev[4 * nk - 1] + ev[4 * nk]), rel_tol=1e-13)
    st = o.initial_state(mu)
    rho = o.rho(st)
    kT = units.from_K(cfg.temperature_K)
    for k in (0, 11, nk - 1):
        e, V = np.linalg.eigh(cfg.model.hk(o.kfrac(k)))
        f = 1 / (1 + np.exp((e - mu) / kT))
        assert np.abs(rho[k] - (V * f) @ V.conj().T).max() < 1e-13
    assert np.allclose(np.trace(rho, axis1=1, axis2=2).real, 4, atol=1e-12)
    # occupiedBelowBandGap selects the mid-gap of the single gap
    mu2 = o.fermi_level(0, below_gap=True)
    e4 = ev.reshape(-1)
    assert ev[4 * nk - 1] < mu2 < ev[4 * nk]


# -------------------------------------------------------------- observer ----
def test_current_against_numpy():
    cfg = cases.small_square(Nk=(15, 15, 1))
    o = Oracle(cfg)
    st = o.initial_state(o.fermi_level(4))
    x = cases.perturbed_state(st, o.n, o.P, o.num_k, seed=7)
    j = o.observe_j(x)
    A = x[:3].real.copy()
    A[cfg.dim:] = 0
    off = cfg.model.lattice @ A / (2 * np.pi)
    lat = cfg.model.lattice
    pos = cfg.model.cells @ lat
    acc = np.zeros(3)
    for k in range(o.num_k):
        kf = o.kfrac(k) + off
        ph = np.exp(2j * np.pi * (cfg.model.cells @ kf))
        H0 = np.tensordot(ph, cfg.model.H, 1)
        rho = o.rho(x)[k]
        T = H0 @ rho - rho @ H0
        for d in range(cfg.dim):
            D = np.tensordot(ph, cfg.model.D[:, d], 1)
            dH = np.tensordot(ph * 1j * pos[:, d], cfg.model.H, 1)
            acc[d] += (1j * np.trace(D @ T) - np.trace(dH @ rho)).real
    assert np.abs(j - acc / o.num_k).max() < 1e-12 * np.abs(acc / o.num_k).max()
