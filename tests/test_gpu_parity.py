"""Parity of the HIP path (through the C ABI) with the CPU oracle.

Tolerances: FP64 throughout.  Single kernels: <= 1e-12 relative.  Whole
propagations replaying the oracle's accepted step sequence: j(t) and the final
density matrices <= 1e-10 relative (the target stated for this path).
"""
import dataclasses
import sys

import numpy as np
import pytest

from sbewh_amd import model
from sbewh_amd.propagator import SBEContext, jacobi_batch
from oracle.oracle_py import Oracle
from tests import cases

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("n", [2, 3, 4, 6, 8, 12, 16])
def test_jacobi_batch(n):
    say(f"[gpu] jacobi n={n}")
    rng = np.random.default_rng(100 + n)
    H = cases.random_hermitian(rng, 257, n)
    sweeps = 12
    U, e = jacobi_batch(H, sweeps)
    I = np.eye(n)
    assert np.abs(U @ np.conj(np.swapaxes(U, 1, 2)) - I).max() < 1e-13
    D = U @ H @ np.conj(np.swapaxes(U, 1, 2))
    off = D - np.einsum("kii->ki", D)[:, :, None] * I
    assert np.abs(off).max() < 1e-12 * np.abs(H).max()
    assert np.allclose(np.einsum("kii->ki", D).real, e, rtol=0, atol=1e-12)
    ref = np.linalg.eigvalsh(H)
    assert np.abs(np.sort(e, 1) - ref).max() < 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("n", [2, 8, 16])
def test_jacobi_batch_degenerate_and_diagonal(n):
    """the rotation rule at its edges: exactly degenerate eigenvalues (pairs
    with a_pp - a_qq = 0 or -0.0 and a non-zero a_pq), a diagonal matrix with
    equal entries and tiny off-diagonal couplings, and a matrix that is
    already diagonal (nothing rotates): U unitary, U H U^dagger diagonal, the
    eigenvalues those of numpy"""
    say(f"[gpu] jacobi edge cases n={n}")
    rng = np.random.default_rng(300 + n)
    mats = []
    # repeated eigenvalues in a random basis
    for _ in range(40):
        ev = np.repeat(rng.normal(size=(n + 1) // 2), 2)[:n]
        Q, _ = np.linalg.qr(rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)))
        mats.append(Q @ np.diag(ev) @ np.conj(Q.T))
    # equal diagonal, tiny couplings (d = 0 exactly in the first rotations), and -0.0 differences
    for s in (1e-3, 1e-9):
        H = np.eye(n, dtype=complex) * 0.7
        H += s * cases.random_hermitian(rng, 1, n)[0] * (1 - np.eye(n))
        mats.append(H)
        Hm = H.copy()
        Hm[np.diag_indices(n)] = np.where(np.arange(n) % 2 == 1, -0.0, 0.0)
        mats.append(Hm)
    # already diagonal
    mats.append(np.diag(rng.normal(size=n)).astype(complex))
    H = np.array(mats)
    U, e = jacobi_batch(H, 12)
    I = np.eye(n)
    assert np.abs(U @ np.conj(np.swapaxes(U, 1, 2)) - I).max() < 1e-13
    D = U @ H @ np.conj(np.swapaxes(U, 1, 2))
    off = D - np.einsum("kii->ki", D)[:, :, None] * I
    scale = np.abs(H).max(axis=(1, 2))
    assert (np.abs(off).max(axis=(1, 2)) < 1e-12 * scale).all()
    ref = np.linalg.eigvalsh(H)
    assert (np.abs(np.sort(e, 1) - ref).max(axis=1) < 1e-12 * scale).all()
    # a diagonal input comes back unchanged (no rotation)
    assert np.array_equal(e[-1], H[-1].diagonal().real)
    assert np.array_equal(U[-1], np.eye(n))


@pytest.mark.parametrize("name", ["chain", "square", "cubic"])
def test_phase_sum(name):
    say(f"[gpu] phase sum {name}")
    cfg = {"chain": cases.chain(), "square": cases.small_square(), "cubic": cases.small_cubic()}[name]
    g = SBEContext(cfg)
    o = Oracle(cfg, "direct")
    A, E = (0.11, -0.07, 0.05), (0.003, 0.002, -0.001)
    m_gpu = g.hk_mesh(A, E)
    m_cpu = o.hk_mesh(A, E)
    assert g.nf == o.nf
    assert rel(m_gpu, m_cpu) < 1e-13


@pytest.mark.parametrize("name", ["chain", "square", "cubic"])
def test_fermi_initial_rhs_observe(name):
    say(f"[gpu] fermi/initial/rhs/observe {name}")
    cfg = {"chain": cases.chain(sweeps=8), "square": cases.small_square(), "cubic": cases.small_cubic()}[name]
    g = SBEContext(cfg)
    o = Oracle(cfg)
    mu_g = g.fermi_level(cfg.occupied_bands)
    mu_o = o.fermi_level(cfg.occupied_bands)
    assert abs(mu_g - mu_o) < 1e-12 * max(1.0, abs(mu_o))
    st_g = g.initial_state(mu_o)
    st_o = o.initial_state(mu_o)
    assert rel(st_g, st_o) < 1e-12
    x = cases.perturbed_state(st_o, o.n, o.P, o.num_k, seed=3)
    t = 0.31 * o.t_start
    d_g = g.rhs(x, t)
    d_o = o.rhs(x, t)
    assert rel(d_g, d_o) < 1e-12, rel(d_g, d_o)
    j_g = g.observe(x)
    j_o = o.observe_j(x)
    assert np.abs(j_g - j_o).max() < 1e-12 * max(np.abs(j_o).max(), 1e-30)
    # second call: the Jacobi starts from the eigenvectors of the first (warm start)
    x2 = x.copy()
    x2[0] += 0.013
    t2 = 0.29 * o.t_start
    d_g2, d_o2 = g.rhs(x2, t2), o.rhs(x2, t2)
    assert rel(d_g2, d_o2) < 1e-12, rel(d_g2, d_o2)


@pytest.mark.parametrize("lr", [1, 2])
def test_wide_last_axis_box(lr, monkeypatch):
    """an R box wider than 16 cells along the last axis (the 1-D chain padded
    to |R| <= 10 with small hoppings, Fourier grid 32): the stage kernels form
    H_0 and H_E one entry at a time, and the few-band bodies take their
    separate-formation branch; RHS against the oracle, cold and warm-started"""
    say(f"[gpu] wide R box, chain, {lr} lane(s) per row")
    monkeypatch.setenv("SBEWH_STAGE_LR", str(lr))
    base = model.chain1d_model()
    cells = [(R, 0, 0) for R in range(-10, 11)]
    n = base.num_wann
    old = {tuple(int(v) for v in c): i for i, c in enumerate(base.cells)}
    H = np.zeros((len(cells), n, n), complex)
    D = np.zeros((len(cells), 3, n, n), complex)
    for i, c in enumerate(cells):
        if c in old:
            H[i], D[i] = base.H[old[c]], base.D[old[c]]
    rng = np.random.default_rng(11)
    for R in range(5, 11):
        h = 1e-4 * (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n)))
        H[cells.index((R, 0, 0))] = h
        H[cells.index((-R, 0, 0))] = h.conj().T
    m = model.TBModel(np.array(cells, np.int32), H, D, base.lattice)
    assert m.hermiticity_error() == 0
    cfg = dataclasses.replace(cases.chain(sweeps=8), model=m, Nf=(32, 1, 1))
    g, o = SBEContext(cfg), Oracle(cfg)
    st = o.initial_state(o.fermi_level(cfg.occupied_bands))
    x = cases.perturbed_state(st, o.n, o.P, o.num_k, seed=13)
    for t in (0.3 * o.t_start, 0.28 * o.t_start):
        assert rel(g.rhs(x, t), o.rhs(x, t)) < 1e-12
    assert g.stage_tuning()["lanes_per_row"] == lr
    g.close()
    o.close()


def test_rhs_no_dephasing_step_mask_cold_jacobi():
    say("[gpu] rhs variants")
    import dataclasses
    base = cases.small_square()
    for T2, w, cold in ((-1.0, 0.025, False), (10.0, -1.0, False), (10.0, 0.025, True)):
        cfg = dataclasses.replace(base, T2_fs=T2, soothing_width_eV=w, cold_jacobi=cold)
        g, o = SBEContext(cfg), Oracle(cfg)
        st = o.initial_state(o.fermi_level(cfg.occupied_bands))
        x = cases.perturbed_state(st, o.n, o.P, o.num_k, seed=5)
        for t in (0.0, 0.01 * o.t_start):
            assert rel(g.rhs(x, t), o.rhs(x, t)) < 1e-12


def check_error_norms(rg, ro):
    """the GPU's error norm of every replayed attempt against the oracle's
    norm of the same accepted attempt (same t, dt; states agree to ~1e-10).
    The norm is a max of cancelling differences: where it is tiny (no field)
    its relative precision is poor, so the bound is absolute in units of the
    acceptance threshold 1, plus relative where the norm decides the step."""
    lo, lg = ro["log"].as_dict(), rg["log"].as_dict()
    eo = lo["err"][lo["accepted"] == 1]
    eg = lg["err"]
    assert len(eg) == len(eo)
    d = np.abs(eg - eo)
    r = d / eo
    big = eo >= 1e-2
    i = int(np.argmax(r))
    say(f"  error norm: abs diff max {d.max():.2e}; rel diff median {np.median(r):.2e}, "
        f"max {r[i]:.2e} at norm {eo[i]:.2e}; rel max where norm >= 1e-2: {r[big].max():.2e} ({big.sum()} of {len(eo)})")
    assert d.max() < 1e-6
    assert r[big].max() < 1e-4


def test_replay_chain_full_pulse():
    """full 8-cycle propagation of the 1-D chain: the GPU replays the oracle's
    accepted dt sequence; j(t) and rho agree to 1e-10; the GPU's own adaptive
    run takes the same step sequence."""
    say("[gpu] replay chain1d")
    cfg = cases.chain(sweeps=5)
    g, o = SBEContext(cfg), Oracle(cfg)
    mu = o.fermi_level(cfg.occupied_bands)
    st0 = o.initial_state(mu)
    so = st0.copy()
    ro = o.integrate(so, cfg.output_count)
    sg = st0.copy()
    rg = g.integrate(sg, cfg.output_count, replay=ro["log"])
    assert rg["steps"] == ro["steps"]
    assert np.array_equal(rg["t"], ro["t"])
    scale = np.abs(ro["j"]).max()
    assert np.abs(rg["j"] - ro["j"]).max() < 1e-10 * scale
    assert rel(sg, so) < 1e-10
    check_error_norms(rg, ro)
    # own step control
    sa = st0.copy()
    ra = g.integrate(sa, cfg.output_count)
    say(f"  steps oracle {ro['steps']} gpu {ra['steps']}")
    assert abs(ra["steps"] - ro["steps"]) <= 2
    assert np.abs(ra["j"] - ro["j"]).max() < 1e-8 * scale


def test_replay_square_short_pulse():
    say("[gpu] replay square2d (15x15, 1 cycle)")
    cfg = cases.small_square(eps_rel=1e-7)
    g, o = SBEContext(cfg), Oracle(cfg)
    mu = o.fermi_level(cfg.occupied_bands)
    st0 = o.initial_state(mu)
    so = st0.copy()
    ro = o.integrate(so, cfg.output_count)
    sg = st0.copy()
    rg = g.integrate(sg, cfg.output_count, replay=ro["log"])
    scale = np.abs(ro["j"]).max()
    say(f"  steps {ro['steps']} rhs {ro['rhs_calls']} vs {rg['rhs_calls']}")
    # a replay runs only the accepted attempts: 6 RHS each + one FSAL start per interval
    assert rg["steps"] == ro["steps"]
    assert rg["rhs_calls"] == 6 * ro["steps"] + cfg.output_count - 1
    assert np.abs(rg["j"] - ro["j"]).max() < 1e-10 * scale
    assert rel(sg, so) < 1e-10
    check_error_norms(rg, ro)


def test_stage_per_launch_equals_attempt_per_launch(monkeypatch):
    """Grids of at least 512 stage blocks launch an attempt's six stages one by
    one, smaller grids in one launch; on the same problem both give the same
    bits (own step control), and the one-by-one path replays the oracle's
    step sequence to 1e-10 like the other."""
    say("[gpu] square2d 15x15: one launch per stage vs one per attempt")
    cfg = cases.small_square(eps_rel=1e-7)
    o = Oracle(cfg)
    mu = o.fermi_level(cfg.occupied_bands)
    st0 = o.initial_state(mu)
    so = st0.copy()
    ro = o.integrate(so, cfg.output_count)
    monkeypatch.setenv("SBEWH_STAGE_LR", "1")   # the same layout both ways
    res = {}
    for fuse in ("0", "512"):
        monkeypatch.setenv("SBEWH_FUSE_BELOW", fuse)
        g = SBEContext(cfg)
        s = st0.copy()
        r = g.integrate(s, cfg.output_count)
        res[fuse] = (r, s)
        if fuse == "0":
            sr = st0.copy()
            rr = g.integrate(sr, cfg.output_count, replay=ro["log"])
            scale = np.abs(ro["j"]).max()
            assert np.abs(rr["j"] - ro["j"]).max() < 1e-10 * scale
            assert rel(sr, so) < 1e-10
            check_error_norms(rr, ro)
        g.close()
    (r0, s0), (r1, s1) = res["0"], res["512"]
    say(f"  steps {r0['steps']} / {r1['steps']}")
    assert r0["steps"] == r1["steps"]
    assert np.array_equal(r0["j"], r1["j"])
    assert np.array_equal(s0, s1)


@pytest.mark.parametrize("nk,lr", [(16, 1), (16, 2), (256, 2)])
def test_one_block_attempt_kernel(nk, lr, monkeypatch):
    """The fused attempt kernel against one launch per stage on the 1-D chain:
    on 16 k a grid of a single stage block (that block takes the last ticket
    and reduces the error norm alone), on 256 k sixteen blocks.  Same bits as
    one launch per stage (own step control), and the oracle's step sequence
    replayed to 1e-10 with matching error norms."""
    say(f"[gpu] chain1d {nk} k, {lr} lane(s) per row: one launch per attempt vs one per stage")
    cfg = cases.chain(Nk=(nk, 1, 1))
    o = Oracle(cfg)
    mu = o.fermi_level(cfg.occupied_bands)
    st0 = o.initial_state(mu)
    so = st0.copy()
    ro = o.integrate(so, cfg.output_count)
    monkeypatch.setenv("SBEWH_STAGE_LR", str(lr))   # the same layout both ways
    res = {}
    for fuse in ("512", "0"):
        monkeypatch.setenv("SBEWH_FUSE_BELOW", fuse)
        g = SBEContext(cfg)
        s = st0.copy()
        r = g.integrate(s, cfg.output_count)
        res[fuse] = (r, s)
        if fuse == "512":
            sr = st0.copy()
            rr = g.integrate(sr, cfg.output_count, replay=ro["log"])
            scale = np.abs(ro["j"]).max()
            assert np.abs(rr["j"] - ro["j"]).max() < 1e-10 * scale
            assert rel(sr, so) < 1e-10
            check_error_norms(rr, ro)
        g.close()
    (r0, s0), (r1, s1) = res["512"], res["0"]
    say(f"  steps {r0['steps']} / {r1['steps']} (oracle {ro['steps']})")
    assert r0["steps"] == r1["steps"]
    assert np.array_equal(r0["j"], r1["j"])
    assert np.array_equal(s0, s1)


@pytest.fixture(scope="module")
def cubic_segment():
    """16-band 9^3 model over a short pulse window (2 um, 8 MV/cm, +-0.05 fs),
    rtol 1e-7: the oracle's run at a converged sweep count (8 cyclic sweeps
    from the identity on every call)"""
    base = cases.small_cubic()
    pul = [model.Pulse("sin2", 0.8, 2000.0, 0.0, 0.05)]
    cfg_o = dataclasses.replace(base, sweeps=8, eps_rel=1e-7, pulses=pul, output_count=8)
    o = Oracle(cfg_o)
    x0 = o.initial_state(o.fermi_level(cfg_o.occupied_bands))
    xo = x0.copy()
    say("[gpu] cubic 9^3 n=16 segment: oracle propagation")
    ro = o.integrate(xo, cfg_o.output_count)
    o.close()
    return cfg_o, x0, xo, ro


@pytest.mark.parametrize("lr", [2, 4])
def test_sixteen_band_attempt_kernel_layouts(monkeypatch, cubic_segment, lr):
    """16 bands, layout LR lanes per row: the attempt in one launch (fused,
    its last block reducing the error norm) gives the same bits as one launch
    per stage under the GPU's own step control, and the fused path replays the
    oracle's step sequence with j(t) and rho <= 1e-10 and matching error norms
    (GPU: at most 5 warm-started sweeps per RHS)."""
    cfg_o, x0, xo, ro = cubic_segment
    cfg = dataclasses.replace(cfg_o, sweeps=5)
    say(f"[gpu] cubic 9^3 n=16 LR={lr}: one launch per attempt vs one per stage")
    monkeypatch.setenv("SBEWH_STAGE_LR", str(lr))
    res = {}
    for fuse in ("100000", "0"):
        monkeypatch.setenv("SBEWH_FUSE_BELOW", fuse)
        g = SBEContext(cfg)
        s = x0.copy()
        r = g.integrate(s, cfg.output_count)
        assert g.stage_tuning()["lanes_per_row"] == lr
        res[fuse] = (r, s)
        if fuse == "100000":
            sr = x0.copy()
            rr = g.integrate(sr, cfg.output_count, replay=ro["log"])
            say(f"  fused replay: j rel {rel(rr['j'], ro['j']):.2e}, rho rel {rel(sr, xo):.2e}")
            assert rel(rr["j"], ro["j"]) <= 1e-10
            assert rel(sr, xo) <= 1e-10
            check_error_norms(rr, ro)
        g.close()
    (r0, s0), (r1, s1) = res["100000"], res["0"]
    say(f"  steps {r0['steps']} / {r1['steps']} (oracle {ro['steps']})")
    assert r0["steps"] == r1["steps"]
    assert np.array_equal(r0["j"], r1["j"])
    assert np.array_equal(s0, s1)
