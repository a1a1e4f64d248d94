"""Wannier90 reader and real-space operator pipeline (library host code,
include/sbewh_w90.h) against the reference's own tests and fixtures
(code/tests/w90util.cpp:8-158 with code/tests/data/example_tb.dat and
example_wsvec.dat, copied to tests/golden/w90/), and a write/read round trip
of a seeded model."""
import os

import numpy as np
import pytest

from sbewh_amd import model, units, w90

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "w90")


def test_read_tb_reference_values():
    # code/tests/w90util.cpp:8-23
    tb = w90.TbFile(os.path.join(GOLDEN, "example_tb.dat"))
    assert tb.num_wann == 2 and tb.num_cells == 21
    found = 0
    for c in range(tb.num_cells):
        r, deg, H, D = tb.cell(c)
        if r == (0, 0, 0):
            found += 1
            assert abs(H[0, 1].real - units.from_eV(0.14439922E+00)) < 1e-10
            assert abs(H[0, 1].imag - units.from_eV(0.22271074E-12)) < 1e-10
            assert abs(D[1, 1, 0].real - units.from_A(-0.87281205E-01)) < 1e-10
            assert abs(D[1, 1, 0].imag - units.from_A(0.80161572E-14)) < 1e-10
    assert found == 1
    # lattice in bohr
    assert abs(tb.lattice[0, 0] - units.from_A(4.3562141609999996)) < 1e-12


def test_read_wsvec_reference_values():
    # code/tests/w90util.cpp:25-41
    ws = w90.WsvecFile(os.path.join(GOLDEN, "example_wsvec.dat"))
    assert ws.count == 21 * 2 * 2
    r, m, n, T = ws.entry(5 * 4 + 1)
    assert r == (-1, 0, 1) and m == 0 and n == 1
    r, m, n, T = ws.entry(20)
    assert len(T) == 4
    assert T[2] == (2, 0, -2)


def _model(cells, H, D, lat=None):
    return model.TBModel(np.array(cells, np.int32), np.array(H, complex), np.array(D, complex),
                         np.eye(3) * 5.0 if lat is None else lat)


def test_symmetrize_reference_values(tmp_path):
    # code/tests/w90util.cpp:43-75 (values in the file's units: the written
    # energies are the model's values times EV, read back divided by it)
    n = 4
    H = np.zeros((3, n, n), complex)
    D = np.zeros((3, 3, n, n), complex)
    H[0, 0, 1] = 2                      # R = (0, 0, 1)
    H[1, 1, 0] = 1 - 10j                # R = (0, 0, -1)
    H[2, 0, 0] = 1 + 224j               # R = 0
    H[2, 1, 2] = 1 + 5j
    H[2, 2, 1] = 2 - 3j
    for d in range(3):
        D[2, d, d, d] = d + 1
    m = _model([(0, 0, 1), (0, 0, -1), (0, 0, 0)], H, D)
    p = str(tmp_path / "sym_tb.dat")
    w90.write_tb(m, p)
    out = w90.build_model(w90.TbFile(p), None, symmetrize=True, dim=3)
    idx = {tuple(c): i for i, c in enumerate(out.cells.tolist())}
    f1, f2, z = out.H[idx[(0, 0, 1)]], out.H[idx[(0, 0, -1)]], out.H[idx[(0, 0, 0)]]
    tol = 1e-14
    assert abs(f1[0, 1].real - 1.5) < tol
    assert abs(f2[1, 0].imag - -5) < tol
    assert abs(z[0, 0].real - 1) < tol and z[0, 0].imag == 0
    assert abs(z[1, 2].real - 1.5) < tol
    assert abs(z[2, 1].imag - -4) < tol
    Dz = out.D[idx[(0, 0, 0)]]
    for d in range(3):
        assert abs(Dz[d, d, d].real - (d + 1)) < tol


def _test_tb(n=2):
    # createTestTb, code/tests/w90util.cpp:79-105
    H = np.zeros((2, n, n), complex)
    D = np.zeros((2, 3, n, n), complex)
    for i in range(n):
        for j in range(n):
            H[0, i, j] = i * n + j
            H[1, i, j] = 10 + i * n + j
            for d in range(3):
                D[0, d, i, j] = d + i * n + j
                D[1, d, i, j] = 10 + d + i * n + j
    return _model([(0, 0, 0), (1, 1, 1)], H, D)


def test_old_realspace_operators(tmp_path):
    # code/tests/w90util.cpp:107-120: division by the degeneracy (1 and 2)
    n = 2
    p = str(tmp_path / "t_tb.dat")
    w90.write_tb(_test_tb(n), p, degeneracy=[1, 2])
    out = w90.build_model(w90.TbFile(p), None, symmetrize=False, dim=3)
    idx = {tuple(c): i for i, c in enumerate(out.cells.tolist())}
    assert abs(out.H[idx[(0, 0, 0)]][1, 0].real - (1 * n + 0)) < 1e-10
    assert abs(out.D[idx[(0, 0, 0)]][2, 1, 0].real - (2 + 1 * n + 0)) < 1e-10
    assert abs(out.H[idx[(1, 1, 1)]][0, 1].real - (10 + 0 * n + 1) / 2) < 1e-10
    assert abs(out.D[idx[(1, 1, 1)]][2, 0, 1].real - (10 + 2 + 0 * n + 1) / 2) < 1e-10


def test_new_realspace_operators(tmp_path):
    # code/tests/w90util.cpp:122-158: redistribution over the wsvec images
    n = 2
    p = str(tmp_path / "t_tb.dat")
    w90.write_tb(_test_tb(n), p, degeneracy=[1, 2])
    q = str(tmp_path / "t_wsvec.dat")
    with open(q, "w") as f:
        f.write("## test wsvec\n")
        for m in range(n):
            for k in range(n):
                f.write(f"0 0 0 {m + 1} {k + 1}\n2\n2 2 2\n0 0 0\n")
        for m in range(n):
            for k in range(n):
                T = [(0, 0, 0)] + ([(1, 1, 1)] if m else [])
                f.write(f"1 1 1 {m + 1} {k + 1}\n{len(T)}\n" + "".join(f"{a} {b} {c}\n" for a, b, c in T))
    out = w90.build_model(w90.TbFile(p), w90.WsvecFile(q), symmetrize=False, dim=3)
    assert out.num_cells == 3
    idx = {tuple(c): i for i, c in enumerate(out.cells.tolist())}
    H0, D0 = out.H[idx[(0, 0, 0)]], out.D[idx[(0, 0, 0)]]
    assert abs(H0[1, 0].real - 0.5 * (1 * n + 0)) < 1e-10
    assert abs(D0[2, 1, 0].real - 0.5 * (2 + 1 * n + 0)) < 1e-10
    H2, D2 = out.H[idx[(2, 2, 2)]], out.D[idx[(2, 2, 2)]]
    assert abs(H2[0, 1].real - 0.5 * (0 * n + 1)) < 1e-10
    assert abs(D2[2, 0, 1].real - 0.5 * (2 + 0 * n + 1)) < 1e-10
    assert abs(H2[1, 0].real - (0.25 * 10 + 0.75 * (1 * n + 0))) < 1e-10
    assert abs(D2[2, 1, 0].real - (0.25 * 10 + 0.75 * (2 + 1 * n + 0))) < 1e-10


def test_example_files_pipeline():
    """the reference's example seed through the whole pipeline: wsvec
    redistribution, symmetrisation, projection; the result is Hermitian in
    the H(-R) = H(R)^dagger sense, and projecting to fewer dimensions keeps
    the k = 0 Hamiltonian (the sum of all blocks)"""
    seed = os.path.join(GOLDEN, "example")
    m3 = w90.load(seed, dim=3)
    assert m3.num_wann == 2
    assert m3.hermiticity_error() < 1e-15
    for dim in (1, 2):
        md = w90.load(seed, dim=dim)
        assert np.all(md.cells[:, dim:] == 0)
        assert md.hermiticity_error() < 1e-15
        assert np.abs(md.H.sum(0) - m3.H.sum(0)).max() < 1e-14
        assert np.abs(md.D.sum(0) - m3.D.sum(0)).max() < 1e-14
    # the old scheme (no wsvec) sums to the same k = 0 Hamiltonian: every
    # Wigner-Seitz point's weight 1/degeneracy is spread over its images
    tb = w90.TbFile(seed + "_tb.dat")
    old = w90.build_model(tb, None, symmetrize=True, dim=3)
    assert np.abs(old.H.sum(0) - m3.H.sum(0)).max() < 1e-14


def test_write_read_round_trip(tmp_path):
    m = model.chain1d_model()
    p = str(tmp_path / "chain_tb.dat")
    w90.write_tb(m, p)
    out = w90.build_model(w90.TbFile(p), None, symmetrize=False, dim=3)
    order = np.lexsort(m.cells.T[::-1])
    assert np.array_equal(out.cells, m.cells[order])
    assert np.abs(out.H - m.H[order]).max() < 1e-15 * np.abs(m.H).max()
    assert np.abs(out.D - m.D[order]).max() < 1e-15 * np.abs(m.D).max()
    assert np.allclose(out.lattice, m.lattice, rtol=1e-15, atol=0)


def test_errors(tmp_path):
    with pytest.raises(w90.W90Error):
        w90.TbFile(str(tmp_path / "missing_tb.dat"))
    bad = tmp_path / "bad_tb.dat"
    bad.write_text("header\n1 0 0\n0 1 0\n")
    with pytest.raises(w90.W90Error, match="lattice vector 2"):
        w90.TbFile(str(bad))
