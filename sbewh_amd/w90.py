"""Wannier90 tight-binding files (<seed>_tb.dat, <seed>_wsvec.dat) as models.

``load(seed, dim)`` is the reference's TightBindingParameter pipeline
(code/src/io/TightBindingParameter.cpp:111-148): read <seed>_tb.dat, use
<seed>_wsvec.dat when it exists (redistribution over the Wigner-Seitz
supercell images), else divide by the degeneracies; symmetrise and/or drop
imaginary parts as requested; fold the cells onto the first ``dim``
directions.  The work is done by the library's host code (include/sbewh_w90.h);
the result is a ``model.TBModel`` in atomic units.

``write_tb(model, path)`` writes a model in the _tb.dat layout the reader
expects (energies in eV, lengths in angstrom), e.g.
to hand a seeded synthetic model to a reference run.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import units
from ._lib import load as load_lib
from .ctypes_types import SbeModel
from .model import TBModel

SYMMETRIZE, ONLY_REAL = 1, 2

_bound = False


def _lib():
    global _bound
    L = load_lib()
    if not _bound:
        P, vp, i = C.POINTER, C.c_void_p, C.c_int
        dp, ip = P(C.c_double), P(C.c_int)
        L.sbe_w90_last_error.restype = C.c_char_p
        L.sbe_w90_tb_read.argtypes = [C.c_char_p, P(vp)]
        L.sbe_w90_tb_free.argtypes = [vp]
        L.sbe_w90_tb_info.argtypes = [vp, ip, ip, dp]
        L.sbe_w90_tb_cell.argtypes = [vp, i, ip, ip, dp, dp]
        L.sbe_w90_wsvec_read.argtypes = [C.c_char_p, P(vp)]
        L.sbe_w90_wsvec_free.argtypes = [vp]
        L.sbe_w90_wsvec_count.argtypes = [vp]
        L.sbe_w90_wsvec_entry.argtypes = [vp, i, ip, ip, ip, ip, ip, i]
        L.sbe_w90_model_build.argtypes = [vp, vp, i, i, P(vp)]
        L.sbe_w90_model_view.argtypes = [vp, P(SbeModel)]
        L.sbe_w90_model_free.argtypes = [vp]
        _bound = True
    return L


class W90Error(RuntimeError):
    pass


def _check(rc):
    if rc < 0:
        raise W90Error(_lib().sbe_w90_last_error().decode())
    return rc


class TbFile:
    """the raw content of a _tb.dat file (w90_read_tb)"""

    def __init__(self, path: str):
        L = _lib()
        h = C.c_void_p()
        _check(L.sbe_w90_tb_read(os.fsencode(path), C.byref(h)))
        self.h = h
        nw, nc = C.c_int(), C.c_int()
        lat = np.zeros(9)
        _check(L.sbe_w90_tb_info(h, C.byref(nw), C.byref(nc), lat.ctypes.data_as(C.POINTER(C.c_double))))
        self.num_wann, self.num_cells, self.lattice = nw.value, nc.value, lat.reshape(3, 3)

    def cell(self, c: int):
        """(R, degeneracy, H[n,n], D[3,n,n]) of cell c as read (atomic units)"""
        n = self.num_wann
        r = (C.c_int * 3)()
        deg = C.c_int()
        H = np.zeros((n, n), complex)
        D = np.zeros((3, n, n), complex)
        _check(_lib().sbe_w90_tb_cell(self.h, c, r, C.byref(deg), H.view(float).ctypes.data_as(C.POINTER(C.c_double)),
                                      D.view(float).ctypes.data_as(C.POINTER(C.c_double))))
        return tuple(r), deg.value, H, D

    def close(self):
        if getattr(self, "h", None):
            _lib().sbe_w90_tb_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class WsvecFile:
    """the raw content of a _wsvec.dat file (w90_read_wsvec)"""

    def __init__(self, path: str):
        L = _lib()
        h = C.c_void_p()
        _check(L.sbe_w90_wsvec_read(os.fsencode(path), C.byref(h)))
        self.h = h
        self.count = _check(L.sbe_w90_wsvec_count(h))

    def entry(self, i: int):
        """(R, m, n, [T ...]) of entry i (m, n zero-based)"""
        L = _lib()
        r = (C.c_int * 3)()
        m, n, cnt = C.c_int(), C.c_int(), C.c_int()
        _check(L.sbe_w90_wsvec_entry(self.h, i, r, C.byref(m), C.byref(n), C.byref(cnt), None, 0))
        T = (C.c_int * (3 * max(1, cnt.value)))()
        _check(L.sbe_w90_wsvec_entry(self.h, i, r, None, None, None, T, cnt.value))
        return tuple(r), m.value, n.value, [tuple(T[3 * j:3 * j + 3]) for j in range(cnt.value)]

    def close(self):
        if getattr(self, "h", None):
            _lib().sbe_w90_wsvec_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def build_model(tb: TbFile, wsvec: WsvecFile | None = None, symmetrize: bool = True, only_real: bool = False,
                dim: int = 3) -> TBModel:
    """real-space operators of a read file as a TBModel (cells sorted)"""
    L = _lib()
    h = C.c_void_p()
    flags = (SYMMETRIZE if symmetrize else 0) | (ONLY_REAL if only_real else 0)
    _check(L.sbe_w90_model_build(tb.h, wsvec.h if wsvec is not None else None, flags, dim, C.byref(h)))
    try:
        v = SbeModel()
        _check(L.sbe_w90_model_view(h, C.byref(v)))
        n, nc = v.num_wann, v.num_cells
        cells = np.ctypeslib.as_array(v.cells, shape=(nc * 3,)).reshape(nc, 3).astype(np.int32).copy()
        H = np.ctypeslib.as_array(v.H, shape=(nc * n * n * 2,)).copy().view(complex).reshape(nc, n, n)
        D = np.ctypeslib.as_array(v.D, shape=(nc * 3 * n * n * 2,)).copy().view(complex).reshape(nc, 3, n, n)
        lat = np.array(v.lattice[:], float).reshape(3, 3)
    finally:
        L.sbe_w90_model_free(h)
    return TBModel(cells, H, D, lat)


def load(seed: str, dim: int = 3, symmetrize: bool = True, only_real: bool = False) -> TBModel:
    """<seed>_tb.dat (+ <seed>_wsvec.dat if present) -> TBModel, the
    reference's evalWannierFiles (code/src/io/TightBindingParameter.cpp:111-148)"""
    tb = TbFile(seed + "_tb.dat")
    ws = WsvecFile(seed + "_wsvec.dat") if os.path.exists(seed + "_wsvec.dat") else None
    try:
        return build_model(tb, ws, symmetrize, only_real, dim)
    finally:
        tb.close()
        if ws is not None:
            ws.close()


def write_tb(model: TBModel, path: str, header: str = "written by sbewh_amd", degeneracy=None) -> None:
    """the model in the _tb.dat layout: lattice (angstrom), num_wann, nrpts,
    degeneracies (1 unless given per cell), then per cell R and the lines
    'm n Re Im' of H (eV) with m running fastest, then per cell R and
    'm n Re(x) Im(x) Re(y) Im(y) Re(z) Im(z)' of the position operator
    (angstrom)"""
    n, nc = model.num_wann, model.num_cells
    deg = [1] * nc if degeneracy is None else [int(x) for x in degeneracy]
    with open(path, "w") as f:
        f.write(header.rstrip("\n") + "\n")
        for a in np.asarray(model.lattice, float):
            f.write("".join(f"{units.to_A(x):26.16f}" for x in a) + "\n")
        f.write(f"{n:12d}\n{nc:12d}\n")
        for i in range(0, nc, 15):
            f.write("".join(f"{deg[i + j]:5d}" for j in range(min(15, nc - i))) + "\n")
        for c in range(nc):
            f.write("\n" + "".join(f"{x:5d}" for x in model.cells[c]) + "\n")
            for b in range(n):
                for a in range(n):
                    h = model.H[c, a, b] * units.EV
                    f.write(f"{a + 1:5d}{b + 1:5d} {h.real:.17e} {h.imag:.17e}\n")
        for c in range(nc):
            f.write("\n" + "".join(f"{x:5d}" for x in model.cells[c]) + "\n")
            for b in range(n):
                for a in range(n):
                    d = model.D[c, :, a, b] * units.ANGSTROM
                    f.write(f"{a + 1:5d}{b + 1:5d}" + "".join(f" {x.real:.17e} {x.imag:.17e}" for x in d) + "\n")
