"""The C-ABI library loads, exports every symbol include/sbewh_hip.h declares,
and rejects invalid configurations with status codes before touching a GPU."""
import ctypes as C
import dataclasses
import os
import re

import numpy as np
import pytest

from sbewh_amd import _lib
from sbewh_amd.ctypes_types import Inputs
from tests import cases

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols(header="sbewh_hip.h"):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(sbe_[A-Za-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    syms = declared_symbols()
    assert len(syms) >= 19
    for s in syms:
        assert hasattr(L, s), s
    assert set(syms) == set(_lib.EXPORTS)
    assert L.sbe_version().decode().startswith("sbewh")
    surr = declared_symbols("sbewh_surrogate.h")
    assert set(surr) == set(_lib.SURROGATE_EXPORTS)
    for s in surr:
        assert hasattr(L, s), s


def _create(cfg, mutate=None):
    inp = Inputs(cfg)
    if mutate:
        mutate(inp)
    h = C.c_void_p()
    L = _lib.load()
    rc = L.sbe_create(C.byref(inp.model), C.byref(inp.grid), C.byref(inp.params), inp.pulses, inp.num_pulses,
                      C.byref(h))
    return rc, L.sbe_last_error(None).decode(), h


def test_invalid_configurations_are_rejected():
    cfg = cases.chain()
    rc, msg, h = _create(dataclasses.replace(cfg, Nk=(250, 1, 1)))
    assert rc == -1 and "not a multiple of Nf" in msg and not h.value
    rc, msg, _ = _create(cfg, lambda i: setattr(i.grid, "dim", 4))
    assert rc == -1 and "dim" in msg
    rc, msg, _ = _create(cfg, lambda i: setattr(i.params, "sweep_count", 99))
    assert rc == -1
    big = dataclasses.replace(cfg.model, H=np.zeros((9, 17, 17), complex), D=np.zeros((9, 3, 17, 17), complex))
    rc, msg, _ = _create(dataclasses.replace(cfg, model=big))
    assert rc == -1 and "num_wann" in msg
    no_origin = dataclasses.replace(cfg.model, cells=cfg.model.cells + np.array([10, 0, 0], np.int32))
    rc, msg, _ = _create(dataclasses.replace(cfg, model=no_origin))
    assert rc == -1 and "R=(0,0,0)" in msg
    rc, msg, _ = _create(cfg, lambda i: setattr(i, "num_pulses", 0))
    assert rc == -1
    rc, msg, _ = _create(cfg, lambda i: setattr(i.params, "flags", 8))
    assert rc == -1 and "flags" in msg
    L = _lib.load()
    assert L.sbe_jacobi_batch(0, 5, 1, None, None, None) == -1
    assert L.sbe_state_size(None) < 0


def test_cpp_propagator_header_links_against_library(tmp_path):
    """include/sbewh_propagator.hpp compiles and links against the built library
    (the program is only built, not run: running needs a GPU)."""
    import shutil
    import subprocess
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    src = tmp_path / "use_header.cpp"
    src.write_text(
        '#include "sbewh_propagator.hpp"\n'
        "int main(int argc, char **) {\n"
        "  if (argc < 5) return 0;\n"
        "  sbe_model m{}; sbe_grid g{}; sbe_params p{}; std::vector<sbe_subpulse> s(1);\n"
        "  sbewh::Context c(m, g, p, s); sbewh::PropagatorHIP prop(c);\n"
        "  sbewh::CPUstate x = prop.calcInitialState(0.0), dx;\n"
        "  prop(x, dx, 0.0); sbewh::ObserverHIP obs(c); obs(x, 0.0);\n"
        "  return (int)sbewh::integrate_const(c, x, 3, obs);\n"
        "}\n")
    libdir = os.path.dirname(_lib.LIB_PATH)
    out = tmp_path / "use_header"
    r = subprocess.run([gxx, "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src),
                        "-L", libdir, "-lsbewh_hip", "-Wl,-rpath," + libdir, "-o", str(out)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
