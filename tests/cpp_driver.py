"""Build and load the C++ functor test driver (tests/cpp/functor_dopri5.cpp),
a host-only shared library that links libsbewh_hip.so."""
import ctypes as C
import os
import shutil
import subprocess

from sbewh_amd import _lib
from sbewh_amd.ctypes_types import SbeGrid, SbeModel, SbeParams, SbeSubpulse

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cpp", "functor_dopri5.cpp")
OUT = os.path.join(HERE, "cpp", "build", "libfunctor_dopri5.so")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")


def build(force: bool = False) -> str:
    """g++ the driver against the built library (host code, seconds)"""
    if not force and os.path.exists(OUT) and os.path.getmtime(OUT) >= max(
            os.path.getmtime(SRC), os.path.getmtime(os.path.join(INCLUDE, "sbewh_propagator.hpp"))):
        return OUT
    gxx = shutil.which("g++")
    if gxx is None:
        raise RuntimeError("g++ not available")
    os.makedirs(os.path.dirname(OUT), exist_ok=True)
    libdir = os.path.dirname(_lib.LIB_PATH)
    subprocess.run([gxx, "-std=c++17", "-O2", "-fPIC", "-shared", "-Wall", "-Werror", "-I", INCLUDE, SRC, "-L", libdir,
                    "-lsbewh_hip", "-Wl,-rpath," + libdir, "-o", OUT + ".tmp"], check=True)
    os.replace(OUT + ".tmp", OUT)
    return OUT


def load():
    _lib.load()  # the driver's dependency, from the same in-tree path
    L = C.CDLL(build())
    P = C.POINTER
    L.functor_dopri5.argtypes = [P(SbeModel), P(SbeGrid), P(SbeParams), P(SbeSubpulse), C.c_int, C.c_double,
                                 C.c_double, C.c_double, C.c_int, C.c_int, P(C.c_double), P(C.c_double)]
    return L
