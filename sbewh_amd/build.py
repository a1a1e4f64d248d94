"""Build libsbewh_hip.so for gfx950 with hipcc (in-tree: sbewh_amd/lib/).

    python -m sbewh_amd.build          # incremental
    python -m sbewh_amd.build --force  # rebuild everything
"""
from __future__ import annotations

import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIBDIR, "libsbewh_hip.so")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# -packed-fp32-ops: keep the compiler from forming packed FP32 instructions
COMMON = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
          "-packed-fp32-ops", "-Wall", "-Wno-unused-result", f"-I{INCLUDE}", f"-I{SRC}"]
SOURCES = ["phase_kernels.hip", "stage_kernels.hip", "rk_kernels.hip", "observer_kernels.hip", "sbe_api.cpp",
           "observer_api.cpp", "scan_api.cpp", "spectrum.cpp", "mlp_kernels.hip", "mlp_api.cpp", "noalloc.hip",
           "w90.cpp", "resident_kernels.hip"]


def _local_includes(path: str, seen=None) -> list:
    """every file reached from `path` through #include "..." lines (the
    source's own headers and include/*.h), so that an edit of any of them
    rebuilds the objects that use it"""
    seen = set() if seen is None else seen
    base = os.path.dirname(path)
    with open(path) as f:
        for line in f:
            line = line.strip()
            if line.startswith("#include") and '"' in line:
                name = line.split('"')[1]
                for root in (base, SRC, INCLUDE):
                    cand = os.path.normpath(os.path.join(root, name))
                    if os.path.exists(cand):
                        if cand not in seen:
                            seen.add(cand)
                            _local_includes(cand, seen)
                        break
    return sorted(seen)


def _newer(a: str, deps: list) -> bool:
    if not os.path.exists(a):
        return False
    t = os.path.getmtime(a)
    return all(os.path.getmtime(d) <= t for d in deps)


# stage_kernels.hip is built as two objects (SBE_STAGE_PART 1 and 2): the
# one-lane-per-row stage kernels (part 2) with the default machine scheduler,
# everything else (part 1) with the ILP-oriented one
UNITS = [(s, p) for s in SOURCES for p in ((1, 2) if s == "stage_kernels.hip" else (0,))]


def _compile(unit, force: bool) -> str:
    src, part = unit
    obj = os.path.join(OBJ, os.path.basename(src) + (f".part{part}" if part else "") + ".o")
    deps = [os.path.join(SRC, src)] + _local_includes(os.path.join(SRC, src)) + [os.path.abspath(__file__)]
    if not force and _newer(obj, deps):
        return obj
    lang = ["-x", "hip"]
    # the host controller's time arithmetic is kept free of FMA contraction so
    # that output times are the plain IEEE sums t0 + step * dt_out
    extra = ["-ffp-contract=off"] if src.endswith(".cpp") else []
    # the stage kernels are issue-bound with a third of their cycles waiting;
    # the ILP-oriented machine scheduler hides more of that latency at the same
    # occupancy (48^3 3-D model: 2.30 -> 2.14 ms per stage launch); the
    # one-lane-per-row kernels keep the default scheduler, which fits them in
    # 162 VGPRs (three waves per SIMD; 150 x 150 2-D model: 108 -> 93 us)
    if part:
        extra += [f"-DSBE_STAGE_PART={part}"]
    if part == 1:
        extra += ["-mllvm", "-amdgpu-sched-strategy=iterative-ilp"]
    cmd = [HIPCC] + COMMON + extra + lang + ["-c", os.path.join(SRC, src), "-o", obj + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    os.replace(obj + ".tmp", obj)
    return obj


def build(force: bool = False, verbose: bool = True) -> str:
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(len(UNITS), os.cpu_count() or 4)) as ex:
        objs = list(ex.map(lambda u: _compile(u, force), UNITS))
    if force or not _newer(LIB, objs):
        cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", LIB + ".tmp"] + objs
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"link failed:\n{r.stderr}")
        os.replace(LIB + ".tmp", LIB)
    build_cli(force or not _newer(CLI, [LIB] + [os.path.join(CLI_SRC, f) for f in os.listdir(CLI_SRC)]))
    if verbose:
        print(f"built {LIB}")
    return LIB


CLI_SRC = os.path.join(SRC, "cli")
CLI = os.path.join(LIBDIR, "sbewh_run")


def build_cli(needed: bool = True) -> str:
    """the command line front end (host C++ over the C ABI), next to the library"""
    if needed:
        cmd = ["g++", "-O2", "-std=c++17", "-Wall", f"-I{INCLUDE}", f"-I{CLI_SRC}", os.path.join(CLI_SRC, "run_sbe.cpp"),
               f"-L{LIBDIR}", "-lsbewh_hip", "-Wl,-rpath,$ORIGIN", "-o", CLI + ".tmp"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"sbewh_run build failed:\n{r.stderr}")
        os.replace(CLI + ".tmp", CLI)
    return CLI


if __name__ == "__main__":
    build(force="--force" in sys.argv)
