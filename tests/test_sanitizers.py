"""AddressSanitizer + UndefinedBehaviorSanitizer builds of the host code (CPU
only): the oracle (plain C) over all its entry points, the Wannier90 reader
and operator pipeline (C++) over the reference's example files, and the
command line front end parsing decks in its Dry mode.  gcc / g++ builds of
the host sources; no GPU code is built with a sanitizer."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HERE = os.path.join(ROOT, "tests", "sanitize")
SAN = ["-fsanitize=address,undefined", "-fno-omit-frame-pointer", "-fno-sanitize-recover=undefined", "-g", "-O1"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")

pytestmark = pytest.mark.skipif(shutil.which("gcc") is None or shutil.which("g++") is None, reason="needs gcc/g++")


def build(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_oracle_under_sanitizers(tmp_path):
    exe = str(tmp_path / "oracle_driver")
    build(["gcc", "-std=gnu11", "-fopenmp", "-fcx-fortran-rules"] + SAN +
          [os.path.join(HERE, "oracle_driver.c"), os.path.join(ROOT, "oracle", "sbe_oracle.c"), "-lm", "-o", exe])
    r = subprocess.run([exe], capture_output=True, text=True, env=dict(ENV, OMP_NUM_THREADS="2"), timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "oracle ok" in r.stdout and "runtime error" not in r.stderr


def test_w90_pipeline_under_sanitizers(tmp_path):
    exe = str(tmp_path / "w90_driver")
    build(["g++", "-std=c++17"] + SAN + ["-I" + os.path.join(ROOT, "include"), os.path.join(HERE, "w90_driver.cpp"),
                                         os.path.join(ROOT, "sbewh_amd", "csrc", "w90.cpp"), "-o", exe])
    r = subprocess.run([exe, os.path.join(ROOT, "tests", "golden", "w90", "example")], capture_output=True, text=True,
                       env=ENV, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "w90 ok" in r.stdout and "runtime error" not in r.stderr


def test_cli_parsing_under_sanitizers(tmp_path):
    from sbewh_amd.build import LIBDIR
    exe = str(tmp_path / "sbewh_run_san")
    cli = os.path.join(ROOT, "sbewh_amd", "csrc", "cli")
    build(["g++", "-std=c++17"] + SAN + ["-I" + os.path.join(ROOT, "include"), "-I" + cli,
                                         os.path.join(cli, "run_sbe.cpp"), "-L" + LIBDIR, "-lsbewh_hip",
                                         "-Wl,-rpath," + LIBDIR, "-o", exe])
    from tests.test_cli import DECK, RUNS
    (tmp_path / "input.txt").write_text(DECK)
    (tmp_path / "runs.txt").write_text(RUNS)
    # leaks of the HIP runtime that the library pulls in are not the host code's
    env = dict(ENV, ASAN_OPTIONS="detect_leaks=0")
    r = subprocess.run([exe, "input.txt", "-Propagation.epsRel=1e-8"], cwd=tmp_path, capture_output=True, text=True,
                       env=env, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "config strong" in r.stdout and "runtime error" not in r.stderr
    r = subprocess.run([exe, "--write-test-npz", str(tmp_path / "x.npz")], capture_output=True, text=True, env=env,
                       timeout=60)
    assert r.returncode == 0, r.stderr
