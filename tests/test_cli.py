"""The command line front end sbewh_run (host C++ over the C ABI) on the CPU:
the reference's input-deck syntax (sections, '#' comments, '\\' continuation,
dotted names, -Group.name=value options; code/src/io/ParameterIO.cpp:137-296),
the SBE parameters with their units and defaults (code/src/io/*Parameter.cpp),
k-space region definitions (code/src/io/KspaceRegionParameter.cpp:78-164),
multi-run files (code/src/io/ParameterIO.cpp:298-336), and the npz container
(checked by numpy.load).  Parsing runs in the Dry run mode, which resolves
every parameter and prints it without touching a GPU."""
import json
import math
import os
import subprocess

import numpy as np
import pytest

from sbewh_amd import units
from sbewh_amd.build import CLI


def run(args, cwd):
    return subprocess.run([CLI] + args, cwd=cwd, capture_output=True, text=True, timeout=60)


DECK = """[General]
runMode = Dry   # parse only
dimensionality = 2

[TightBinding]
wannierSeed = ../data/model
occupiedBands = 4

[FourierTransform]
Nf = 10

[Pulse]
pol = 1, 0, 0
Emax = 2 # V/nm
lambda = 3000 # nm
FWHM = 30.02 # fs
type = sin2
tStart = -1
tEnd = 1
multiRunFile = runs.txt

[Propagation]
Nk = 40 40
outputCount = 64
T2 = 2 fs
soothingWidth =  0.05

[KspaceRegions]
storeDensityMatrix = 1
densityOutputStride = 6
Ks = F 0.33333333 0.33333333 0  0.1
slice = M 0.3333333 0.33333333 0 \\
          0.6666666 0.66666666 0 \\
          50 I 0.100000
area  = F 0 0 0 \\
          1 0 0 \\
          0 1 0 \\
          20 20 E 0.05
"""

RUNS = """Emax = 99   # before the first section: discarded by the first reset
[weak]
Emax = 0.5
[strong]
Emax = 3
cep = 1.5
"""


def test_parameter_list():
    r = subprocess.run([CLI, "-par"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0
    assert "Propagation.epsRel = 1e-10" in r.stdout
    assert "Pulse.Emax3 = 0" in r.stdout


def test_dry_run_resolves_deck(tmp_path):
    (tmp_path / "input.txt").write_text(DECK)
    (tmp_path / "runs.txt").write_text(RUNS)
    r = run(["input.txt", "-Propagation.epsRel=1e-8", "-Diagonalization.maxSweepCount=7"], tmp_path)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    cfg = json.loads(lines[0])
    assert cfg["dim"] == 2 and cfg["Nk"] == [40, 40, 1] and cfg["Nf"] == [10, 10, 1]
    assert cfg["outputCount"] == 64 and cfg["sweeps"] == 7 and cfg["epsRel"] == 1e-8
    assert cfg["epsAbs"] == 1e-10                                   # default
    assert math.isclose(cfg["F2"], 1 / units.from_fs(2.0), rel_tol=1e-15)
    assert math.isclose(cfg["soothingWidth"], units.from_eV(0.05), rel_tol=1e-15)
    # occupationSmearingWidth unset: the soothing width (code/src/io/PropagationParameter.cpp:104-108)
    assert math.isclose(cfg["smearing"], units.from_eV(0.05), rel_tol=1e-15)
    assert math.isclose(cfg["kT"], units.from_K(300.0), rel_tol=1e-15)
    assert cfg["seed"] == "../data/model"
    # the whole zone, then the named regions in name order (std::map: Ks < area < slice)
    regions = [ln.split() for ln in lines if ln.startswith("region ")]
    assert [g[1] for g in regions] == ["expectationValues", "Ks", "area", "slice"]
    by = {g[1]: g for g in regions}
    assert by["expectationValues"][2] == "M" and by["expectationValues"][4] == "inf"
    assert by["Ks"][2] == "F" and float(by["Ks"][4]) == 0.1
    assert by["slice"][2] == "M" and by["slice"][6:9] == ["50", "1", "1"]
    assert by["area"][6:9] == ["20", "20", "1"]
    configs = {ln.split()[1].rstrip(":"): ln.split() for ln in lines if ln.startswith("config ")}
    assert set(configs) == {"weak", "strong"}
    assert math.isclose(float(configs["weak"][5]), units.from_V_nm(0.5), rel_tol=1e-15)
    assert math.isclose(float(configs["strong"][5]), units.from_V_nm(3.0), rel_tol=1e-15)
    assert float(configs["strong"][9]) == 1.5 and float(configs["weak"][9]) == 0.0
    assert math.isclose(float(configs["weak"][7]), units.omega_from_lambda_nm(3000.0), rel_tol=1e-14)


def test_single_value_nk_and_subpulses(tmp_path):
    deck = """[General]
dimensionality = 3
[Pulse]
Emax = 1
lambda = 2000
FWHM = 10
type = gauss
tStart = -2
tEnd = 2
pol = 1, 0, 0
Emax1 = 0.3
pol1 = 0 1 0
[Propagation]
Nk = 24
"""
    (tmp_path / "input.txt").write_text(deck)
    r = run([], tmp_path)                     # input.txt by default
    assert r.returncode == 0, r.stderr
    cfg = json.loads(r.stdout.splitlines()[0])
    assert cfg["Nk"] == [24, 24, 24] and cfg["pulses"] == 1   # only sub-pulse 1 is defined
    assert cfg["runMode"] == "Dry"


@pytest.mark.parametrize("line,msg", [
    ("[Propagation]\nNkk = 3\n", "Name 'Nkk' not known"),
    ("[Nope]\n", "subgroup 'Nope' not known"),
    ("[Propagation]\nT2 = abc\n", "Could not read value of 'abc'"),
    ("[KspaceRegions]\nx = Q 0 0 0 0.1\n", "mode must be either"),
    ("[KspaceRegions]\nx = M 0 0 0 1.5\n", "smaller 1"),
    ("[KspaceRegions]\nx = M 0 0 0 1 0 0 1 E 0.1\n", "bigger than 1"),
])
def test_deck_errors(tmp_path, line, msg):
    (tmp_path / "input.txt").write_text("[Pulse]\npol = 1,0,0\ntype = sin2\nEmax = 1\nlambda = 800\nFWHM = 5\n"
                                        "tStart = -1\ntEnd = 1\n" + line)
    r = run(["input.txt"], tmp_path)
    assert r.returncode == 1
    assert msg in r.stderr


def test_npz_container(tmp_path):
    p = str(tmp_path / "x.npz")
    r = subprocess.run([CLI, "--write-test-npz", p], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0
    z = np.load(p, allow_pickle=False)
    assert sorted(z.files) == ["pulse/E", "region/wan", "t", "tDensIndices"]
    assert np.array_equal(z["t"], [0.0, 0.5, 1.25])
    assert z["pulse/E"].shape == (3, 2) and np.allclose(z["pulse/E"].ravel(), 0.1 * np.arange(6) - 0.2)
    assert z["region/wan"].dtype == np.complex128 and z["region/wan"][1, 0] == -3 + 0.25j
    assert z["tDensIndices"].dtype == np.uint32 and z["tDensIndices"][1] == 4000000000
