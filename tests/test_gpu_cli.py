"""sbewh_run end to end on the GPU: a deck in the reference's syntax over a
Wannier90 file written from the seeded 1-D chain, with k-space regions and
density matrices, against the same run through the Python API (same library,
same inputs: equal to rounding of the unit conversions); and a multi-run file
whose interrupted configuration alone is rerun by General.finishMultiRuns
(code/src/main/runSBE.cpp:19-72,205-216)."""
import os
import subprocess
import sys

import numpy as np
import pytest

from sbewh_amd import model, units, w90
from sbewh_amd.build import CLI
from sbewh_amd.propagator import FullObserver, Region, SBEContext, WHOLE_ZONE

pytestmark = pytest.mark.gpu

DECK = """[General]
runMode = SBE
dimensionality = 1
[TightBinding]
wannierSeed = chain
occupiedBands = 1
[FourierTransform]
Nf = 16
[Diagonalization]
maxSweepCount = 8
[Pulse]
pol = 1, 0, 0
Emax = 0.5
lambda = 3200
FWHM = 5.3
type = sin2
tStart = -1
tEnd = 1
{multi}
[Propagation]
Nk = 64
outputCount = 33
T2 = 5
epsAbs = 1e-12
epsRel = 1e-8
[KspaceRegions]
storeDensityMatrix = 1
densityOutputStride = 8
pt = M 0.25 0 0 0.1
line = F 0 0 0 \\
         0.5 0 0 \\
         4 E 0.1
"""


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def setup(tmp_path, multi=""):
    w90.write_tb(model.chain1d_model(), str(tmp_path / "chain_tb.dat"))
    (tmp_path / "input.txt").write_text(DECK.format(multi=multi))


def run(args, cwd):
    say(f"[gpu] sbewh_run {' '.join(args)}")
    r = subprocess.run([CLI] + args, cwd=cwd, capture_output=True, text=True, timeout=300)
    say(r.stdout[-2000:], r.stderr[-2000:])
    return r


def test_cli_run_matches_python_api(tmp_path):
    setup(tmp_path)
    r = run(["input.txt"], tmp_path)
    assert r.returncode == 0
    z = np.load(tmp_path / "runs" / "001" / "data.npz", allow_pickle=False)
    assert os.path.exists(tmp_path / "runs" / "001" / "args.txt")
    assert os.path.exists(tmp_path / "runs" / "001" / "input.txt")
    # the same run through the Python API
    m = w90.load(str(tmp_path / "chain"), dim=1)
    cfg = model.Config("cli", m, 1, (64, 1, 1), (16, 1, 1), [model.Pulse("sin2", 0.5, 3200.0, 0.0, 5.3)], T2_fs=5.0,
                       eps_abs=1e-12, eps_rel=1e-8, sweeps=8, occupied_bands=1, output_count=33)
    g = SBEContext(cfg)
    x = g.initial_state(g.fermi_level(1))
    regs = [WHOLE_ZONE, Region((0.25, 0, 0), 0.1, True)] + [Region((0.5 * i / 4, 0, 0), 0.1, False) for i in range(4)]
    ob = FullObserver(g, regs, smearing_width=units.from_eV(0.025), density_stride=8, capacity=33)
    g.integrate(x, 33, observer=ob)
    ref = ob.collectExpectationValues()
    tol = 1e-12

    def close(a, b):
        return np.abs(a - b).max() <= tol * max(np.abs(b).max(), 1e-300)

    assert close(z["t"], ref["t"]) and close(z["pulse/E"], ref["E"]) and close(z["pulse/A"], ref["A"])
    assert close(z["occHamMin"], ref["occupationHamMin"]) and close(z["occHamMax"], ref["occupationHamMax"])
    assert np.array_equal(z["tDensIndices"], ref["tDensIndices"]) and z["tDensIndices"].dtype == np.uint32
    assert np.allclose(z["lattice"], m.lattice, rtol=1e-15, atol=0)
    # whole zone: sample 0; pt: sample 1; line: samples 2..5 (shape (Nt, 4, ...))
    assert close(z["expectationValues/j"], ref["j"][:, 0])
    assert close(z["expectationValues/weight"], ref["kPointWeightSum"][:, 0])
    assert close(z["pt/occH"], ref["occupationHam"][:, 1]) and close(z["pt/occW"], ref["occupationWan"][:, 1])
    assert z["line/j"].shape == (33, 4, 3) and close(z["line/j"], ref["j"][:, 2:6])
    assert z["line/wan"].shape == (5, 4, 2, 2) and close(z["line/wan"], ref["wan"][:, 2:6])
    g.close()


def test_cli_multi_run_resume(tmp_path, monkeypatch):
    # the resumed run repeats b bit for bit only with the same stage layout
    # (each process times the layouts on its first RHS)
    monkeypatch.setenv("SBEWH_STAGE_LR", "1")
    setup(tmp_path, multi="multiRunFile = runs.txt")
    (tmp_path / "runs.txt").write_text("[a]\nEmax = 0.3\n[b]\nEmax = 0.6\ncep = 0.5\n")
    r = run(["input.txt", "-Propagation.outputCount=17"], tmp_path)
    assert r.returncode == 0 and "Configuration: a" in r.stdout and "Configuration: b" in r.stdout
    d = tmp_path / "runs" / "001"
    assert (d / ".a").read_text() == "1" and (d / ".b").read_text() == "1"
    za = np.load(d / "a.npz", allow_pickle=False)
    zb = np.load(d / "b.npz", allow_pickle=False)
    assert za["t"].shape == (17,)
    assert np.abs(zb["pulse/E"]).max() > 1.5 * np.abs(za["pulse/E"]).max()
    # configuration b interrupted: its results gone, its lock file empty and unlocked
    jb = zb["expectationValues/j"].copy()
    os.remove(d / "b.npz")
    (d / ".b").write_text("")
    mtime_a = os.path.getmtime(d / "a.npz")
    r = run(["-General.finishMultiRuns=1"], tmp_path)
    assert r.returncode == 0
    assert "Configuration: b" in r.stdout and "Configuration: a" not in r.stdout
    assert os.path.getmtime(d / "a.npz") == mtime_a
    zb2 = np.load(d / "b.npz", allow_pickle=False)
    assert np.array_equal(zb2["expectationValues/j"], jb)
    assert (d / ".b").read_text() == "1"
    # nothing left to run
    r = run(["-General.finishMultiRuns=1"], tmp_path)
    assert r.returncode == 0 and "Configuration:" not in r.stdout
