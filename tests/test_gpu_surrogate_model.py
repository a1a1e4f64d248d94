"""The trained surrogate ensemble (sbewh_amd/data/surrogate_ensemble.npz) on
the spectra the batched solver generated (sbewh_amd/data/
surrogate_square2d_15x15.npz; the first 512 rows were held out of training):
median per-harmonic relative log10-yield error <= 5 %, the time per answer
against the solver's time per spectrum on the same grid, and the solver
fallback outside the training box."""
import os
import sys
import time

import numpy as np
import pytest

from sbewh_amd.surrogate import DATA_DIR, Surrogate, relative_log_error, solve

pytestmark = pytest.mark.gpu

DATA = os.path.join(DATA_DIR, "surrogate_square2d_15x15.npz")
ENS = os.path.join(DATA_DIR, "surrogate_ensemble.npz")
N_TEST = 512


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_trained_ensemble_accuracy_and_gain():
    z = np.load(DATA, allow_pickle=False)
    P, Y = z["params"], z["log_yields"]
    s = Surrogate.load(ENS)
    s.predict(P[:8])  # first call allocates the batch buffers
    t0 = time.perf_counter()
    mean, spread = s.predict(P[:N_TEST])
    t_net = time.perf_counter() - t0
    err = relative_log_error(mean, Y[:N_TEST])
    med = float(np.median(err))
    solver_per_spectrum = float(z["seconds"].sum()) / len(P)   # batched solver, same grid
    gain = solver_per_spectrum / (t_net / N_TEST)
    say(f"  median rel log10-yield error {med:.4f}, p90 {np.quantile(err, 0.9):.4f}; "
        f"{t_net * 1e3:.1f} ms for {N_TEST} answers; solver {solver_per_spectrum:.3f} s/spectrum; gain {gain:.3g}x; "
        f"spread > tol: {(spread.max(1) > s.tol).mean():.3f}")
    assert med <= 0.05
    # without the network every answer costs a solver run
    assert gain > 1e3
    # the ensemble spread tracks the error: answers it would hand to the solver are the worse ones
    bad = spread.max(1) > s.tol
    if bad.any() and (~bad).any():
        assert np.median(err[bad]) >= np.median(err[~bad])


def test_fallback_outside_training_box_uses_the_solver(monkeypatch):
    # one stage layout for both solver runs, so that they are bit-identical
    monkeypatch.setenv("SBEWH_STAGE_LR", "1")
    z = np.load(DATA, allow_pickle=False)
    inside = z["params"][0]
    outside = np.array([1.0, 0.3, 0.5, 2000.0, 5.0, 10.0])   # 5 cycles: below the trained 6..12
    s = Surrogate.load(ENS)
    out, used, info = s.query(np.stack([outside, inside]))
    assert used[0] and not used[1]
    ly, _, steps = solve(outside[None, :])
    assert steps[0] > 0
    assert np.abs(out[0] - ly[0]).max() < 1e-9
    mean, _ = s.predict(inside[None, :])
    assert np.array_equal(out[1], mean[0])
