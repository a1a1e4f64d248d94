"""The batched scan (sbe_scan_*) reproduces independent single-instance runs.

Four instances of the 2-D 8-band model with uneven grids (15^2, 30^2), three
pulse shapes and two dephasing times advance together; each must match
sbe_integrate on a context built from the same configuration: the same number
of accepted steps and RHS calls, j(t) and the final density matrices to
<= 1e-12 relative (the kernels are shared, so the step sequences coincide),
and j(t) of one instance against the CPU oracle replaying its step sequence
to <= 1e-10 relative.
"""
import dataclasses
import sys

import numpy as np
import pytest

from sbewh_amd import model
from sbewh_amd.propagator import SBEContext
from sbewh_amd.scan import Scan
from oracle.oracle_py import Oracle
from tests import cases

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def small_scan():
    base = cases.small_square(cycles=1.0)
    out = []
    specs = [((15, 15, 1), 1.0, 0.3, 0.0, 10.0), ((30, 30, 1), 1.4, 0.8, 1.0, 10.0),
             ((15, 15, 1), 0.5, 0.0, 2.0, 6.0), ((30, 30, 1), 1.0, 1.0, 3.0, 6.0)]
    for i, (Nk, E0, eps, cep, T2) in enumerate(specs):
        pul = model.elliptic_sin2(E0, eps, 3000.0, 1.0, cep=cep)
        out.append(dataclasses.replace(base, name=f"s{i}", Nk=Nk, pulses=pul, T2_fs=T2, output_count=48))
    return out


def test_scan_matches_single_instance_runs():
    cfgs = small_scan()
    say("[gpu] scan of 4 instances (15^2 / 30^2)")
    sc = Scan(cfgs)
    r = sc.run()
    assert (r["steps"] > 0).all()
    for i, cfg in enumerate(cfgs):
        say(f"[gpu] single-instance run {i} Nk={cfg.Nk}")
        g = SBEContext(cfg)
        mu = g.fermi_level(cfg.occupied_bands)
        x = g.initial_state(mu)
        s = g.integrate(x, cfg.output_count)
        assert r["steps"][i] == s["steps"], (i, r["steps"][i], s["steps"])
        assert r["rhs_calls"][i] == s["rhs_calls"]
        assert np.array_equal(r["t"][i], s["t"])
        assert rel(r["j"][i], s["j"]) <= 1e-12, (i, rel(r["j"][i], s["j"]))
        assert rel(sc.final_state(i), x) <= 1e-12
    tun = sc.tuning()
    assert tun["lanes_per_row"] in (1, 2, 4)


def test_scan_instance_against_oracle_replay():
    cfg = small_scan()[1]
    say("[gpu] scan instance vs oracle replay")
    sc = Scan([cfg])
    r = sc.run()
    g = SBEContext(cfg)
    mu = g.fermi_level(cfg.occupied_bands)
    x = g.initial_state(mu)
    s = g.integrate(x, cfg.output_count)
    o = Oracle(cfg)
    xo = o.initial_state(o.fermi_level(cfg.occupied_bands))
    ro = o.integrate(xo, cfg.output_count, replay=s["log"])
    assert rel(r["j"][0], ro["j"]) <= 1e-10
    assert rel(sc.final_state(0), xo) <= 1e-10
