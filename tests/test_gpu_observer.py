"""The full observer on the GPU (sbe_observer_*, through the C ABI) against the
oracle's ObserverCPU restatement (itself pinned in test_oracle_observer.py):
k-space regions (whole zone, moving and fixed frame, a line of narrow
samples), Hamiltonian occupations with and without energy smearing, Wannier
occupations, min / max over k, region-averaged density matrices.  Bound:
1e-12 relative to the largest value of each quantity (FP64 single
evaluation); a whole propagation replaying the oracle's steps: 1e-10.
"""
import math
import sys

import numpy as np
import pytest

from oracle.oracle_py import Oracle
from sbewh_amd import units
from sbewh_amd.propagator import FullObserver, Region, SBEContext, WHOLE_ZONE
from tests import cases

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def regions_for(dim):
    line = [Region((0.1 + 0.2 * i, 0.3 if dim > 1 else 0.0, 0.0), 0.08, True, "line") for i in range(4)]
    return [WHOLE_ZONE, Region((0.25, 0.5 if dim > 1 else 0.0, 0.0), 0.2, True, "m"),
            Region((0.6, 0.1 if dim > 1 else 0.0, 0.2 if dim > 2 else 0.0), 0.3, False, "f")] + line


def as_tuples(regs):
    return [(r.k, r.fwhm, r.moving_frame) for r in regs]


def close(a, b, tol):
    scale = max(np.abs(b).max(), 1e-30)
    return np.abs(a - b).max() <= tol * scale, float(np.abs(a - b).max() / scale)


@pytest.mark.parametrize("name", ["chain", "square", "cubic"])
@pytest.mark.parametrize("smear_eV", [0.025, -1.0])
def test_observer_record_matches_oracle(name, smear_eV):
    cfg = {"chain": cases.chain(sweeps=10), "square": cases.small_square(sweeps=10),
           "cubic": cases.small_cubic(sweeps=12)}[name]
    say(f"[gpu] full observer {name}, smearing {smear_eV} eV")
    g, o = SBEContext(cfg), Oracle(cfg)
    st = o.initial_state(o.fermi_level(cfg.occupied_bands))
    x = cases.perturbed_state(st, o.n, o.P, o.num_k, seed=9)
    regs = regions_for(cfg.dim)
    w = units.from_eV(smear_eV) if smear_eV > 0 else -1.0
    ob = FullObserver(g, regs, smearing_width=w, density_stride=1, capacity=4)
    t = 0.37 * o.t_start
    ob.record(x, t)
    r = ob.collectExpectationValues()
    ref = o.observe_full(x, as_tuples(regs), w, density=True)
    assert r["t"][0] == t
    assert np.array_equal(r["A"][0], x[:3].real)
    assert np.allclose(r["E"][0], o.E(t), rtol=1e-14, atol=0)
    for key, rk in (("kPointWeightSum", "weight_sum"), ("j", "j"), ("occupationHam", "occ_ham"),
                    ("occupationWan", "occ_wan"), ("occupationHamMin", "occ_min"), ("occupationHamMax", "occ_max")):
        ok, err = close(r[key][0], ref[rk], 1e-12)
        say(f"  {key}: {err:.2e}")
        assert ok, (key, err)
    ok, err = close(r["wan"][0], ref["wan"], 1e-12)
    assert ok, ("wan", err)
    assert list(r["tDensIndices"]) == [0]
    ob.close()
    g.close()


def test_observer_density_stride_and_capacity():
    cfg = cases.chain(sweeps=8)
    g = SBEContext(cfg)
    x = g.initial_state(g.fermi_level(1))
    ob = FullObserver(g, [WHOLE_ZONE], density_stride=2, capacity=3)
    for i in range(3):
        ob.record(x, 0.1 * i)
    r = ob.collectExpectationValues()
    assert list(r["tDensIndices"]) == [0, 2]
    assert r["wan"].shape[0] == 2
    with pytest.raises(Exception):
        ob.record(x, 1.0)
    ob.reset()
    ob.record(x, 2.0)
    assert ob.count() == (1, 1)
    ob.close()
    g.close()


def test_integrate_observed_replays_oracle():
    """the 1-D chain's full pulse with the full observer at every output:
    the whole-zone current equals the integrator's j output; the last record
    matches the oracle's observer on the oracle's final state (replayed step
    sequence) to 1e-10"""
    cfg = cases.chain(sweeps=8)
    say("[gpu] chain1d integrate with the full observer")
    g, o = SBEContext(cfg), Oracle(cfg)
    st0 = o.initial_state(o.fermi_level(1))
    so = st0.copy()
    ro = o.integrate(so, cfg.output_count)
    regs = regions_for(1)
    w = units.from_eV(0.025)
    ob = FullObserver(g, regs, smearing_width=w, density_stride=64, capacity=cfg.output_count)
    sg = st0.copy()
    rg = g.integrate(sg, cfg.output_count, replay=ro["log"], observer=ob)
    r = ob.collectExpectationValues()
    assert r["t"].shape == (cfg.output_count,)
    assert np.array_equal(r["t"], rg["t"])
    ok, err = close(r["j"][:, 0], rg["j"], 1e-13)
    assert ok, err
    ref = o.observe_full(so, as_tuples(regs), w, density=True)
    for key, rk in (("j", "j"), ("occupationHam", "occ_ham"), ("occupationWan", "occ_wan"),
                    ("occupationHamMin", "occ_min"), ("occupationHamMax", "occ_max")):
        ok, err = close(r[key][-1], ref[rk], 1e-10)
        say(f"  last record {key}: {err:.2e}")
        assert ok, (key, err)
    assert list(r["tDensIndices"]) == list(range(0, cfg.output_count, 64))
    ob.close()
    g.close()
