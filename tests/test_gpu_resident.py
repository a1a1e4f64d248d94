"""The resident integrator (Config.device_controller: the whole
integrate_const of a 1-D instance in one cooperative launch, the step-size
control on the GPU) against the host controller and the CPU oracle.

Both controllers apply the same step rules to the same error norms; the
device's pow / cos / sin may round differently from the host's, so the step
sequences may part at that level: steps within 2, j(t) and rho within 1e-7
relative of each other (the integrator's own tolerance is rtol 1e-8), output
times identical, and j(t) within 1e-7 of the oracle's own step control.
"""
import dataclasses
import sys

import numpy as np
import pytest

from oracle.oracle_py import Oracle
from sbewh_amd.propagator import SBEContext
from tests import cases

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("nk", [256, 16])
def test_resident_chain_matches_host_controller_and_oracle(nk):
    """the 1-D 2-band chain, full pulse (1024 outputs) on 256 k (16 blocks)
    and 16 k (one block)"""
    cfg = cases.chain(Nk=(nk, 1, 1))
    say(f"[gpu] resident integrator: chain1d {nk} k")
    o = Oracle(cfg)
    x0 = o.initial_state(o.fermi_level(cfg.occupied_bands))
    xo = x0.copy()
    ro = o.integrate(xo, cfg.output_count)
    g = SBEContext(cfg)
    xh = x0.copy()
    rh = g.integrate(xh, cfg.output_count, log_capacity=0)
    assert not rh["device_controller"]
    g.close()
    gd = SBEContext(dataclasses.replace(cfg, device_controller=True))
    xd = x0.copy()
    rd = gd.integrate(xd, cfg.output_count, log_capacity=0)
    say(f"  resident: {rd['steps']} steps, {rd['rhs_calls']} RHS; host {rh['steps']} / {rh['rhs_calls']}; "
        f"oracle {ro['steps']}; j rel host {rel(rd['j'], rh['j']):.2e} oracle {rel(rd['j'], ro['j']):.2e}")
    assert rd["device_controller"]
    assert np.array_equal(rd["t"], rh["t"])
    assert abs(rd["steps"] - rh["steps"]) <= 2
    assert rel(rd["j"], rh["j"]) <= 1e-7
    assert rel(xd, xh) <= 1e-7
    assert rel(rd["j"], ro["j"]) <= 1e-7
    # a second integration on the same context starts again from the identity
    xd2 = x0.copy()
    rd2 = gd.integrate(xd2, cfg.output_count, log_capacity=0)
    assert rd2["device_controller"]
    assert np.array_equal(rd2["j"], rd["j"]) and np.array_equal(xd2, xd)
    gd.close()
    o.close()


def test_resident_not_taken_for_2d_or_with_step_log():
    """2-D instances and integrations with a step log run on the host
    controller whatever the flag says"""
    say("[gpu] resident integrator: fallbacks")
    cfg = dataclasses.replace(cases.small_square(Nk=(15, 15, 1)), device_controller=True)
    g = SBEContext(cfg)
    o = Oracle(cfg)
    x = o.initial_state(o.fermi_level(cfg.occupied_bands))
    r = g.integrate(x.copy(), 8, log_capacity=0)
    assert not r["device_controller"]
    g.close()
    o.close()
    c1 = dataclasses.replace(cases.chain(), device_controller=True, output_count=16)
    g = SBEContext(c1)
    o = Oracle(c1)
    x = o.initial_state(o.fermi_level(c1.occupied_bands))
    r = g.integrate(x.copy(), 16)   # with a step log
    assert not r["device_controller"]
    g.close()
    o.close()
