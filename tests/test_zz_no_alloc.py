"""sbe_observe_device_noalloc (no device allocation; per-k partials in a
file-scope device array) against sbe_observe on the same state: identical
bits, on two calls in a row whose second has fewer k-points, and a clear
error for a grid beyond the array's size."""
import ctypes as C
import dataclasses
import sys

import numpy as np
import pytest

from sbewh_amd import hiprt
from sbewh_amd._lib import load
from sbewh_amd.propagator import SBEContext
from tests import cases

pytestmark = pytest.mark.gpu


def noalloc_observe(ctx, x):
    L = load()
    flat = x.view(np.float64)
    d_x = hiprt.DeviceBuffer.from_array(flat)
    d_j = hiprt.DeviceBuffer.from_array(np.full(3, np.nan))
    rc = L.sbe_observe_device_noalloc(ctx.h, C.c_void_p(d_x.ptr), C.c_void_p(d_j.ptr), None)
    assert rc == 0, L.sbe_last_error(ctx.h).decode()
    hiprt.synchronize()
    j = d_j.download(np.float64, 3)
    d_x.free()
    d_j.free()
    return j


def test_noalloc_observe_equals_observe_twice_smaller_second():
    for cfg in (cases.small_square(Nk=(30, 30, 1)), cases.chain()):
        print(f"[gpu] noalloc observe {cfg.name} Nk={cfg.Nk}", flush=True)
        sys.stdout.flush()
        g = SBEContext(cfg)
        st = g.initial_state(g.fermi_level(cfg.occupied_bands))
        x = cases.perturbed_state(st, g.n, g.P, g.num_k, seed=4)
        j_ref = g.observe(x)
        j = noalloc_observe(g, x)
        assert np.array_equal(j, j_ref), (j, j_ref)
        g.close()


def test_noalloc_rejects_grid_beyond_scratch():
    cfg = dataclasses.replace(cases.chain(), Nk=(262160, 1, 1))
    g = SBEContext(cfg)
    L = load()
    d_x = hiprt.DeviceBuffer(16)
    d_j = hiprt.DeviceBuffer(24)
    rc = L.sbe_observe_device_noalloc(g.h, C.c_void_p(d_x.ptr), C.c_void_p(d_j.ptr), None)
    assert rc < 0
    assert "at most 262144 k-points" in L.sbe_last_error(g.h).decode()
    g.close()
