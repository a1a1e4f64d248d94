"""The drop-in boundary on the GPU: the reference GPU path's unpadded state
layout (SBE_FLAG_GPU_LAYOUT, code/src/CuPropagator.cu:50-52,111-114), the
k-index map, the device-buffer RHS, and the C++ functor with the reference's
odeint System signature driven by a plain Dormand-Prince loop, all against
the oracle.  n = 2 (the 1-D chain) is the case where the two layouts differ
(n*n = 4 is padded to 8 in the CPU layout)."""
import ctypes as C
import dataclasses

import numpy as np
import pytest

from sbewh_amd.ctypes_types import Inputs, dptr
from sbewh_amd.model import Pulse
from sbewh_amd.propagator import SBEContext
from oracle.oracle_py import Oracle
from sbewh_amd import hiprt as hip_mem
from tests import cases, cpp_driver
from tests.test_oracle_dopri5 import uniform_log

pytestmark = pytest.mark.gpu


def chain_short(gpu_layout):
    pul = [Pulse("sin2", 0.5, 3200.0, 0.0, 5.3, t_start=-1, t_end=1)]
    return dataclasses.replace(cases.chain(sweeps=5), pulses=pul, output_count=17, gpu_layout=gpu_layout)


def to_cpu_layout(x_gpu, nk, n):
    """GPU layout (P = n*n) -> CPU layout (P = 8 ceil(n*n/8)), same k order"""
    P = 8 * ((n * n + 7) // 8)
    out = np.zeros(8 + nk * P, complex)
    out[:8] = x_gpu[:8]
    out[8:].reshape(nk, P)[:, :n * n] = x_gpu[8:].reshape(nk, n * n)
    return out


def to_gpu_layout(x_cpu, nk, n):
    P = 8 * ((n * n + 7) // 8)
    return np.concatenate([x_cpu[:8], x_cpu[8:].reshape(nk, P)[:, :n * n].reshape(-1)])


@pytest.mark.parametrize("name", ["chain", "square"])
def test_gpu_layout_rhs_equals_cpu_layout(name, monkeypatch):
    if name == "chain":
        cfg = cases.chain(sweeps=5)
        occ = 1
    else:
        cfg = cases.small_square(Nk=(15, 15, 1))
        occ = 4
    cc = SBEContext(cfg)
    cg = SBEContext(dataclasses.replace(cfg, gpu_layout=True))
    n, nk = cc.n, cc.num_k
    assert cg.P == n * n and cg.state_size == 8 + nk * n * n
    assert cc.P == 8 * ((n * n + 7) // 8)
    mu = cc.fermi_level(occ)
    assert cg.fermi_level(occ) == mu
    xc = cc.initial_state(mu)
    xg = cg.initial_state(mu)
    assert np.array_equal(to_gpu_layout(xc, nk, n), xg)
    x = cases.perturbed_state(xc, n, cc.P, nk, seed=5)
    t = 0.3 * cc.time_window[0]
    dc = cc.rhs(x, t)
    # bit-identical results need the same stage layout: each context times
    # the layouts on its first RHS, so the second takes the first's choice
    monkeypatch.setenv("SBEWH_STAGE_LR", str(cc.stage_tuning()["lanes_per_row"]))
    dg = cg.rhs(to_gpu_layout(x, nk, n), t)
    assert cg.stage_tuning()["lanes_per_row"] == cc.stage_tuning()["lanes_per_row"]
    assert np.array_equal(to_gpu_layout(dc, nk, n), dg)
    assert np.array_equal(cc.observe(x), cg.observe(to_gpu_layout(x, nk, n)))
    o = Oracle(cfg)
    do = o.rhs(x, t)
    assert np.abs(to_cpu_layout(dg, nk, n) - do).max() < 1e-12 * np.abs(do).max()


def test_k_index_map_matches_fourier_block_order():
    for cfg in (cases.chain(), cases.small_square(Nk=(30, 30, 1)), cases.small_cubic()):
        c = SBEContext(cfg)
        o = Oracle(cfg)
        g = c.k_index_map()
        Nk = np.array([cfg.Nk[d] if d < cfg.dim else 1 for d in range(3)])
        kf = np.array([o.kfrac(k) for k in range(o.num_k)])
        assert np.array_equal(g / Nk, kf)
        # a permutation of the grid
        flat = (g[:, 0] * Nk[1] + g[:, 1]) * Nk[2] + g[:, 2]
        assert np.array_equal(np.sort(flat), np.arange(o.num_k))
        c.close()


def test_rhs_device_gpu_layout_on_device_buffers():
    """sbe_rhs_device on caller-owned device buffers in the GPU layout (the
    CuPropagator GPUstate: interleaved doubles), enqueued on a caller stream,
    equals sbe_rhs on host buffers bit for bit"""
    cfg = chain_short(True)
    cg = SBEContext(cfg)
    n, nk = cg.n, cg.num_k
    x = cg.initial_state(cg.fermi_level(1))
    x[8:] += cases.random_hermitian(np.random.default_rng(3), nk, n, 0.01).reshape(-1)
    x[0] = 0.2
    t = 0.4 * cg.time_window[0]
    dx = cg.rhs(x, t)
    flat = x.view(np.float64)
    d_in = hip_mem.DeviceBuffer.from_array(flat)
    d_out = hip_mem.DeviceBuffer.from_array(np.zeros_like(flat))
    s = hip_mem.Stream()
    cg.rhs_device(d_in.ptr, d_out.ptr, t, s.handle)
    s.synchronize()
    assert np.array_equal(d_out.download(np.float64, flat.size), dx.view(np.float64))
    s.destroy()
    d_in.free()
    d_out.free()
    cg.close()


def test_integrate_gpu_layout_equals_cpu_layout(monkeypatch):
    cc, cg = SBEContext(chain_short(False)), SBEContext(chain_short(True))
    n, nk = cc.n, cc.num_k
    xc = cc.initial_state(cc.fermi_level(1))
    xg = to_gpu_layout(xc, nk, n)
    rc = cc.integrate(xc, 17)
    # the same stage layout both ways (each context times them on its first RHS)
    monkeypatch.setenv("SBEWH_STAGE_LR", str(cc.stage_tuning()["lanes_per_row"]))
    rg = cg.integrate(xg, 17)
    assert rc["steps"] == rg["steps"]
    assert np.array_equal(rc["j"], rg["j"])
    assert np.array_equal(to_gpu_layout(xc, nk, n), xg)


@pytest.mark.parametrize("gpu_layout", [True, False])
def test_cpp_functor_plain_dopri5_against_oracle(gpu_layout):
    """sbewh::PropagatorHIP (reference System signature) in a plain C++
    Dormand-Prince loop, 8 equal steps per output interval, against the oracle
    replaying the same steps: rho and j(t) to <= 1e-10 relative"""
    cfg = chain_short(gpu_layout)
    L = cpp_driver.load()
    inp = Inputs(cfg)
    o = Oracle(cfg)
    mu = o.fermi_level(1)
    t0, t1 = o.t_start, o.t_end
    outputs, m = cfg.output_count, 8
    dt_out = (t1 - t0) / (outputs - 1)
    n, nk = o.n, o.num_k
    size = 8 + nk * (n * n if gpu_layout else o.P)
    st = np.zeros(size, complex)
    j = np.zeros((outputs, 3))
    print(f"[cpp functor] chain n=2 gpu_layout={gpu_layout} outputs={outputs} m={m}", flush=True)
    rc = L.functor_dopri5(C.byref(inp.model), C.byref(inp.grid), C.byref(inp.params), inp.pulses, inp.num_pulses, mu,
                          t0, dt_out, outputs, m, dptr(st.view(np.float64)), dptr(j.reshape(-1)))
    assert rc == 0
    x = o.initial_state(mu)
    r = o.integrate(x, outputs, replay=uniform_log(t0, t1, outputs, m))
    st_cpu = to_cpu_layout(st, nk, n) if gpu_layout else st
    rho_g, rho_o = st_cpu[8:].reshape(nk, o.P)[:, :n * n], o.rho(x).reshape(nk, n * n)
    assert np.abs(rho_g - rho_o).max() < 1e-10 * np.abs(rho_o).max()
    # A returns to ~0 at the end of the pulse: scale by its peak along the pulse
    A_peak = max(np.abs(o.A(t)).max() for t in np.linspace(t0, t1, 201))
    assert np.abs(st_cpu[:3] - x[:3]).max() < 1e-10 * A_peak
    assert np.abs(j - r["j"]).max() < 1e-10 * np.abs(r["j"]).max()
