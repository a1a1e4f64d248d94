"""Pins of the oracle's Dormand-Prince controller (Boost.Odeint semantics).

* exact solution: with E = 0 and no dephasing rho_k(t) = e^{-iHt} rho_k(0) e^{iHt};
* order: fixed-step replays at h and h/2 show global error ratio 2^5 (a wrong
  tableau entry breaks fifth order) and embedded-estimate ratio 2^5;
* controller rule: every logged attempt obeys the restated accept / reject /
  grow / truncate rules; RHS count = 6 attempts + one fresh FSAL derivative
  per output interval; output times t0 + i dt_out.
"""
import dataclasses
import math

import numpy as np
import pytest
from scipy.linalg import expm

from sbewh_amd import units
from sbewh_amd.ctypes_types import StepLog, time_window
from sbewh_amd.model import Pulse
from oracle.oracle_py import Oracle
from tests import cases


def free_cfg(output_count=17, eps_rel=1e-10, eps_abs=1e-14):
    """chain model, zero field (E0 = 0), no dephasing, 16 k-points."""
    pul = [Pulse("sin2", 0.0, 3200.0, 0.0, 2.0)]
    return dataclasses.replace(cases.chain(sweeps=10), pulses=pul, T2_fs=-1.0, Nk=(16, 1, 1), Nf=(16, 1, 1),
                               output_count=output_count, eps_rel=eps_rel, eps_abs=eps_abs)


def exact(cfg, o, x0, T):
    out = x0.copy()
    for k in range(o.num_k):
        H = cfg.model.hk(o.kfrac(k))
        U = expm(-1j * H * T)
        rho = o.rho(x0)[k]
        out[8 + o.P * k: 8 + o.P * k + o.n * o.n] = (U @ rho @ U.conj().T).reshape(-1)
    return out


def start_state(o):
    st = o.initial_state(o.fermi_level(1))
    return cases.perturbed_state(st, o.n, o.P, o.num_k, seed=11, amp=0.1) * np.r_[np.zeros(8), np.ones(o.state_size - 8)]


def test_adaptive_matches_exact_solution():
    cfg = free_cfg()
    o = Oracle(cfg)
    x0 = start_state(o)
    x = x0.copy()
    r = o.integrate(x, cfg.output_count)
    T = o.t_end - o.t_start
    ref = exact(cfg, o, x0, T)
    assert np.abs(x - ref).max() < 1e-8
    t0, t1 = time_window(cfg)
    dt_out = (t1 - t0) / (cfg.output_count - 1)
    assert np.array_equal(r["t"], t0 + np.arange(cfg.output_count) * dt_out)


def uniform_log(t0, t1, outputs, m):
    """replay log with m equal steps per output interval."""
    dt_out = (t1 - t0) / (outputs - 1)
    h = dt_out / m
    log = StepLog(m * (outputs - 1) + 1)
    idx = 0
    for i in range(outputs - 1):
        a = t0 + i * dt_out
        for j in range(m):
            log.t[idx] = a + j * h
            log.dt[idx] = h
            log.accepted[idx] = 1
            idx += 1
    log.c.count = idx
    return log


def test_fifth_order_and_embedded_estimate():
    cfg = free_cfg(output_count=3)
    o = Oracle(cfg)
    x0 = start_state(o)
    t0, t1 = o.t_start, o.t_end
    ref = exact(cfg, o, x0, t1 - t0)
    errs, first_err = [], []
    for m in (40, 80):
        x = x0.copy()
        r = o.integrate(x, 3, replay=uniform_log(t0, t1, 3, m))
        errs.append(np.abs(x - ref).max())
        first_err.append(r["log"].err[0])
    ratio = errs[0] / errs[1]
    assert 24 < ratio < 40, ratio            # 2^5 = 32
    eratio = first_err[0] / first_err[1]
    assert 24 < eratio < 40, eratio          # local estimate O(h^5)


def test_controller_rules_and_rhs_count():
    cfg = dataclasses.replace(cases.chain(sweeps=5), Nk=(32, 1, 1), Nf=(16, 1, 1), output_count=65)
    o = Oracle(cfg)
    x = o.initial_state(o.fermi_level(1))
    r = o.integrate(x, cfg.output_count)
    log = r["log"]
    n = log.count
    t, dt, err, acc = log.t[:n], log.dt[:n], log.err[:n], log.accepted[:n]
    assert np.all(err[acc == 1] <= 1.0) and np.all(err[acc == 0] > 1.0)
    t0, t1 = o.t_start, o.t_end
    dt_out = (t1 - t0) / (cfg.output_count - 1)
    # interval ends as integrate_const forms them: (t0 + i dt_out) + dt_out
    ends = (t0 + np.arange(cfg.output_count - 1) * dt_out) + dt_out
    checked = 0
    for i in range(n - 1):
        if acc[i] == 0:
            expect = dt[i] * max(0.9 * err[i] ** (-1 / 3), 0.2)
            assert t[i + 1] == t[i] and math.isclose(dt[i + 1], expect, rel_tol=1e-15)
            checked += 1
        else:
            grow = dt[i] * 0.9 * max(err[i], 5.0 ** -5) ** (-1 / 5) if err[i] < 0.5 else dt[i]
            assert abs(t[i + 1] - (t[i] + dt[i])) < 1e-11  # interval starts are recomputed
            end = ends[np.searchsorted(ends, t[i + 1] + 1e-9)]
            truncated = dt[i + 1] == end - t[i + 1]
            assert truncated or math.isclose(dt[i + 1], grow, rel_tol=1e-15)
            checked += 1
    assert checked == n - 1
    assert r["rhs_calls"] == 6 * n + (cfg.output_count - 1)
    assert r["steps"] == int(acc.sum())


def test_replay_reproduces_run_bit_exactly():
    cfg = dataclasses.replace(cases.chain(sweeps=5), Nk=(32, 1, 1), Nf=(16, 1, 1), output_count=65)
    o = Oracle(cfg)
    x0 = o.initial_state(o.fermi_level(1))
    xa, xb = x0.copy(), x0.copy()
    ra = o.integrate(xa, cfg.output_count)
    rb = o.integrate(xb, cfg.output_count, replay=ra["log"])
    assert rb["steps"] == ra["steps"]
    assert np.array_equal(xa, xb) and np.array_equal(ra["j"], rb["j"])
    assert rb["rhs_calls"] == 6 * ra["steps"] + cfg.output_count - 1
