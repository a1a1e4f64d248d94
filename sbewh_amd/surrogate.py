"""Learned surrogate: pulse parameters -> log10 yields of harmonics 1..40.

An ensemble of MLPs (6 -> 256 -> 256 -> 256 -> 40, SiLU) trained and evaluated
on the GPU by the library's own BF16 / FP32-accumulate kernels
(include/sbewh_surrogate.h), on spectra that the batched solver generates for
the 2-D 8-band model.  The solver stays in charge: queries outside the
training box, or where the members disagree by more than a tolerance, are
answered by the solver (one batched scan over all of them).

Inputs (in this order): E0 [V/nm], ellipticity, CEP [rad], wavelength [nm],
cycles, T2 [fs].  Yields: Y_h = sum of S(omega) over
[h w0 - w w0, h w0 + w w0) with w = 0.5, S the normalised high-harmonic
spectrum of j(t) (the reference's post-processing chain, sbewh_amd.spectrum).
"""
from __future__ import annotations

import argparse
import ctypes as C
import dataclasses
import json
import math
import os
import time

import numpy as np

from . import model, units
from ._lib import SbeError, load
from .spectrum import harmonic_yields, hhg_spectrum

PARAM_NAMES = ("E0_V_nm", "ellipticity", "cep", "lambda_nm", "cycles", "T2_fs")
RANGES = np.array([[0.3, 1.5], [0.0, 1.0], [0.0, 2 * math.pi], [2000.0, 4000.0], [6.0, 12.0], [5.0, 20.0]])
DIMS = (6, 256, 256, 256, 40)
HMAX = 40
YIELD_WIDTH = 0.5
LOG_FLOOR = 1e-30
DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")


# ---- the ensemble (ctypes over sbe_mlp_*) -------------------------------------
def _fp(a):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_float))


class MLPEnsemble:
    def __init__(self, dims=DIMS, members: int = 5, seed: int = 2005):
        L = load()
        self.dims = tuple(int(d) for d in dims)
        self.members = int(members)
        d = (C.c_int * len(self.dims))(*self.dims)
        h = C.c_void_p()
        rc = L.sbe_mlp_create(self.members, len(self.dims) - 1, d, C.c_ulonglong(seed), C.byref(h))
        if rc < 0:
            raise SbeError(f"sbe_mlp_create failed ({rc}): {L.sbe_mlp_last_error(None).decode()}")
        self.h = h
        self.P = int(L.sbe_mlp_param_count(h))

    def _check(self, rc):
        if rc < 0:
            raise SbeError(f"sbe_mlp error {rc}: {load().sbe_mlp_last_error(self.h).decode()}")

    def close(self):
        if getattr(self, "h", None):
            load().sbe_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self, member: int) -> np.ndarray:
        p = np.zeros(self.P, np.float32)
        self._check(load().sbe_mlp_get_params(self.h, member, _fp(p)))
        return p

    def set_params(self, member: int, p: np.ndarray):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.P
        self._check(load().sbe_mlp_set_params(self.h, member, _fp(p)))

    def forward(self, x: np.ndarray) -> np.ndarray:
        """[members, n, d_out]"""
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros((self.members, x.shape[0], self.dims[-1]), np.float32)
        self._check(load().sbe_mlp_forward(self.h, x.shape[0], _fp(x), _fp(y)))
        return y

    def gradient(self, x: np.ndarray, t: np.ndarray):
        """(loss [members], grad [members, P]) on one batch, no update"""
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        loss = np.zeros(self.members, np.float32)
        g = np.zeros((self.members, self.P), np.float32)
        self._check(load().sbe_mlp_gradient(self.h, x.shape[0], _fp(x), _fp(t), _fp(loss), _fp(g)))
        return loss, g

    def train(self, x: np.ndarray, t: np.ndarray, batch: int = 256, epochs: int = 200, lr: float = 2e-3,
              seed: int = 7):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        hist = np.zeros((epochs, self.members), np.float32)
        secs = C.c_double()
        self._check(load().sbe_mlp_train(self.h, x.shape[0], _fp(x), _fp(t), batch, epochs, C.c_float(lr),
                                         C.c_ulonglong(seed), _fp(hist), C.byref(secs)))
        return hist, secs.value


# ---- solver side: pulses, spectra, yields ------------------------------------
def sample_params(count: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return RANGES[:, 0] + rng.uniform(size=(count, 6)) * (RANGES[:, 1] - RANGES[:, 0])


def config_for(p, base=None, grid: int = 15, output_count: int = 1024):
    """solver configuration of one parameter vector on the 2-D 8-band model"""
    base = base or model.config("square2d")
    E0, eps, cep, lam, cyc, T2 = (float(v) for v in p)
    pul = model.elliptic_sin2(E0, eps, lam, cyc, cep=cep)
    return dataclasses.replace(base, name="surrogate", Nk=(grid, grid, 1), pulses=pul, T2_fs=T2,
                               output_count=output_count)


def log_yields(t: np.ndarray, j: np.ndarray, lambda_nm: float) -> np.ndarray:
    omega0 = units.omega_from_lambda_nm(lambda_nm)
    omax_eV = units.to_eV(omega0) * (HMAX + 1.0)
    w, S = hhg_spectrum(t, j, omega_max_eV=omax_eV)
    y = harmonic_yields(w, S, omega0, width=YIELD_WIDTH, hmax=HMAX)
    return np.log10(np.maximum(y, LOG_FLOOR))


def solve(params: np.ndarray, grid: int = 15, base=None, output_count: int = 1024, chunk: int = 0):
    """log10 yields [n, 40] of the solver (batched scans of `chunk` pulses,
    all in one when 0, with a progress line per scan), seconds, steps"""
    from .scan import Scan
    base = base or model.config("square2d")
    cfgs = [config_for(p, base, grid, output_count) for p in params]
    chunk = chunk or len(cfgs)
    t0 = time.perf_counter()
    ly, steps = [], []
    for a in range(0, len(cfgs), chunk):
        part = cfgs[a:a + chunk]
        sc = Scan(part)
        r = sc.run()
        sc.close()
        ly += [log_yields(r["t"][i], r["j"][i], part[i].pulses[0].lambda_nm) for i in range(len(part))]
        steps += list(r["steps"])
        if chunk < len(cfgs):
            print(f"[solve] {grid}^2 k: {a + len(part)} of {len(cfgs)} spectra, {time.perf_counter() - t0:.1f} s",
                  flush=True)
    return np.array(ly), time.perf_counter() - t0, np.array(steps)


# ---- the surrogate -------------------------------------------------------------
def normalise_inputs(p: np.ndarray) -> np.ndarray:
    return (2.0 * (np.asarray(p, float) - RANGES[:, 0]) / (RANGES[:, 1] - RANGES[:, 0]) - 1.0).astype(np.float32)


def in_hull(p: np.ndarray) -> np.ndarray:
    p = np.atleast_2d(p)
    return np.all((p >= RANGES[:, 0]) & (p <= RANGES[:, 1]), axis=1)


class Surrogate:
    """Ensemble + normalisation + fallback rule."""

    def __init__(self, ens: MLPEnsemble, y_mean: np.ndarray, y_std: np.ndarray, tol: float = 0.25):
        self.ens, self.y_mean, self.y_std, self.tol = ens, np.asarray(y_mean), np.asarray(y_std), tol

    @classmethod
    def fit(cls, params, logy, members=5, epochs=400, batch=256, lr=2e-3, seed=2005, tol=0.25):
        y_mean, y_std = logy.mean(0), logy.std(0) + 1e-6
        ens = MLPEnsemble(DIMS, members, seed)
        hist, secs = ens.train(normalise_inputs(params), ((logy - y_mean) / y_std).astype(np.float32), batch,
                               epochs, lr, seed + 1)
        s = cls(ens, y_mean, y_std, tol)
        s.history, s.train_seconds = hist, secs
        return s

    def predict(self, params):
        """(mean log10 yields [n, 40], member spread [n, 40] in log10 units)"""
        y = self.ens.forward(normalise_inputs(params)).astype(np.float64) * self.y_std + self.y_mean
        return y.mean(0), y.std(0)

    def query(self, params, grid: int = 15, base=None):
        """log10 yields for every row of params; rows outside the training box
        or with member spread above tol (log10 units) go to the solver.
        Returns (yields, used_solver mask, info)."""
        params = np.atleast_2d(np.asarray(params, float))
        t0 = time.perf_counter()
        mean, spread = self.predict(params)
        t_net = time.perf_counter() - t0
        fallback = ~in_hull(params) | (spread.max(1) > self.tol)
        out = mean.copy()
        t_solver = 0.0
        if fallback.any():
            ly, t_solver, _ = solve(params[fallback], grid, base)
            out[fallback] = ly
        return out, fallback, {"seconds_network": t_net, "seconds_solver": t_solver}

    def save(self, path):
        np.savez_compressed(path, params=np.stack([self.ens.get_params(e) for e in range(self.ens.members)]),
                            y_mean=self.y_mean, y_std=self.y_std, tol=self.tol, dims=np.array(DIMS))

    @classmethod
    def load(cls, path):
        z = np.load(path, allow_pickle=False)
        p = z["params"]
        ens = MLPEnsemble(tuple(int(v) for v in z["dims"]), p.shape[0])
        for e in range(p.shape[0]):
            ens.set_params(e, p[e])
        return cls(ens, z["y_mean"], z["y_std"], float(z["tol"]))


def relative_log_error(pred: np.ndarray, ref: np.ndarray) -> np.ndarray:
    """per-harmonic relative error of the log10 yield, |pred - ref| / |ref|"""
    return np.abs(pred - ref) / np.maximum(np.abs(ref), 1e-12)


def _main():
    ap = argparse.ArgumentParser(description="surrogate data generation / training")
    sub = ap.add_subparsers(dest="cmd", required=True)
    g = sub.add_parser("gen", help="solver spectra for sampled pulse parameters")
    g.add_argument("--count", type=int, default=1024)
    g.add_argument("--seed", type=int, default=5000)
    g.add_argument("--grid", type=int, default=15)
    g.add_argument("--chunk", type=int, default=0, help="pulses per batched scan (0: all in one)")
    g.add_argument("--out", required=True)
    k = sub.add_parser("kconv", help="k-grid convergence of the log10 yields on held-out pulses")
    k.add_argument("--count", type=int, default=32)
    k.add_argument("--seed", type=int, default=9000)
    k.add_argument("--grids", default="15,30,45,60,150")
    k.add_argument("--chunk", type=int, default=8)
    k.add_argument("--ensemble", default=os.path.join(DATA_DIR, "surrogate_ensemble.npz"))
    k.add_argument("--out", required=True)
    t = sub.add_parser("train", help="train the ensemble on generated data")
    t.add_argument("--data", nargs="+", required=True)
    t.add_argument("--test-fraction", type=float, default=0.125)
    t.add_argument("--epochs", type=int, default=400)
    t.add_argument("--batch", type=int, default=256)
    t.add_argument("--lr", type=float, default=2e-3)
    t.add_argument("--out", required=True)
    a = ap.parse_args()
    if a.cmd == "kconv":
        params = sample_params(a.count, a.seed)
        grids = [int(g) for g in a.grids.split(",")]
        res = {}
        for g in grids:
            ly, secs, steps = solve(params, g, chunk=a.chunk)
            res[g] = ly
            print(f"[kconv] {g}^2 k: {secs:.1f} s", flush=True)
        ref = res[max(grids)]
        s = Surrogate.load(a.ensemble)
        mean, _ = s.predict(params)
        report = {"count": a.count, "reference_grid": max(grids),
                  "median_abs_log10_vs_reference": {str(g): float(np.median(np.abs(res[g] - ref))) for g in grids},
                  "p90_abs_log10_vs_reference": {str(g): float(np.quantile(np.abs(res[g] - ref), 0.9))
                                                 for g in grids},
                  "surrogate_median_abs_log10_vs_reference": float(np.median(np.abs(mean - ref))),
                  "surrogate_median_rel_log10_vs_reference": float(np.median(relative_log_error(mean, ref)))}
        np.savez_compressed(a.out, params=params, grids=np.array(grids),
                            **{f"log_yields_{g}": res[g] for g in grids})
        print(json.dumps(report), flush=True)
    elif a.cmd == "gen":
        params = sample_params(a.count, a.seed)
        print(f"[gen] {a.count} spectra on {a.grid}^2 k-points, seed {a.seed}", flush=True)
        ly, secs, steps = solve(params, a.grid, chunk=a.chunk)
        np.savez_compressed(a.out, params=params, log_yields=ly, seconds=secs, steps=steps, grid=a.grid,
                            seed=a.seed)
        print(json.dumps({"count": a.count, "seconds": secs, "spectra_per_hour": a.count / secs * 3600,
                          "steps_mean": float(np.mean(steps))}), flush=True)
    else:
        P = np.concatenate([np.load(f, allow_pickle=False)["params"] for f in a.data])
        Y = np.concatenate([np.load(f, allow_pickle=False)["log_yields"] for f in a.data])
        n_test = int(len(P) * a.test_fraction)
        s = Surrogate.fit(P[n_test:], Y[n_test:], epochs=a.epochs, batch=a.batch, lr=a.lr)
        mean, spread = s.predict(P[:n_test])
        err = relative_log_error(mean, Y[:n_test])
        s.save(a.out)
        print(json.dumps({"train": len(P) - n_test, "test": n_test, "train_seconds": s.train_seconds,
                          "final_loss": s.history[-1].tolist(), "median_rel_log_error": float(np.median(err)),
                          "p90_rel_log_error": float(np.quantile(err, 0.9)),
                          "median_abs_log10_error": float(np.median(np.abs(mean - Y[:n_test])))}), flush=True)


if __name__ == "__main__":
    _main()
