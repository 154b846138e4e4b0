"""Helpers shared by the parity tests: load golden fixtures and rebuild inputs.

The fixtures under tests/golden/ were written from the reference's own NumPy
path by tests/golden/make_golden.py.  They hold the seeded noise, geometry,
parameters and expected outputs; the initial state is rebuilt here from the
noise through the equilibrium closed form.
"""

from __future__ import annotations

import json
import pathlib

import numpy as np

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"

E9 = np.array([[0, 0], [1, 0], [0, 1], [-1, 0], [0, -1],
               [1, 1], [-1, 1], [-1, -1], [1, -1]])
W9 = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)
E19 = np.array([
    [0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
    [1, 1, 0], [-1, 1, 0], [1, -1, 0], [-1, -1, 0],
    [1, 0, 1], [-1, 0, 1], [1, 0, -1], [-1, 0, -1],
    [0, 1, 1], [0, -1, 1], [0, 1, -1], [0, -1, -1]])
W19 = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)

CASES_2D = sorted(p.stem[len("ref2d_"):] for p in GOLDEN.glob("ref2d_*.npz"))
CASES_3D = sorted(p.stem[len("ref3d_"):] for p in GOLDEN.glob("ref3d_*.npz"))
SNAP2 = (1, 40)
SNAP3 = (1, 6)


def load(dim: int, name: str) -> dict:
    with np.load(GOLDEN / f"ref{dim}d_{name}.npz", allow_pickle=False) as d:
        out = {k: d[k] for k in d.files}
    out["kw"] = json.loads(str(out.pop("params_json")))
    return out


def feq(E, W, rho, *u):
    usq = sum(c * c for c in u)
    out = np.empty((len(W),) + np.shape(rho))
    for i in range(len(W)):
        eu = sum(E[i, a] * u[a] for a in range(len(u)))
        out[i] = W[i] * rho * (1.0 + 3.0 * eu + 4.5 * eu * eu - 1.5 * usq)
    return out


def initial_f(case: dict) -> np.ndarray:
    """feq(1, u0, noise) -- the harness state written into solver.f."""
    u0 = float(case["u0"])
    if "noise_uz" in case:
        uy, uz = case["noise_uy"], case["noise_uz"]
        return feq(E19, W19, np.ones(uy.shape), np.full(uy.shape, u0), uy, uz)
    uy = case["noise_uy"]
    return feq(E9, W9, np.ones(uy.shape), np.full(uy.shape, u0), uy)


class SyntheticInflow:
    """Restates the synthetic inflow generators (aero/lbm/solver.py:259-270,
    aero/lbm/solver3d.py:297-321) for the oracle, which takes planes as input."""

    def __init__(self, shape, u0, intensity, seed=12345):
        self.rng = np.random.default_rng(seed)
        self.shape = shape
        self.u0, self.intensity = u0, intensity
        if len(shape) == 1:
            self.fy = np.zeros(shape)
        else:
            self.fy, self.fz = np.zeros(shape), np.zeros(shape)

    def next(self):
        target = self.intensity * self.u0
        if len(self.shape) == 1:
            noise = self.rng.standard_normal(self.shape[0])
            filt = (noise + np.roll(noise, 1) + np.roll(noise, -1)) / 3.0
            self.fy = 0.75 * self.fy + 0.25 * filt
            self.fy -= float(np.mean(self.fy))
            rms = float(np.sqrt(np.mean(self.fy ** 2)))
            if rms > 1e-12 and target > 0.0:
                self.fy *= target / rms
            return np.full(self.shape, self.u0), self.fy.copy()
        ny_, nz_ = self.rng.standard_normal(self.shape), self.rng.standard_normal(self.shape)

        def filt(a):
            return (a + np.roll(a, 1, axis=0) + np.roll(a, -1, axis=0)
                    + np.roll(a, 1, axis=1) + np.roll(a, -1, axis=1)) / 5.0
        self.fy = 0.8 * self.fy + 0.2 * filt(ny_)
        self.fz = 0.8 * self.fz + 0.2 * filt(nz_)
        self.fy -= float(np.mean(self.fy))
        self.fz -= float(np.mean(self.fz))
        rms = float(np.sqrt(np.mean(self.fy ** 2 + self.fz ** 2)))
        if rms > 1e-12 and target > 0.0:
            self.fy *= target / rms
            self.fz *= target / rms
        return np.full(self.shape, self.u0), self.fy.copy(), self.fz.copy()


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
