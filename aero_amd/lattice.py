"""Lattice constants and host-side helpers (D2Q9, D3Q19, MRT bases).

Mirrors the reference's public lattice modules so callers can import the same
names: aero/lbm/d2q9.py, aero/lbm/d3q19.py, aero/lbm/trt2d.py,
aero/lbm/mrt2d.py, aero/lbm/mrt3d.py, aero/lbm/physics.py.  The step itself
runs in the HIP library; these helpers serve setup, checkpoints and tests.
"""

from __future__ import annotations

import numpy as np

# ---- D2Q9 (aero/lbm/d2q9.py:15-41); indexing 6 2 5 / 3 0 1 / 7 4 8
Q = 9
E = np.array([[0, 0], [1, 0], [0, 1], [-1, 0], [0, -1],
              [1, 1], [-1, 1], [-1, -1], [1, -1]], dtype=np.int8)
W = np.array([4 / 9, 1 / 9, 1 / 9, 1 / 9, 1 / 9, 1 / 36, 1 / 36, 1 / 36, 1 / 36], dtype=np.float64)
OPP = np.array([0, 3, 4, 1, 2, 7, 8, 5, 6], dtype=np.int8)
CS2 = 1.0 / 3.0

# ---- D3Q19 (aero/lbm/d3q19.py:34-76)
Q3 = 19
E3 = np.array([
    [0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
    [1, 1, 0], [-1, 1, 0], [1, -1, 0], [-1, -1, 0],
    [1, 0, 1], [-1, 0, 1], [1, 0, -1], [-1, 0, -1],
    [0, 1, 1], [0, -1, 1], [0, 1, -1], [0, -1, -1]], dtype=np.int8)
W3 = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12, dtype=np.float64)
OPP3 = np.array([0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15], dtype=np.int8)
Y_MIR3 = np.array([0, 1, 2, 4, 3, 5, 6, 9, 10, 7, 8, 11, 12, 13, 14, 16, 15, 18, 17], dtype=np.int8)
CS2_3 = 1.0 / 3.0

# ---- MRT bases.  D2Q9: Lallemand & Luo (aero/lbm/mrt2d.py:36-48).
M9 = np.array([
    [1, 1, 1, 1, 1, 1, 1, 1, 1],
    [-4, -1, -1, -1, -1, 2, 2, 2, 2],
    [4, -2, -2, -2, -2, 1, 1, 1, 1],
    [0, 1, 0, -1, 0, 1, -1, -1, 1],
    [0, -2, 0, 2, 0, 1, -1, -1, 1],
    [0, 0, 1, 0, -1, 1, 1, -1, -1],
    [0, 0, -2, 0, 2, 1, 1, -1, -1],
    [0, 1, -1, 1, -1, 0, 0, 0, 0],
    [0, 0, 0, 0, 0, 1, -1, 1, -1]], dtype=np.float64)
M9_inv = np.linalg.inv(M9)

# D3Q19 basis as written in aero/lbm/mrt3d.py:46-87.  Its rows 14-16 are
# combinations of the momentum and energy-flux rows, so the matrix has rank 16
# and the reference inverts it with the Moore-Penrose pseudo-inverse
# (mrt3d.py:89-91); the same pseudo-inverse is used here so the collision
# operator (including its momentum projection) matches.
M19 = np.array([
    [1] * 19,
    [-30, -11, -11, -11, -11, -11, -11, 8, 8, 8, 8, 8, 8, 8, 8, 8, 8, 8, 8],
    [12, -4, -4, -4, -4, -4, -4, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1],
    [0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0],
    [0, -4, 4, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0],
    [0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1],
    [0, 0, 0, -4, 4, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1],
    [0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1],
    [0, 0, 0, 0, 0, -4, 4, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1],
    [0, 2, 2, -1, -1, -1, -1, 1, 1, 1, 1, 1, 1, 1, 1, -2, -2, -2, -2],
    [0, 0, 0, 1, 1, -1, -1, 1, 1, 1, 1, -1, -1, -1, -1, 0, 0, 0, 0],
    [0, 0, 0, 0, 0, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 0, 0, 0, 0],
    [0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, -1, -1, 1],
    [0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0],
    [0, -2, 2, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0],
    [0, 0, 0, -2, 2, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1],
    [0, 0, 0, 0, 0, -2, 2, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1],
    [0, 0, 0, 0, 0, 0, 0, 1, -1, 1, -1, -1, 1, -1, 1, 0, 0, 0, 0],
    [0, 0, 0, 0, 0, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, -1, 1, -1, 1]], dtype=np.float64)
M19_inv = np.linalg.pinv(M19)


def build_s_vec(omega: float, se: float = 1.64, sep: float = 1.54) -> np.ndarray:
    """D2Q9 MRT relaxation rates (aero/lbm/mrt2d.py:51-77)."""
    sv = omega
    sq = 8.0 * (2.0 - sv) / (8.0 - sv)
    return np.array([1.0, se, sep, 1.0, sq, 1.0, sq, sv, sv], dtype=np.float64)


def build_s3_vec(omega: float, se: float = 1.19, sep: float = 1.40, sq: float = 1.20) -> np.ndarray:
    """D3Q19 MRT relaxation rates (aero/lbm/mrt3d.py:94-135)."""
    sv = omega
    return np.array([1.0, se, sep, 1.0, sq, 1.0, sq, 1.0, sq, sv, sv, sv, sv, sv,
                     1.0, 1.0, 1.0, 1.0, 1.0], dtype=np.float64)


def trt_taus(omega: float, magic_lambda: float = 0.25) -> tuple[float, float]:
    """(tau+, tau-) of the reference TRT (aero/lbm/trt2d.py:8-18)."""
    tau = 1.0 / omega
    lam = float(magic_lambda)
    tau_plus = lam * tau + (1.0 - lam)
    tau_minus = (tau_plus * tau - 0.5) / (tau_plus - 0.5)
    return tau_plus, tau_minus


def base_nu_from_omega(omega: float) -> float:
    """aero/lbm/physics.py:27-29"""
    return (1.0 / omega - 0.5) / 3.0


def compute_feq(rho, ux, uy) -> np.ndarray:
    """D2Q9 equilibrium (aero/lbm/d2q9.py:66-89)."""
    usq = ux * ux + uy * uy
    feq = np.empty((Q,) + np.shape(rho), dtype=np.float64)
    for i in range(Q):
        eu = E[i, 0] * ux + E[i, 1] * uy
        feq[i] = W[i] * rho * (1.0 + 3.0 * eu + 4.5 * eu * eu - 1.5 * usq)
    return feq


def compute_feq_3d(rho, ux, uy, uz) -> np.ndarray:
    """D3Q19 equilibrium (aero/lbm/d3q19.py:105-129)."""
    usq = ux * ux + uy * uy + uz * uz
    feq = np.empty((Q3,) + np.shape(rho), dtype=np.float64)
    for i in range(Q3):
        eu = E3[i, 0] * ux + E3[i, 1] * uy + E3[i, 2] * uz
        feq[i] = W3[i] * rho * (1.0 + 3.0 * eu + 4.5 * eu * eu - 1.5 * usq)
    return feq


def compute_macroscopic(f: np.ndarray):
    """rho, ux, uy of a (9, Ny, Nx) field (aero/lbm/d2q9.py:44-63)."""
    rho = f.sum(axis=0)
    inv = np.where(rho > 0, 1.0 / np.where(rho > 0, rho, 1.0), 0.0)
    ux = inv * np.einsum("i,iyx->yx", E[:, 0].astype(np.float64), f)
    uy = inv * np.einsum("i,iyx->yx", E[:, 1].astype(np.float64), f)
    return rho, ux, uy


def compute_macroscopic_3d(f: np.ndarray):
    """rho, ux, uy, uz of a (19, Nz, Ny, Nx) field (aero/lbm/d3q19.py:80-102)."""
    rho = f.sum(axis=0)
    inv = np.where(rho > 0, 1.0 / np.where(rho > 0, rho, 1.0), 0.0)
    ux = inv * np.einsum("i,izyx->zyx", E3[:, 0].astype(np.float64), f)
    uy = inv * np.einsum("i,izyx->zyx", E3[:, 1].astype(np.float64), f)
    uz = inv * np.einsum("i,izyx->zyx", E3[:, 2].astype(np.float64), f)
    return rho, ux, uy, uz


def stream_inverse(g: np.ndarray, E_: np.ndarray) -> np.ndarray:
    """Inverse of the periodic push stream: out[i](x) = g[i](x + e_i)."""
    out = np.empty_like(g)
    nd = g.ndim - 1
    for i in range(g.shape[0]):
        shifts = tuple(-int(E_[i, a]) for a in reversed(range(nd)))
        out[i] = np.roll(g[i], shifts, axis=tuple(range(nd)))
    return out
