"""CPU oracle for the LBM step path -- TEST INFRASTRUCTURE ONLY.

This module composes the float64 C restatement in ``lbm_oracle.c`` into the
reference's step, in the reference's exact order:

* 2D ``Solver._step``   aero/lbm/solver.py:276-464
* 3D ``Solver3D._step`` aero/lbm/solver3d.py:325-541

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg may import it, and only as the checker (or the timed CPU
baseline).  The product package never imports it.

Pinning (see tests/test_oracle.py):
* every composed step is compared with golden vectors written from the
  reference's own NumPy path (tests/golden/make_golden.py) -- f, g,
  f_outlet_prev after 1 and 40 (2D) / 1 and 6 (3D) steps and every per-step
  coefficient, for BGK/TRT/MRT, bounce-back/Bouzidi, all inlets, outlets,
  sponge, walls, LES (Smagorinsky, WALE, Van Driest), thermal + buoyancy,
  SEM, synthetic inflow, recycling and body force;
* lattice tables and feq against closed-form moments; MRT inverses against
  numpy's inv/pinv; links against brute force on small grids;
* the local Smagorinsky closure (no reference counterpart) against the
  Chapman-Enskog identity: populations built as feq + f^neq with
  Pi^neq = -2 rho c_s^2 tau* S for a chosen strain S must return exactly
  tau* = tau0 + 3 C^2 |S|, and C = 0 must return tau0.
"""

from __future__ import annotations

import ctypes
import os
import pathlib
from typing import Optional

import numpy as np

_HERE = pathlib.Path(__file__).resolve().parent
_LIB_PATH = _HERE / "_build" / "liblbm_oracle.so"

# ------------------------------------------------------------------ lattices
# D2Q9: aero/lbm/d2q9.py:18-41
E9 = np.array([[0, 0], [1, 0], [0, 1], [-1, 0], [0, -1],
               [1, 1], [-1, 1], [-1, -1], [1, -1]], dtype=np.int32)
W9 = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4, dtype=np.float64)
OPP9 = np.array([0, 3, 4, 1, 2, 7, 8, 5, 6], dtype=np.int32)
# D3Q19: aero/lbm/d3q19.py:36-73
E19 = np.array([
    [0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
    [1, 1, 0], [-1, 1, 0], [1, -1, 0], [-1, -1, 0],
    [1, 0, 1], [-1, 0, 1], [1, 0, -1], [-1, 0, -1],
    [0, 1, 1], [0, -1, 1], [0, 1, -1], [0, -1, -1]], dtype=np.int32)
W19 = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12, dtype=np.float64)
OPP19 = np.array([0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15], dtype=np.int32)

# MRT bases: aero/lbm/mrt2d.py:36-48 and aero/lbm/mrt3d.py:46-91
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
M9_INV = np.linalg.inv(M9)
M19_INV = np.linalg.pinv(M19)


def mrt_s_2d(omega, se=1.64, sep=1.54):
    """aero/lbm/mrt2d.py:51-77"""
    sq = 8.0 * (2.0 - omega) / (8.0 - omega)
    return np.array([1.0, se, sep, 1.0, sq, 1.0, sq, omega, omega])


def mrt_s_3d(omega, se=1.19, sep=1.40, sq=1.20):
    """aero/lbm/mrt3d.py:94-135"""
    return np.array([1.0, se, sep, 1.0, sq, 1.0, sq, 1.0, sq] + [omega] * 5 + [1.0] * 5)


def sponge_sigma(nx, thickness, sigma_max):
    """aero/lbm/sponge.py:8-21"""
    sigma = np.zeros(nx)
    if thickness <= 0 or sigma_max <= 0.0:
        return sigma
    start = max(nx - thickness, 0)
    for x in range(start, nx):
        frac = (x - start + 1) / max(thickness, 1)
        sigma[x] = sigma_max * frac * frac
    return sigma


def feq(E, W, rho, *u):
    """Equilibrium for every direction (d2q9.py:66-89, d3q19.py:105-129)."""
    usq = sum(c * c for c in u)
    out = np.empty((len(W),) + np.shape(rho))
    for i in range(len(W)):
        eu = sum(E[i, a] * u[a] for a in range(len(u)))
        out[i] = W[i] * rho * (1.0 + 3.0 * eu + 4.5 * eu * eu - 1.5 * usq)
    return out


# ----------------------------------------------------------------- library
def build_library(force: bool = False) -> pathlib.Path:
    """Compile lbm_oracle.c with gcc + OpenMP into oracle/_build/."""
    src = _HERE / "lbm_oracle.c"
    _LIB_PATH.parent.mkdir(exist_ok=True)
    if force or not _LIB_PATH.exists() or _LIB_PATH.stat().st_mtime < src.stat().st_mtime:
        import subprocess
        tmp = str(_LIB_PATH) + ".tmp"
        subprocess.check_call(["gcc", "-O2", "-fopenmp", "-fPIC", "-shared", "-std=c11",
                               "-D_GNU_SOURCE", str(src), "-o", tmp, "-lm"])
        os.replace(tmp, _LIB_PATH)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            build_library()
        _lib = ctypes.CDLL(str(_LIB_PATH))
    return _lib


def _p(a):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.c_void_p)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def build_links(solid: np.ndarray, phi: Optional[np.ndarray] = None):
    """Surface links and q (boundary.py:232-280, boundary3d.py:279-330).

    Returns links with one column per axis as the reference does:
    2D (L, 3) = (i, y, x); 3D (L, 4) = (i, z, y, x).
    """
    is3d = solid.ndim == 3
    s3 = solid if is3d else solid[None]
    nz, ny, nx = s3.shape
    E = E19 if is3d else E9
    ex, ey = _i32(E[:, 0]), _i32(E[:, 1])
    ez = _i32(E[:, 2]) if is3d else np.zeros(9, np.int32)
    sol = np.ascontiguousarray(s3, dtype=np.uint8)
    L = lib()
    L.orc_build_links.restype = ctypes.c_longlong
    n = L.orc_build_links(len(E), nz, ny, nx, _p(sol), _p(ex), _p(ey), _p(ez), None,
                          ctypes.c_longlong(0))
    out = np.zeros((n, 4), np.int32)
    L.orc_build_links(len(E), nz, ny, nx, _p(sol), _p(ex), _p(ey), _p(ez), _p(out),
                      ctypes.c_longlong(n))
    links = out if is3d else out[:, [0, 2, 3]].copy()
    if phi is not None and n:
        idx = tuple(links[:, k] for k in range(1, links.shape[1]))
        q = np.clip(phi[idx].astype(np.float32), 1e-4, 1.0)
    else:
        q = np.full(n, 0.5, dtype=np.float32)
    return links, q


def link_fractions(links: np.ndarray, phi0: np.ndarray) -> np.ndarray:
    """Linear-Bouzidi wall fractions q = phi(x) / (phi(x) - phi(x + e_i)) from a
    signed distance sampled at integer cell coordinates (not a reference
    operator; pinned by the closed form for planes and spheres in
    tests/test_oracle.py)."""
    is3d = links.shape[1] == 4
    E = E19 if is3d else E9
    ph = phi0 if is3d else phi0[None]
    out = np.empty(len(links), dtype=np.float32)
    for k, row in enumerate(links):
        i = int(row[0])
        z, y, x = (int(row[1]), int(row[2]), int(row[3])) if is3d else (0, int(row[1]), int(row[2]))
        e = E[i]
        zn = min(max(z + (int(e[2]) if is3d else 0), 0), ph.shape[0] - 1)
        yn = min(max(y + int(e[1]), 0), ph.shape[1] - 1)
        xn = min(max(x + int(e[0]), 0), ph.shape[2] - 1)
        a, b = float(ph[z, y, x]), float(ph[zn, yn, xn])
        q = a / (a - b) if a != b else 0.5
        out[k] = np.float32(min(max(q, 1e-4), 1.0))
    return out


# --------------------------------------------------------------- the solver
class OracleSolver:
    """Reference-order float64 step, 2D (nz = None) or 3D.

    Parameters mirror the reference constructors (solver.py:75-116,
    solver3d.py:71-119).  ``f`` is the reference's post-boundary,
    pre-collision state.  Synthetic inflow and SEM planes are supplied by the
    caller through ``inlet_planes`` (a callable returning per-step planes), so
    the oracle never depends on a random generator.
    """

    def __init__(self, solid, omega, u0, D, *, rho0=1.0, wall_bc="slip", inlet_bc="velocity",
                 outlet_bc="convective", streamwise_bc="open", rho_in=1.0, rho_out=1.0,
                 collision="bgk", inlet_perturbation=0.0, trt_lambda=0.25, sponge_thickness=0,
                 sponge_strength=0.1, les=False, les_cs=0.16, les_model="smagorinsky", phi=None,
                 bouzidi=False, van_driest=False, van_driest_A=25.0, body_force=(0.0, 0.0, 0.0),
                 wall_velocity_top=0.0, wall_velocity_bottom=0.0, inlet_mode="zouhe",
                 thermal=False, T_hot=1.0, T_cold=0.0, alpha_T=1e-3, buoyancy=False,
                 g_gravity=0.0, beta=1e-3, T_ref=0.5, links=None, q_vals=None, trt_variant="reference",
                 bouzidi_variant="reference", ibm=False, les_form="fd"):
        self.is3d = solid.ndim == 3
        self.solid = np.ascontiguousarray(solid, dtype=bool)
        s3 = self.solid if self.is3d else self.solid[None]
        self.nz, self.ny, self.nx = s3.shape
        self.solid_u8 = np.ascontiguousarray(s3, dtype=np.uint8)
        E = E19 if self.is3d else E9
        self.Q = len(E)
        self.ex, self.ey = _i32(E[:, 0]), _i32(E[:, 1])
        self.ez = _i32(E[:, 2]) if self.is3d else np.zeros(9, np.int32)
        self.w = (W19 if self.is3d else W9).copy()
        self.opp = (OPP19 if self.is3d else OPP9).copy()
        self.omega, self.u0, self.D, self.rho0 = float(omega), float(u0), float(D), float(rho0)
        self.wall_bc, self.inlet_bc, self.outlet_bc = wall_bc, inlet_bc, outlet_bc
        self.streamwise_bc = streamwise_bc
        self.rho_in, self.rho_out = float(rho_in), float(rho_out)
        self.collision = {"bgk": 0, "trt": 1, "mrt": 2}[collision]
        self.inlet_perturbation = max(float(inlet_perturbation), 0.0)
        self.trt_lambda = float(trt_lambda)
        self.trt_variant = {"reference": 0, "lambda": 1}[trt_variant]
        self.sigma = sponge_sigma(self.nx, int(sponge_thickness), float(sponge_strength))
        self.sponge_on = int(sponge_thickness) > 0
        self.les, self.les_cs = bool(les), float(les_cs)
        self.les_model = {"smagorinsky": 0, "wale": 1}[les_model]
        self.les_form = {"fd": 0, "local": 1}[les_form]
        self.phi = phi
        # Guo IBM (solver.py:280-285, 379-383): forcing of band cells before the
        # collision when phi is given; the surface links never bounce back
        self.ibm = bool(ibm)
        self.van_driest, self.van_driest_A = bool(van_driest), float(van_driest_A)
        self.body_force = tuple(float(v) for v in body_force)
        self.u_top, self.u_bot = float(wall_velocity_top), float(wall_velocity_bottom)
        self.inlet_mode = inlet_mode  # "zouhe" | "field" | "sem"
        self.thermal, self.buoyancy = bool(thermal), bool(buoyancy)
        self.T_hot, self.T_cold, self.T_ref = float(T_hot), float(T_cold), float(T_ref)
        self.omega_T = 1.0 / (3.0 * float(alpha_T) + 0.5)
        self.g_gravity, self.beta = float(g_gravity), float(beta)
        self.tau = 1.0 / self.omega
        self.base_nu = (self.tau - 0.5) / 3.0
        if links is None:
            links, q_vals = build_links(self.solid, phi if bouzidi else None)
            if bouzidi and bouzidi_variant == "linear" and phi is not None:
                q_vals = link_fractions(links, phi)
        self.links_ref = links
        self.q_vals = q_vals
        l4 = np.zeros((len(links), 4), np.int32)
        if self.is3d:
            l4[:] = links
        else:
            l4[:, 0], l4[:, 2], l4[:, 3] = links[:, 0], links[:, 1], links[:, 2]
        self.links4 = l4
        self.bouzidi = bool(bouzidi)
        self.bouzidi_variant = {"reference": 1, "linear": 2}[bouzidi_variant]
        if self.collision == 2:
            self.M = M19 if self.is3d else M9
            self.Minv = M19_INV if self.is3d else M9_INV
            self.s = mrt_s_3d(self.omega) if self.is3d else mrt_s_2d(self.omega)
        else:
            self.M = self.Minv = self.s = None
        # state
        shape = (self.Q, self.nz, self.ny, self.nx)
        rho = np.full(shape[1:], self.rho0)
        zeros = np.zeros(shape[1:])
        vel = (np.full(shape[1:], self.u0),) + (zeros,) * (2 if self.is3d else 1)
        self.f = feq(E, self.w, rho, *vel)
        self.f_outlet_prev = self.f[:, :, :, -1].copy()
        self.g = None
        if self.thermal:
            self.g = np.empty(shape)
            for i in range(self.Q):
                self.g[i] = self.w[i] * self.T_ref
        ys = np.where(self.solid)
        if self.is3d:
            cz, cy, cx = (float(np.mean(a)) if a.size else 0.5 * (n - 1)
                          for a, n in zip(ys, (self.nz, self.ny, self.nx)))
        else:
            cy, cx = (float(np.mean(a)) if a.size else 0.5 * (n - 1)
                      for a, n in zip(ys, (self.ny, self.nx)))
            cz = 0.0
        self.center = (cx, cy, cz)
        self.step_count = 0
        self.last = {}
        self._wall_tables()

    # walls -----------------------------------------------------------------
    def _wall_tables(self):
        """Slip/noslip/moving pair tables (boundary.py:162-226, boundary3d.py:205-272)."""
        if self.is3d:
            bot = {"slip": [(3, 4), (7, 9), (8, 10), (15, 16), (17, 18)],
                   "noslip": [(3, 4), (7, 10), (8, 9), (15, 18), (17, 16)]}
            top = {"slip": [(4, 3), (9, 7), (10, 8), (16, 15), (18, 17)],
                   "noslip": [(4, 3), (10, 7), (9, 8), (18, 15), (16, 17)]}
            sgn_b = [0, 1, -1, 0, 0]
            sgn_t = [0, -1, 1, 0, 0]
        else:
            bot = {"slip": [(2, 4), (5, 8), (6, 7)], "noslip": [(2, 4), (5, 7), (6, 8)]}
            top = {"slip": [(4, 2), (8, 5), (7, 6)], "noslip": [(4, 2), (7, 5), (8, 6)]}
            sgn_b = [0, 1, -1]
            sgn_t = [0, -1, 1]
        kind = "noslip" if self.wall_bc == "moving" else self.wall_bc
        if kind not in bot:
            self.walls = None
            return
        b, t = bot[kind], top[kind]
        zeros = [0] * len(b)
        self.walls = (len(b), _i32([p[0] for p in b]), _i32([p[1] for p in b]),
                      _i32(sgn_b if self.wall_bc == "moving" else zeros),
                      _i32([p[0] for p in t]), _i32([p[1] for p in t]),
                      _i32(sgn_t if self.wall_bc == "moving" else zeros))

    # helpers ---------------------------------------------------------------
    def _dims(self):
        return self.Q, self.nz, self.ny, self.nx

    def macroscopic(self, f):
        Q, nz, ny, nx = self._dims()
        n = nz * ny * nx
        rho, ux, uy, uz = (np.empty(n) for _ in range(4))
        lib().orc_macroscopic(Q, nz, ny, nx, _p(f), _p(self.ex), _p(self.ey), _p(self.ez),
                              _p(rho), _p(ux), _p(uy), _p(uz))
        sh = (nz, ny, nx)
        return rho.reshape(sh), ux.reshape(sh), uy.reshape(sh), uz.reshape(sh)

    def omega_field(self, f):
        """LES relaxation field from the pre-collision state (les.py:273-344),
        or, for les_form "local", the local Smagorinsky closure on the
        non-equilibrium stress of f (orc_les_local_omega)."""
        Q, nz, ny, nx = self._dims()
        damp = None
        if self.van_driest and self.phi is not None:
            ph = self.phi if self.is3d else self.phi[None]
            damp = _f64((1.0 - np.exp(-np.abs(ph) / self.van_driest_A)) ** 2)
        om = np.empty((nz, ny, nx))
        if self.les_form == 1:
            lib().orc_les_local_omega(Q, nz, ny, nx, _p(_f64(f)), _p(self.solid_u8), _p(self.ex), _p(self.ey),
                                      _p(self.ez), ctypes.c_double(self.omega), ctypes.c_double(self.les_cs),
                                      _p(damp), _p(om))
            return om
        _, ux, uy, uz = self.macroscopic(f)
        lib().orc_les_omega(3 if self.is3d else 2, nz, ny, nx, _p(ux), _p(uy), _p(uz),
                            _p(self.solid_u8), ctypes.c_double(self.base_nu),
                            ctypes.c_double(self.les_cs), self.les_model, _p(damp), _p(om))
        return om

    def collide(self, f, omega_field=None):
        Q, nz, ny, nx = self._dims()
        out = np.empty_like(f)
        lib().orc_collide(Q, nz, ny, nx, _p(f), _p(out), _p(self.solid_u8), _p(self.ex),
                          _p(self.ey), _p(self.ez), _p(self.w), _p(self.opp), self.collision,
                          ctypes.c_double(self.omega), _p(omega_field),
                          ctypes.c_double(self.trt_lambda), _p(self.M), _p(self.Minv), _p(self.s),
                          self.trt_variant)
        return out

    def stream(self, src):
        Q, nz, ny, nx = self._dims()
        dst = np.empty_like(src)
        lib().orc_stream(Q, nz, ny, nx, _p(src), _p(dst), _p(self.ex), _p(self.ey), _p(self.ez))
        return dst

    # the step ----------------------------------------------------------------
    def step(self, plane=None):
        """One reference step; returns the raw force vector (12 floats).

        ``plane``: (ux, uy[, uz]) target planes for inlet_mode "field", or the
        SEM fluctuation triple (u', v', w') for "sem".
        """
        Q, nz, ny, nx = self._dims()
        L = lib()
        d = ctypes.c_double
        f = _f64(self.f)
        if self.ibm and self.phi is not None:
            f = f.copy()
            phi3 = _f64(np.asarray(self.phi, dtype=np.float64).reshape(nz, ny, nx))
            L.orc_ibm_guo(Q, nz, ny, nx, _p(f), _p(phi3), _p(self.solid_u8), d(self.tau), _p(self.ex), _p(self.ey),
                          _p(self.ez), _p(self.w))
        om = None
        if self.les and self.collision != 2:  # MRT never receives omega_field
            om = self.omega_field(f)
        fpre = self.collide(f, om)
        fnew = self.stream(fpre)
        if self.is3d and any(abs(v) > 0.0 for v in self.body_force):
            L.orc_body_force(Q, nz, ny, nx, _p(fnew), _p(self.solid_u8), d(self.tau),
                             d(self.body_force[0]), d(self.body_force[1]), d(self.body_force[2]),
                             _p(self.ex), _p(self.ey), _p(self.ez), _p(self.w))
        if not self.ibm:
            L.orc_bounce_back(Q, nz, ny, nx, _p(fnew), _p(fpre), ctypes.c_longlong(len(self.links4)),
                              _p(self.links4), _p(self.q_vals) if self.bouzidi else None,
                              _p(self.ex), _p(self.ey), _p(self.ez), _p(self.opp), self.bouzidi_variant)
        step_no = self.step_count + 1
        if self.is3d:
            if self.streamwise_bc == "open":
                if self.inlet_bc == "velocity":
                    mode = {"zouhe": 0, "field": 1, "sem": 2}[self.inlet_mode]
                    pu = pv = pw = None
                    if mode:
                        pu, pv, pw = (_f64(a) for a in plane)
                    L.orc_inlet_3d(nz, ny, nx, _p(fnew), mode, d(self.u0), d(0.0),
                                   d(self.inlet_perturbation), ctypes.c_longlong(step_no),
                                   _p(pu), _p(pv), _p(pw))
                self._outlet(fnew)
            elif self.streamwise_bc == "recycling":
                L.orc_recycling_3d(nz, ny, nx, _p(fnew), d(self.u0), _p(self.ex), _p(self.ey),
                                   _p(self.ez), _p(self.w))
        else:
            if self.inlet_bc == "velocity":
                if self.inlet_mode == "field":
                    ux_t, uy_t = (_f64(a) for a in plane)
                    L.orc_inlet_velocity_field_2d(ny, nx, _p(fnew), _p(ux_t), _p(uy_t))
                else:
                    L.orc_inlet_zouhe_2d(ny, nx, _p(fnew), d(self.u0), d(self.inlet_perturbation),
                                         ctypes.c_longlong(step_no))
            elif self.inlet_bc == "pressure":
                L.orc_inlet_pressure_2d(ny, nx, _p(fnew), d(self.rho_in))
            if self.outlet_bc == "pressure":
                L.orc_outlet_pressure_2d(ny, nx, _p(fnew), d(self.rho_out))
            else:
                self._outlet(fnew)
        if self.sponge_on:
            L.orc_sponge(Q, nz, ny, nx, _p(fnew), _p(self.sigma), d(self.u0), _p(self.ex), _p(self.w))
        if self.walls is not None:
            n, bd, bs, bg, td, ts, tg = self.walls
            L.orc_walls(Q, nz, ny, nx, _p(fnew), n, _p(bd), _p(bs), _p(bg), _p(td), _p(ts), _p(tg),
                        d(self.u_top), d(self.u_bot))
        if self.thermal:
            gout = np.empty_like(fnew)
            L.orc_thermal(Q, nz, ny, nx, _p(fnew), _p(_f64(self.g)), _p(gout), _p(self.solid_u8),
                          _p(self.ex), _p(self.ey), _p(self.ez), _p(self.w), d(self.omega_T),
                          int(self.buoyancy), d(self.g_gravity), d(self.beta), d(self.T_ref),
                          d(self.T_hot), d(self.T_cold))
            self.g = gout
        self.f = fnew
        self.step_count += 1
        out = np.zeros(12)
        nprof = nz if self.is3d else ny
        prof = np.zeros((nprof, 3))
        L.orc_forces(Q, nz, ny, nx, _p(fpre), _p(fnew), ctypes.c_longlong(len(self.links4)),
                     _p(self.links4), _p(self.ex), _p(self.ey), _p(self.ez), _p(self.w),
                     _p(self.opp), d(self.center[0]), d(self.center[1]), d(self.center[2]),
                     _p(out), _p(prof))
        self.last = dict(forces=out, profile=prof)
        return out

    def _outlet(self, fnew):
        Q, nz, ny, nx = self._dims()
        if self.outlet_bc == "convective":
            prev = _f64(self.f_outlet_prev)
            lib().orc_outlet_convective(Q, nz, ny, nx, _p(fnew), _p(prev), ctypes.c_double(self.u0))
            self.f_outlet_prev = prev
        elif self.outlet_bc == "zerogradient" or self.is3d:
            lib().orc_outlet_zero_gradient(Q, nz, ny, nx, _p(fnew))

    # coefficients -------------------------------------------------------------
    def coefficients(self, raw):
        """Cd, Cl(y), Clz, Cmx, Cmy, Cmz, Cd_p, Cd_v from raw forces
        (forces.py:78-139 with D and D^2; forces3d.py:364-439 with D^2 and D^3)."""
        rho0, u0, D = self.rho0, self.u0, self.D
        if self.is3d:
            fd = 0.5 * rho0 * u0 * u0 * D * D
            md = 0.5 * rho0 * u0 * u0 * D * D * D
        else:
            fd = 0.5 * rho0 * u0 * u0 * D
            md = 0.5 * rho0 * u0 * u0 * D * D
        if fd == 0.0:
            return np.zeros(8)
        return np.array([raw[0] / fd, raw[1] / fd, raw[2] / fd, raw[3] / md, raw[4] / md,
                         raw[5] / md, raw[6] / fd, raw[9] / fd])

    # field views in the reference's own shapes ----------------------------------
    def f_ref(self):
        return self.f if self.is3d else self.f[:, 0]

    def g_ref(self):
        return None if self.g is None else (self.g if self.is3d else self.g[:, 0])

    def prev_ref(self):
        return self.f_outlet_prev if self.is3d else self.f_outlet_prev[:, 0]

    def set_state(self, f, prev=None, g=None):
        self.f = _f64(f if self.is3d else f[:, None])
        self.f_outlet_prev = (self.f[:, :, :, -1].copy() if prev is None else
                              _f64(prev if self.is3d else prev[:, None]))
        if g is not None:
            self.g = _f64(g if self.is3d else g[:, None])


def kernel_only_seconds(solver: OracleSolver, steps: int) -> float:
    """Wall time of the collide + copy + stream part only, for the CPU
    baseline, timed inside the C library (orc_kernel_only) on copies of the
    solver's state; the solver itself is left unchanged."""
    Q, nz, ny, nx = solver._dims()
    f = np.array(solver.f, dtype=np.float64, order="C")
    a = np.empty_like(f)
    b = np.empty_like(f)
    a.fill(0.0)  # first touch of the work arrays outside the timed loop
    b.fill(0.0)
    sec = ctypes.c_double(0.0)
    lib().orc_kernel_only(Q, nz, ny, nx, _p(f), _p(a), _p(b), _p(solver.solid_u8), _p(solver.ex), _p(solver.ey),
                          _p(solver.ez), _p(solver.w), _p(solver.opp), solver.collision,
                          ctypes.c_double(solver.omega), ctypes.c_double(solver.trt_lambda), _p(solver.M),
                          _p(solver.Minv), _p(solver.s), solver.trt_variant, int(steps), ctypes.byref(sec))
    return float(sec.value)


def set_threads(n: int) -> None:
    """OpenMP threads of the library's following calls."""
    lib().orc_set_threads(int(n))
