"""D2Q9 wind-tunnel solver on the MI355X step library.

Drop-in for the reference's ``aero.lbm.solver.Solver`` (aero/lbm/solver.py:74-727):
same constructor arguments, ``_step()`` / ``step()`` / ``run()``, public
state (``f``, ``g``, ``surface_links``, ``q_vals``, histories, ``step_count``)
and the same ``.npz`` checkpoint layout.  Every step runs in hand-written HIP
kernels (include/aero_hip.h); this class only prepares inputs and collects the
FP64 coefficient history the device reduces.
"""

from __future__ import annotations

import json
from typing import Callable, Optional

import numpy as np

from . import _abi
from ._base import (DeviceSolver, checkpoint_params, extra_params_entry, les_code, resolve_backend, run_loop,
                    warn_hdf5)
from .diagnostics import coefficient_spectrum, detect_statistical_stationarity
from .inflow import SyntheticInflow2D
from .lattice import M9_inv, build_s_vec, compute_feq, base_nu_from_omega
from .links import build_surface_links, link_fractions


class Solver(DeviceSolver):
    _Q = 9
    _is3d = False

    def __init__(self, Ny: int, Nx: int, solid: np.ndarray, omega: float, u0: float, D: float,
                 rho0: float = 1.0, wall_bc: str = "slip", inlet_bc: str = "velocity",
                 outlet_bc: str = "convective", rho_in: float = 1.0, rho_out: float = 1.0,
                 backend: str = "auto", collision: str = "bgk", inlet_perturbation: float = 0.0,
                 trt_lambda: float = 0.25, sponge_thickness: int = 0, sponge_strength: float = 0.1,
                 les: bool = False, les_cs: float = 0.16, les_model: str = "smagorinsky",
                 ibm_enabled: bool = False, phi: Optional[np.ndarray] = None, bouzidi: bool = False,
                 van_driest: bool = False, van_driest_A: float = 25.0, wall_velocity_top: float = 0.0,
                 wall_velocity_bottom: float = 0.0, synthetic_inflow: bool = False,
                 synthetic_inflow_intensity: float = 0.03, synthetic_inflow_seed: int = 12345,
                 thermal: bool = False, T_hot: float = 1.0, T_cold: float = 0.0, alpha_T: float = 1e-3,
                 buoyancy: bool = False, g_gravity: float = 0.0, beta: float = 1e-3, T_ref: float = 0.5,
                 storage: str = "f64", trt_variant: str = "reference",
                 bouzidi_variant: str = "reference", les_form: str = "fd"):
        """Arguments as the reference (solver.py:75-142), plus:

        storage     -- population format on the device: "f64" (reference
                       precision), "f32", "f16" or "bf16" (shifted half storage,
                       FP32 arithmetic; "bf16" rounds its stores
                       stochastically, from bits that depend only on the
                       cell, the step number and the population).
        trt_variant -- "reference" uses the reference's TRT rates
                       (trt2d.py:8-18), which are unstable for tau below ~0.76;
                       "lambda" uses the standard magic-parameter TRT
                       (tau+ = tau, tau- = trt_lambda/(tau - 1/2) + 1/2).
        bouzidi_variant -- "reference" keeps the reference's interpolated
                       bounce-back (q = phi at the fluid cell, boundary.py:303-352);
                       "linear" is the Bouzidi-Firdaouss-Lallemand rule with q the
                       link fraction from ``phi`` sampled at the integer cell
                       coordinates (e.g. ``geometry.sdf_field(..., offset=0.0)``).
        les_form    -- "fd" (the reference's finite-difference Smagorinsky /
                       WALE, les.py:273-344, a split step) or "local":
                       Smagorinsky from each cell's non-equilibrium stress,
                       fused into the step, tau = (tau0 + sqrt(tau0^2 +
                       18 C^2 |Pi^neq| / rho)) / 2 with C = les_cs times the
                       Van Driest factor; BGK and TRT (MRT ignores LES, as in
                       the reference)."""
        self.Ny, self.Nx = int(Ny), int(Nx)
        self.solid = np.ascontiguousarray(solid, dtype=np.bool_)
        if self.solid.shape != (self.Ny, self.Nx):
            raise ValueError(f"solid must have shape ({Ny}, {Nx})")
        self.fluid = ~self.solid
        self.omega, self.u0, self.D, self.rho0 = float(omega), float(u0), float(D), float(rho0)
        self.wall_bc, self.inlet_bc, self.outlet_bc = wall_bc, inlet_bc, outlet_bc
        self.rho_in, self.rho_out = float(rho_in), float(rho_out)
        self.inlet_perturbation = max(float(inlet_perturbation), 0.0)
        self.trt_lambda = float(trt_lambda)
        self.sponge_thickness, self.sponge_strength = int(sponge_thickness), float(sponge_strength)
        self.les, self.les_cs = bool(les), float(les_cs)
        if les_model not in ("smagorinsky", "wale"):
            raise ValueError("les_model must be 'smagorinsky' or 'wale'.")
        self.les_model = str(les_model)
        self._les_code = les_code(self.les, self.les_model, les_form)
        self.les_form = les_form
        self.ibm_enabled = bool(ibm_enabled)
        self.phi = phi
        self.bouzidi = bool(bouzidi)
        self.van_driest, self.van_driest_A = bool(van_driest), float(van_driest_A)
        self.wall_velocity_top, self.wall_velocity_bottom = float(wall_velocity_top), float(wall_velocity_bottom)
        self.synthetic_inflow = bool(synthetic_inflow)
        self.synthetic_inflow_intensity = float(synthetic_inflow_intensity)
        self._inflow = SyntheticInflow2D(self.Ny, self.u0, self.synthetic_inflow_intensity, synthetic_inflow_seed)
        self.thermal = bool(thermal)
        self.T_hot, self.T_cold, self.alpha_T = float(T_hot), float(T_cold), float(alpha_T)
        self.omega_T = float(1.0 / (3.0 * alpha_T + 0.5))
        self.buoyancy, self.g_gravity, self.beta, self.T_ref = bool(buoyancy), float(g_gravity), float(beta), float(T_ref)
        self.tau = 1.0 / omega
        self._base_nu = base_nu_from_omega(omega)
        if collision not in ("bgk", "mrt", "trt"):
            raise ValueError(f"Unknown collision '{collision}'. Choose 'bgk', 'mrt', or 'trt'.")
        self.collision = collision
        self.backend = resolve_backend(backend)
        if storage not in _abi.STORE:
            raise ValueError(f"Unknown storage '{storage}'. Choose one of {sorted(_abi.STORE)}.")
        self.storage = storage
        if trt_variant not in ("reference", "lambda"):
            raise ValueError("trt_variant must be 'reference' or 'lambda'.")
        self.trt_variant = trt_variant
        if bouzidi_variant not in ("reference", "linear"):
            raise ValueError("bouzidi_variant must be 'reference' or 'linear'.")
        self.bouzidi_variant = bouzidi_variant
        self._raise_on_step = None
        if self.les and self.les_model == "wale" and self.van_driest and phi is not None:
            # the reference's WALE loop cannot take a damped (per-cell) constant
            # (les.py:515 -> wale_nu_sgs_2d) and fails on the first step
            self._raise_on_step = ValueError("setting an array element with a sequence.")

        _phi_q = phi if (bouzidi and phi is not None) else None
        self.surface_links, self.q_vals = build_surface_links(self.solid, _phi_q)
        if bouzidi and bouzidi_variant == "linear" and phi is not None:
            self.q_vals = link_fractions(self.surface_links, phi)

        ys, xs = np.where(self.solid)
        self._ref_center_y = float(np.mean(ys)) if ys.size else 0.5 * (Ny - 1)
        self._ref_center_x = float(np.mean(xs)) if xs.size else 0.5 * (Nx - 1)

        d = _abi.AeroDesc()
        d.q, d.nx, d.ny, d.nz = 9, self.Nx, self.Ny, 1
        d.collision = {"bgk": _abi.BGK, "trt": _abi.TRT, "mrt": _abi.MRT}[collision]
        d.storage = _abi.STORE[storage]
        if inlet_bc == "velocity":
            d.inlet = _abi.IN_FIELD if self.synthetic_inflow else _abi.IN_ZOUHE
        elif inlet_bc == "pressure":
            d.inlet = _abi.IN_PRESSURE
        else:
            d.inlet = _abi.IN_NONE
        d.outlet = {"convective": _abi.OUT_CONVECTIVE, "pressure": _abi.OUT_PRESSURE,
                    "zerogradient": _abi.OUT_ZEROGRAD}.get(outlet_bc, _abi.OUT_NONE)
        d.wall = {"slip": _abi.WALL_SLIP, "noslip": _abi.WALL_NOSLIP,
                  "moving": _abi.WALL_MOVING}.get(wall_bc, _abi.WALL_NONE)
        d.bouzidi = 0 if not self.bouzidi else (2 if self.bouzidi_variant == "linear" else 1)
        d.trt_variant = 1 if self.trt_variant == "lambda" else 0
        d.sponge_thickness = self.sponge_thickness
        d.les = self._les_code
        d.thermal, d.buoyancy, d.body_force_on = int(self.thermal), int(self.buoyancy), 0
        d.omega, d.u0, d.rho0, d.D, d.trt_lambda = self.omega, self.u0, self.rho0, self.D, self.trt_lambda
        d.rho_in, d.rho_out, d.inlet_perturbation = self.rho_in, self.rho_out, self.inlet_perturbation
        d.sponge_strength, d.les_cs = self.sponge_strength, self.les_cs
        d.u_wall_top, d.u_wall_bottom = self.wall_velocity_top, self.wall_velocity_bottom
        d.alpha_T, d.T_hot, d.T_cold, d.T_ref = self.alpha_T, self.T_hot, self.T_cold, self.T_ref
        d.g_gravity, d.beta = self.g_gravity, self.beta
        d.center[0], d.center[1], d.center[2] = self._ref_center_x, self._ref_center_y, 0.0
        s = build_s_vec(self.omega)
        for k in range(9):
            d.mrt_s[k] = s[k]
        for k, v in enumerate(M9_inv.ravel()):
            d.mrt_minv[k] = v
        damp = None
        if self.les and self.van_driest and phi is not None:
            damp = (1.0 - np.exp(-np.abs(phi) / self.van_driest_A)) ** 2
        l4 = np.zeros((len(self.surface_links), 4), np.int32)
        l4[:, 0], l4[:, 2], l4[:, 3] = self.surface_links[:, 0], self.surface_links[:, 1], self.surface_links[:, 2]
        self._init_device(d, self.solid[None], l4, self.q_vals, damp)
        if self.ibm_enabled:
            # Guo IBM: links stop bouncing back; forcing only where phi is given
            # (the reference's `ibm_enabled and phi is not None`)
            self._set_ibm(self.phi)

        # f = feq(rho0, u0, 0) (solver.py:218-223), evaluated on the device;
        # the outlet memory is its last column
        self.set_equilibrium()
        self._f_outlet_prev = compute_feq(np.full(Ny, float(rho0)), np.full(Ny, float(u0)), np.zeros(Ny))
        if self.thermal:
            self.g = np.broadcast_to(
                (np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4) * self.T_ref)[:, None, None],
                (9, Ny, Nx)).copy()
        self.step_count = 0
        self.rho = self.ux = self.uy = None
        self.Cd_history, self.Cl_history, self.Cm_history = [], [], []
        self.Cd_p_history, self.Cd_v_history = [], []

    def _ref_shape(self):
        return (9, self.Ny, self.Nx)

    def _planes(self, n):
        if not (self.synthetic_inflow and self.inlet_bc == "velocity"):
            return None
        planes = np.zeros((n, 3, self.Ny))
        planes[:, 0] = self.u0
        planes[:, 1] = self._inflow.next_n(n)
        return planes

    # ---------------------------------------------------------------- step
    def _consume(self, rows: np.ndarray):
        last = rows[-1]
        self._last_cd_p = float(last[6])
        self._last_cd_v = float(last[7])
        self._last_cm = float(last[5])

    def _record(self, rows: np.ndarray):
        """Append the coefficient rows of a run of steps to the histories."""
        self.Cd_history.extend(rows[:, 0].tolist())
        self.Cl_history.extend(rows[:, 1].tolist())
        self.Cm_history.extend(rows[:, 5].tolist())
        self.Cd_p_history.extend(rows[:, 6].tolist())
        self.Cd_v_history.extend(rows[:, 7].tolist())
        self._consume(rows)

    @property
    def _last_force_profile(self):
        cells, forces = self._link_forces()
        fx = np.zeros(self.Ny)
        fy = np.zeros(self.Ny)
        if len(cells):
            y = (cells // self.Nx) % self.Ny
            np.add.at(fx, y, forces[:, 0])
            np.add.at(fy, y, forces[:, 1])
        return {"y": list(range(self.Ny)), "fx": fx.tolist(), "fy": fy.tolist()}

    def _step(self) -> tuple:
        """One reference step (solver.py:276-464); returns (Cd, Cl)."""
        rows = self._advance(1)
        self._consume(rows)
        return float(rows[0, 0]), float(rows[0, 1])

    step = _step

    # ---------------------------------------------------------------- run
    def run(self, steps: int, check_every: int = 500, verbose: bool = True, callback: Optional[Callable] = None,
            checkpoint_every: Optional[int] = None, checkpoint_dir: Optional[str] = None,
            auto_stop: bool = False, convergence_window: int = 5000, convergence_tol: float = 0.01,
            strouhal_window: int = 4096, strouhal_tol: float = 0.05, hdf5_path: Optional[str] = None,
            hdf5_every: Optional[int] = None) -> dict:
        """Reference run loop (solver.py:470-632) in device-resident chunks."""
        if verbose:
            print(f"  Backend   : {self.backend}")
            print(f"  Collision : {self.collision}")
        warn_hdf5(hdf5_path)
        avg_window = max(1, steps // 5)

        on_chunk = self._record

        def on_check(step, total, elapsed):
            self._device_check(step)
            w = min(check_every, len(self.Cd_history))
            if verbose:
                print(f"  step {step:6d}/{total}  Cd={float(np.mean(self.Cd_history[-w:])):+.4f}  "
                      f"Cl={float(np.mean(self.Cl_history[-w:])):+.4f}  [{step / elapsed:.0f} steps/s]",
                      flush=True)
            if callback is not None:
                rho, ux, uy, _ = self._macroscopic()
                callback(step, self.Cd_history[-1], self.Cl_history[-1], rho[0], ux[0], uy[0])

        def stop_fn():
            return detect_statistical_stationarity(
                self.Cd_history, self.Cl_history, self.D, self.u0, convergence_window=convergence_window,
                convergence_tol=convergence_tol, strouhal_window=strouhal_window, strouhal_tol=strouhal_tol)

        stop_reason, last_conv, last_st = run_loop(
            self, steps, check_every, verbose, checkpoint_every, checkpoint_dir,
            "checkpoint_{step:08d}.npz", on_chunk, on_check, stop_fn if auto_stop else None)

        rho, ux, uy, _ = self._macroscopic()
        self.rho, self.ux, self.uy = rho[0], ux[0], uy[0]
        a = lambda h: np.array(h[-avg_window:]) if h else np.zeros(1)  # noqa: E731
        scalar = scalar_stats = None
        if self.thermal:
            scalar = self.g.sum(axis=0)
            scalar_stats = {"enabled": True, "mean": float(np.mean(scalar)), "std": float(np.std(scalar)),
                            "min": float(np.min(scalar)), "max": float(np.max(scalar)),
                            "wall_hot": self.T_hot, "wall_cold": self.T_cold, "diffusivity": self.alpha_T,
                            "buoyancy": self.buoyancy, "gravity": self.g_gravity, "beta": self.beta,
                            "reference": self.T_ref}
        Cd, Cl, Cm = a(self.Cd_history), a(self.Cl_history), a(self.Cm_history)
        return {
            "Cd_mean": float(np.mean(Cd)), "Cl_mean": float(np.mean(Cl)), "Cm_mean": float(np.mean(Cm)),
            "Cd_std": float(np.std(Cd)), "Cl_std": float(np.std(Cl)), "Cm_std": float(np.std(Cm)),
            "Cd_p_mean": float(np.mean(a(self.Cd_p_history))), "Cd_v_mean": float(np.mean(a(self.Cd_v_history))),
            "Cd_history": self.Cd_history, "Cl_history": self.Cl_history, "Cm_history": self.Cm_history,
            "Cd_p_history": self.Cd_p_history, "Cd_v_history": self.Cd_v_history,
            "observables": {"cd_spectrum": coefficient_spectrum(self.Cd_history),
                            "cl_spectrum": coefficient_spectrum(self.Cl_history),
                            "cm_spectrum": coefficient_spectrum(self.Cm_history),
                            "force_profile_last": self._last_force_profile if self.step_count else None},
            "steps_completed": self.step_count, "stop_reason": stop_reason,
            "convergence_report": last_conv, "strouhal_report": last_st,
            "rho": self.rho, "ux": self.ux, "uy": self.uy, "scalar": scalar, "scalar_stats": scalar_stats,
        }

    # ---------------------------------------------------------- checkpoints
    def _params(self) -> dict:
        return dict(
            Ny=self.Ny, Nx=self.Nx, omega=self.omega, u0=self.u0, D=self.D, rho0=self.rho0,
            wall_bc=self.wall_bc, inlet_bc=self.inlet_bc, outlet_bc=self.outlet_bc, rho_in=self.rho_in,
            rho_out=self.rho_out, inlet_perturbation=self.inlet_perturbation, trt_lambda=self.trt_lambda,
            sponge_thickness=self.sponge_thickness, sponge_strength=self.sponge_strength, les=self.les,
            les_cs=self.les_cs, ibm_enabled=self.ibm_enabled, les_model=self.les_model, bouzidi=self.bouzidi,
            van_driest=self.van_driest, van_driest_A=self.van_driest_A,
            wall_velocity_top=self.wall_velocity_top, wall_velocity_bottom=self.wall_velocity_bottom,
            synthetic_inflow=self.synthetic_inflow, synthetic_inflow_intensity=self.synthetic_inflow_intensity,
            thermal=self.thermal, T_hot=self.T_hot, T_cold=self.T_cold, alpha_T=self.alpha_T,
            buoyancy=self.buoyancy, g_gravity=self.g_gravity, beta=self.beta, T_ref=self.T_ref,
            collision=self.collision, backend="auto")

    def _extra_params(self) -> dict:
        return dict(storage=self.storage, trt_variant=self.trt_variant, bouzidi_variant=self.bouzidi_variant,
                    les_form=self.les_form)

    def save_checkpoint(self, path: str) -> None:
        """Same keys and layout as the reference (solver.py:638-686)."""
        kw = dict(f=self.f, f_outlet_prev=self._f_outlet_prev, solid=self.solid,
                  Cd_history=np.array(self.Cd_history, dtype=np.float64),
                  Cl_history=np.array(self.Cl_history, dtype=np.float64),
                  Cm_history=np.array(self.Cm_history, dtype=np.float64),
                  step_count=np.array(self.step_count, dtype=np.int64), surface_links=self.surface_links,
                  q_vals=self.q_vals, params_json=np.array(json.dumps(self._params())),
                  **extra_params_entry(self._extra_params()))
        if self.phi is not None:
            kw["phi"] = self.phi
        if self.thermal:
            kw["g"] = self.g
        if self.synthetic_inflow:
            # not a reference key (its loader ignores it): the inflow generator's
            # state, so a resumed run continues the same inflow
            kw["inflow_state_json"] = np.array(json.dumps(self._inflow.state()))
        np.savez_compressed(path, **kw)

    def load_checkpoint(self, path: str) -> None:
        """solver.py:688-709"""
        data = np.load(path, allow_pickle=False)
        if "surface_links" in data or "q_vals" in data:
            links = data["surface_links"] if "surface_links" in data else self.surface_links
            q = data["q_vals"] if "q_vals" in data else self.q_vals
            self.surface_links, self.q_vals = links, q
            l4 = np.zeros((len(links), 4), np.int32)
            l4[:, 0], l4[:, 2], l4[:, 3] = links[:, 0], links[:, 1], links[:, 2]
            sol = np.ascontiguousarray(self.solid[None], dtype=np.uint8)
            import ctypes
            self._ctx.call("aero_set_geometry", _abi.ptr(sol), ctypes.c_int64(len(l4)), _abi.ptr(l4),
                           _abi.ptr(np.ascontiguousarray(q, dtype=np.float32)))
            if self.ibm_enabled:
                # a new geometry ends immersed-boundary mode: mark the band again
                self._set_ibm(self.phi)
        self.f = data["f"]
        self._f_outlet_prev = data["f_outlet_prev"]
        self.Cd_history = list(data["Cd_history"])
        self.Cl_history = list(data["Cl_history"])
        if "Cm_history" in data:
            self.Cm_history = list(data["Cm_history"])
        self.step_count = int(data["step_count"])
        if "g" in data and self.thermal:
            self.g = data["g"]
        if "inflow_state_json" in data and self.synthetic_inflow:
            self._inflow.set_state(json.loads(str(data["inflow_state_json"])))

    @classmethod
    def from_checkpoint(cls, path: str) -> "Solver":
        """solver.py:711-727"""
        data = np.load(path, allow_pickle=False)
        params = checkpoint_params(data)
        solver = cls(solid=data["solid"], phi=data["phi"] if "phi" in data else None, **params)
        solver.load_checkpoint(path)
        return solver
