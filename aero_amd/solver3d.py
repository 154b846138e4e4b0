"""D3Q19 wind-tunnel solver on the MI355X step library.

Drop-in for the reference's ``aero.lbm.solver3d.Solver3D``
(aero/lbm/solver3d.py:70-838): same constructor arguments, step/run API,
public state and checkpoint layout.  x is streamwise, y vertical (walls),
z spanwise (periodic through the streaming).
"""

from __future__ import annotations

import ctypes
import json
from typing import Callable, List, Optional

import numpy as np

from . import _abi
from ._base import (DeviceSolver, checkpoint_params, extra_params_entry, les_code, resolve_backend, run_loop,
                    warn_hdf5)
from .diagnostics import coefficient_spectrum, detect_statistical_stationarity
from .inflow import SEMInlet, SyntheticInflow3D
from .lattice import M19_inv, W3, build_s3_vec, compute_feq_3d, base_nu_from_omega
from .links import build_surface_links_3d, link_fractions


class Solver3D(DeviceSolver):
    _Q = 19
    _is3d = True

    def __init__(self, Nz: int, Ny: int, Nx: int, solid: np.ndarray, omega: float, u0: float, D: float,
                 rho0: float = 1.0, wall_bc: str = "slip", inlet_bc: str = "velocity",
                 outlet_bc: str = "convective", streamwise_bc: str = "open", backend: str = "auto",
                 collision: str = "bgk", inlet_perturbation: float = 0.0, trt_lambda: float = 0.25,
                 sponge_thickness: int = 0, sponge_strength: float = 0.1, les: bool = False,
                 les_cs: float = 0.16, les_model: str = "smagorinsky", ibm_enabled: bool = False,
                 phi: Optional[np.ndarray] = None, bouzidi: bool = False, van_driest: bool = False,
                 van_driest_A: float = 25.0, body_force_x: float = 0.0, body_force_y: float = 0.0,
                 body_force_z: float = 0.0, wall_velocity_top: float = 0.0, wall_velocity_bottom: float = 0.0,
                 synthetic_inflow: bool = False, synthetic_inflow_intensity: float = 0.03,
                 synthetic_inflow_seed: int = 12345, sem_inlet: bool = False, sem_Tu: float = 0.05,
                 sem_L_int: float = 10.0, sem_N: int = 200, thermal: bool = False, T_hot: float = 1.0,
                 T_cold: float = 0.0, alpha_T: float = 1e-3, buoyancy: bool = False, g_gravity: float = 0.0,
                 beta: float = 1e-3, T_ref: float = 0.5, storage: str = "f64", trt_variant: str = "reference",
                 bouzidi_variant: str = "reference", les_form: str = "fd"):
        """Arguments as the reference (solver3d.py:71-135) plus ``storage``,
        ``trt_variant``, ``bouzidi_variant`` and ``les_form`` (see Solver)."""
        self.Nz, self.Ny, self.Nx = int(Nz), int(Ny), int(Nx)
        self.solid = np.ascontiguousarray(solid, dtype=np.bool_)
        if self.solid.shape != (self.Nz, self.Ny, self.Nx):
            raise ValueError(f"solid must have shape ({Nz}, {Ny}, {Nx})")
        self.fluid = ~self.solid
        self.omega, self.u0, self.D, self.rho0 = float(omega), float(u0), float(D), float(rho0)
        self.wall_bc, self.inlet_bc, self.outlet_bc = wall_bc, inlet_bc, outlet_bc
        if streamwise_bc not in ("open", "periodic", "recycling"):
            raise ValueError(f"Unknown streamwise_bc '{streamwise_bc}'. Choose 'open', 'periodic', or 'recycling'.")
        self.streamwise_bc = streamwise_bc
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
        self.body_force_x, self.body_force_y, self.body_force_z = float(body_force_x), float(body_force_y), float(body_force_z)
        self.wall_velocity_top, self.wall_velocity_bottom = float(wall_velocity_top), float(wall_velocity_bottom)
        self.synthetic_inflow = bool(synthetic_inflow)
        self.synthetic_inflow_intensity = float(synthetic_inflow_intensity)
        self._inflow = SyntheticInflow3D(self.Nz, self.Ny, self.u0, self.synthetic_inflow_intensity,
                                         synthetic_inflow_seed)
        self.sem_inlet = bool(sem_inlet)
        self._sem = SEMInlet(self.Ny, self.Nz, self.u0, float(sem_Tu), float(sem_L_int), int(sem_N)) \
            if self.sem_inlet else None
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
            self._raise_on_step = ValueError("setting an array element with a sequence.")

        _phi_q = phi if (bouzidi and phi is not None) else None
        self.surface_links, self.q_vals = build_surface_links_3d(self.solid, _phi_q)
        if bouzidi and bouzidi_variant == "linear" and phi is not None:
            self.q_vals = link_fractions(self.surface_links, phi)
        zs, ys, xs = np.where(self.solid)
        self._ref_center_z = float(np.mean(zs)) if zs.size else 0.5 * (Nz - 1)
        self._ref_center_y = float(np.mean(ys)) if ys.size else 0.5 * (Ny - 1)
        self._ref_center_x = float(np.mean(xs)) if xs.size else 0.5 * (Nx - 1)

        d = _abi.AeroDesc()
        d.q, d.nx, d.ny, d.nz = 19, self.Nx, self.Ny, self.Nz
        d.collision = {"bgk": _abi.BGK, "trt": _abi.TRT, "mrt": _abi.MRT}[collision]
        d.storage = _abi.STORE[storage]
        d.inlet, d.outlet = _abi.IN_NONE, _abi.OUT_NONE
        if streamwise_bc == "open":
            if inlet_bc == "velocity":
                d.inlet = (_abi.IN_SEM if self._sem is not None else
                           _abi.IN_FIELD if self.synthetic_inflow else _abi.IN_ZOUHE)
            d.outlet = _abi.OUT_CONVECTIVE if outlet_bc == "convective" else _abi.OUT_ZEROGRAD
        elif streamwise_bc == "recycling":
            d.inlet = _abi.IN_RECYCLING
        d.wall = {"slip": _abi.WALL_SLIP, "noslip": _abi.WALL_NOSLIP,
                  "moving": _abi.WALL_MOVING}.get(wall_bc, _abi.WALL_NONE)
        d.bouzidi = 0 if not self.bouzidi else (2 if self.bouzidi_variant == "linear" else 1)
        d.trt_variant = 1 if self.trt_variant == "lambda" else 0
        d.sponge_thickness = self.sponge_thickness
        d.les = self._les_code
        d.thermal, d.buoyancy = int(self.thermal), int(self.buoyancy)
        d.body_force_on = int(any(abs(v) > 0.0 for v in (self.body_force_x, self.body_force_y, self.body_force_z)))
        d.body_force[0], d.body_force[1], d.body_force[2] = self.body_force_x, self.body_force_y, self.body_force_z
        d.omega, d.u0, d.rho0, d.D, d.trt_lambda = self.omega, self.u0, self.rho0, self.D, self.trt_lambda
        d.rho_in, d.rho_out, d.inlet_perturbation = 1.0, 1.0, self.inlet_perturbation
        d.sponge_strength, d.les_cs = self.sponge_strength, self.les_cs
        d.u_wall_top, d.u_wall_bottom = self.wall_velocity_top, self.wall_velocity_bottom
        d.alpha_T, d.T_hot, d.T_cold, d.T_ref = self.alpha_T, self.T_hot, self.T_cold, self.T_ref
        d.g_gravity, d.beta = self.g_gravity, self.beta
        d.center[0], d.center[1], d.center[2] = self._ref_center_x, self._ref_center_y, self._ref_center_z
        s = build_s3_vec(self.omega)
        for k in range(19):
            d.mrt_s[k] = s[k]
        for k, v in enumerate(M19_inv.ravel()):
            d.mrt_minv[k] = v
        damp = None
        if self.les and self.van_driest and phi is not None:
            damp = (1.0 - np.exp(-np.abs(phi) / self.van_driest_A)) ** 2
        self._init_device(d, self.solid, self.surface_links, self.q_vals, damp)
        if self.ibm_enabled:
            # Guo IBM: links stop bouncing back; forcing only where phi is given
            # (the reference's `ibm_enabled and phi is not None`)
            self._set_ibm(self.phi)
        self._sem_device = False
        if self._sem is not None and d.inlet == _abi.IN_SEM:
            # the SEM planes are formed on the device from the eddy states the
            # host advances (SEMInlet.step keeps the reference's random draws)
            self._ctx.call("aero_set_sem", ctypes.c_int32(self._sem.N), ctypes.c_double(self._sem.sigma),
                           ctypes.c_double(self._sem.Tu))
            self._sem_device = True

        sh = (Nz, Ny, Nx)
        # f = feq(rho0, u0, 0) (solver3d.py:239-247), evaluated on the device;
        # the outlet memory is its last x plane
        self.set_equilibrium()
        ps = (Nz, Ny)
        self._f_outlet_prev = compute_feq_3d(np.full(ps, float(rho0)), np.full(ps, float(u0)), np.zeros(ps),
                                             np.zeros(ps))
        if self.thermal:
            self.g = np.broadcast_to((W3 * self.T_ref)[:, None, None, None], (19,) + sh).copy()
        self.step_count = 0
        self.Cd_history: List[float] = []
        self.Cly_history: List[float] = []
        self.Clz_history: List[float] = []
        self.Cmx_history: List[float] = []
        self.Cmy_history: List[float] = []
        self.Cmz_history: List[float] = []
        self.Cd_p_history: List[float] = []
        self.Cd_v_history: List[float] = []

    def _ref_shape(self):
        return (19, self.Nz, self.Ny, self.Nx)

    def _planes(self, n):
        if self.streamwise_bc != "open" or self.inlet_bc != "velocity":
            return None
        if self._sem is not None and self._sem_device:
            # per step: the eddy state after that step's advance (aero_set_sem)
            snaps = np.empty((n, 6, self._sem.N))
            for k in range(n):
                self._sem.step()
                snaps[k, 0], snaps[k, 1], snaps[k, 2] = self._sem.xk, self._sem.yk, self._sem.zk
                snaps[k, 3:] = self._sem.eps
            return snaps
        if self._sem is not None:
            planes = np.empty((n, 3, self.Nz, self.Ny))
            for k in range(n):
                self._sem.step()
                planes[k] = np.stack(self._sem.fluctuation())
            return planes
        if self.synthetic_inflow:
            planes = np.empty((n, 3, self.Nz, self.Ny))
            planes[:, 0] = self.u0
            planes[:, 1], planes[:, 2] = self._inflow.next_n(n)
            return planes
        return None

    def _consume(self, rows):
        last = rows[-1]
        self._last_cd_p, self._last_cd_v = float(last[6]), float(last[7])
        self._last_cmx, self._last_cmy, self._last_cmz = float(last[3]), float(last[4]), float(last[5])

    @property
    def _last_spanwise_profile(self):
        cells, forces = self._link_forces()
        prof = {k: np.zeros(self.Nz) for k in ("fx", "fy", "fz")}
        if len(cells):
            z = cells // (self.Nx * self.Ny)
            for k, col in (("fx", 0), ("fy", 1), ("fz", 2)):
                np.add.at(prof[k], z, forces[:, col])
        return {"z": list(range(self.Nz)), **{k: v.tolist() for k, v in prof.items()}}

    _HISTS = ("Cd_history", "Cly_history", "Clz_history", "Cmx_history", "Cmy_history", "Cmz_history",
              "Cd_p_history", "Cd_v_history")

    def _record(self, rows: np.ndarray):
        """Append the coefficient rows of a run of steps to the histories."""
        for c, name in enumerate(self._HISTS):
            getattr(self, name).extend(rows[:, c].tolist())
        self._consume(rows)

    def _step(self) -> tuple:
        """One reference step (solver3d.py:325-541); returns (Cd, Cl_y, Cl_z)."""
        rows = self._advance(1)
        self._consume(rows)
        return float(rows[0, 0]), float(rows[0, 1]), float(rows[0, 2])

    step = _step

    def run(self, steps: int, check_every: int = 500, verbose: bool = True, callback: Optional[Callable] = None,
            checkpoint_every: Optional[int] = None, checkpoint_dir: Optional[str] = None,
            auto_stop: bool = False, convergence_window: int = 5000, convergence_tol: float = 0.01,
            strouhal_window: int = 4096, strouhal_tol: float = 0.05, hdf5_path: Optional[str] = None,
            hdf5_every: Optional[int] = None) -> dict:
        """Reference 3D run loop (solver3d.py:545-723) in device-resident chunks."""
        if verbose:
            print(f"  Backend   : {self.backend}  (3D D3Q19)")
            print(f"  Collision : {self.collision}")
        warn_hdf5(hdf5_path)
        avg_window = max(1, steps // 5)
        on_chunk = self._record

        def on_check(step, total, elapsed):
            self._device_check(step)  # z-midplane, as solver3d.py:608-615
            w = min(check_every, len(self.Cd_history))
            if verbose:
                print(f"  step {step:6d}/{total}  Cd={float(np.mean(self.Cd_history[-w:])):+.4f}  "
                      f"Cl_y={float(np.mean(self.Cly_history[-w:])):+.4f}  [{step / elapsed:.0f} steps/s]",
                      flush=True)
            if callback is not None:
                callback(step, self.Cd_history[-1], self.Cly_history[-1], self.Clz_history[-1])

        def stop_fn():
            return detect_statistical_stationarity(
                self.Cd_history, self.lift_history(strouhal_window), self.D, self.u0,
                convergence_window=convergence_window,
                convergence_tol=convergence_tol, strouhal_window=strouhal_window, strouhal_tol=strouhal_tol)

        stop_reason, last_conv, last_st = run_loop(
            self, steps, check_every, verbose, checkpoint_every, checkpoint_dir,
            "checkpoint3d_{step:08d}.npz", on_chunk, on_check, stop_fn if auto_stop else None)

        rho, ux, uy, uz = self._macroscopic()
        a = lambda h: np.array(h[-avg_window:]) if h else np.zeros(1)  # noqa: E731
        scalar = scalar_stats = None
        if self.thermal:
            scalar = self.g.sum(axis=0)
            scalar_stats = {"enabled": True, "mean": float(np.mean(scalar)), "std": float(np.std(scalar)),
                            "min": float(np.min(scalar)), "max": float(np.max(scalar)),
                            "wall_hot": self.T_hot, "wall_cold": self.T_cold, "diffusivity": self.alpha_T,
                            "buoyancy": self.buoyancy, "gravity": self.g_gravity, "beta": self.beta,
                            "reference": self.T_ref}
        out = {}
        for key, name in (("Cd", "Cd_history"), ("Cly", "Cly_history"), ("Clz", "Clz_history")):
            arr = a(getattr(self, name))
            out[f"{key}_mean"] = float(np.mean(arr))
            out[f"{key}_std"] = float(np.std(arr))
        for key, name in (("Cmx", "Cmx_history"), ("Cmy", "Cmy_history"), ("Cmz", "Cmz_history")):
            arr = a(getattr(self, name))
            out[f"{key}_mean"] = float(np.mean(arr))
            out[f"{key}_std"] = float(np.std(arr))
        out["Cd_p_mean"] = float(np.mean(a(self.Cd_p_history)))
        out["Cd_v_mean"] = float(np.mean(a(self.Cd_v_history)))
        for name in self._HISTS:
            out[name] = getattr(self, name)
        out["observables"] = {
            "cd_spectrum": coefficient_spectrum(self.Cd_history),
            "cly_spectrum": coefficient_spectrum(self.Cly_history),
            "clz_spectrum": coefficient_spectrum(self.Clz_history),
            "cmz_spectrum": coefficient_spectrum(self.Cmz_history),
            "spanwise_force_profile_last": self._last_spanwise_profile if self.step_count else None,
        }
        out.update({"steps_completed": self.step_count, "stop_reason": stop_reason,
                    "convergence_report": last_conv, "strouhal_report": last_st,
                    "rho": rho, "ux": ux, "uy": uy, "uz": uz, "scalar": scalar, "scalar_stats": scalar_stats})
        return out

    # ---------------------------------------------------------- checkpoints
    def _params(self) -> dict:
        return dict(
            Nz=self.Nz, Ny=self.Ny, Nx=self.Nx, omega=self.omega, u0=self.u0, D=self.D, rho0=self.rho0,
            wall_bc=self.wall_bc, inlet_bc=self.inlet_bc, outlet_bc=self.outlet_bc,
            streamwise_bc=self.streamwise_bc, backend="numpy", collision=self.collision,
            inlet_perturbation=self.inlet_perturbation, trt_lambda=self.trt_lambda,
            sponge_thickness=self.sponge_thickness, sponge_strength=self.sponge_strength, les=self.les,
            les_cs=self.les_cs, les_model=self.les_model, ibm_enabled=self.ibm_enabled, bouzidi=self.bouzidi,
            van_driest=self.van_driest, van_driest_A=self.van_driest_A, body_force_x=self.body_force_x,
            body_force_y=self.body_force_y, body_force_z=self.body_force_z,
            wall_velocity_top=self.wall_velocity_top, wall_velocity_bottom=self.wall_velocity_bottom,
            synthetic_inflow=self.synthetic_inflow, synthetic_inflow_intensity=self.synthetic_inflow_intensity,
            thermal=self.thermal, T_hot=self.T_hot, T_cold=self.T_cold, alpha_T=self.alpha_T,
            buoyancy=self.buoyancy, g_gravity=self.g_gravity, beta=self.beta, T_ref=self.T_ref)

    def _extra_params(self) -> dict:
        return dict(storage=self.storage, trt_variant=self.trt_variant, bouzidi_variant=self.bouzidi_variant,
                    les_form=self.les_form)

    def save_checkpoint(self, path: str) -> None:
        """Same keys and layout as the reference (solver3d.py:727-783)."""
        kw = dict(f=self.f, f_outlet_prev=self._f_outlet_prev, solid=self.solid,
                  surface_links=self.surface_links, q_vals=self.q_vals,
                  Cd_history=np.array(self.Cd_history, dtype=np.float64),
                  Cly_history=np.array(self.Cly_history, dtype=np.float64),
                  Clz_history=np.array(self.Clz_history, dtype=np.float64),
                  Cmx_history=np.array(self.Cmx_history, dtype=np.float64),
                  Cmy_history=np.array(self.Cmy_history, dtype=np.float64),
                  Cmz_history=np.array(self.Cmz_history, dtype=np.float64),
                  step_count=np.array(self.step_count, dtype=np.int64),
                  params_json=np.array(json.dumps(self._params())),
                  **extra_params_entry(self._extra_params()))
        if self.phi is not None:
            kw["phi"] = self.phi
        if self.thermal:
            kw["g"] = self.g
        if self._sem is not None:
            # not a reference key (its loader ignores it): the SEM eddies and
            # random state, so a resumed run continues the same inlet turbulence
            kw["sem_state_json"] = np.array(json.dumps(self._sem.state()))
        if self.synthetic_inflow:
            # not a reference key (its loader ignores it): the inflow generator's
            # state, so a resumed run continues the same inflow
            kw["inflow_state_json"] = np.array(json.dumps(self._inflow.state()))
        np.savez_compressed(path, **kw)

    def load_checkpoint(self, path: str) -> None:
        """solver3d.py:785-814"""
        data = np.load(path, allow_pickle=False)
        if "surface_links" in data or "q_vals" in data:
            links = data["surface_links"] if "surface_links" in data else self.surface_links
            q = data["q_vals"] if "q_vals" in data else self.q_vals
            self.surface_links, self.q_vals = links, q
            sol = np.ascontiguousarray(self.solid, dtype=np.uint8)
            l4 = np.ascontiguousarray(links, dtype=np.int32)
            self._ctx.call("aero_set_geometry", _abi.ptr(sol), ctypes.c_int64(len(l4)), _abi.ptr(l4),
                           _abi.ptr(np.ascontiguousarray(q, dtype=np.float32)))
            if self.ibm_enabled:
                # a new geometry ends immersed-boundary mode: mark the band again
                self._set_ibm(self.phi)
        self.f = data["f"]
        self._f_outlet_prev = data["f_outlet_prev"]
        self.Cd_history = list(data["Cd_history"])
        self.Cly_history = list(data["Cly_history"])
        self.Clz_history = list(data["Clz_history"])
        for name in ("Cmx_history", "Cmy_history", "Cmz_history"):
            if name in data:
                setattr(self, name, list(data[name]))
        self.step_count = int(data["step_count"])
        if "g" in data and self.thermal:
            self.g = data["g"]
        if "sem_state_json" in data and self._sem is not None:
            self._sem.set_state(json.loads(str(data["sem_state_json"])))
        if "inflow_state_json" in data and self.synthetic_inflow:
            self._inflow.set_state(json.loads(str(data["inflow_state_json"])))

    @classmethod
    def from_checkpoint(cls, path: str, backend: Optional[str] = None) -> "Solver3D":
        """solver3d.py:816-838"""
        data = np.load(path, allow_pickle=False)
        params = checkpoint_params(data)
        if backend is not None:
            params["backend"] = backend
        if "sem_state_json" in data:
            st = json.loads(str(data["sem_state_json"]))
            params.update(sem_inlet=True, sem_Tu=st["Tu"], sem_L_int=st["L_int"], sem_N=st["N"])
        solver = cls(solid=data["solid"], phi=data["phi"] if "phi" in data else None, **params)
        solver.load_checkpoint(path)
        return solver
