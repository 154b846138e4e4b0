"""Shared machinery of Solver (D2Q9) and Solver3D (D3Q19).

The device context holds every population array; the Python object keeps the
reference's public attributes (histories, links, parameters) and turns the
reference's per-step Python loop into chunked ``aero_step`` calls, syncing
with the host only at check points (aero/lbm/solver.py:470-632).
"""

from __future__ import annotations

import ctypes
import json
import time
import warnings
import weakref
from typing import Optional

import numpy as np

from . import _abi
from .diagnostics import check_stability_stats

VALID_BACKENDS = ("auto", "hip", "numpy", "numba", "cupy")


def resolve_backend(backend: str) -> str:
    """The reference selects numpy/numba/cupy kernels (solver.py:192-203,
    solver3d.py:197-221).  Here every name runs the same HIP kernels; unknown
    names raise ValueError exactly as the reference does."""
    if backend not in VALID_BACKENDS:
        raise ValueError(f"Unknown backend '{backend}'. Choose 'auto', 'numpy', 'numba', 'cupy' or 'hip'.")
    return "hip"


# Constructor arguments this build adds to the reference's (storage format and
# the labelled operator variants).  Checkpoints keep them under a key of their
# own, so ``params_json`` holds only the reference's constructor arguments and
# the reference's from_checkpoint (aero/lbm/solver.py:711-727,
# solver3d.py:816-838) can rebuild a solver from a checkpoint written here.
EXTRA_PARAMS_KEY = "aero_amd_params_json"
EXTRA_PARAM_NAMES = ("storage", "trt_variant", "bouzidi_variant", "les_form")

LES_FORMS = ("fd", "local")


def les_code(les: bool, les_model: str, les_form: str) -> int:
    """The device LES mode.  les_form "fd" is the reference's finite-difference
    Smagorinsky / WALE on the pre-collision velocity (les.py:273-344), run as a
    split step; "local" is Smagorinsky from each cell's non-equilibrium
    stress, fused into the step (and batchable).  WALE needs the rotation part
    of the velocity gradient, which the local stress does not carry, so it
    has no local form."""
    if les_form not in LES_FORMS:
        raise ValueError(f"les_form must be one of {LES_FORMS}.")
    if not les:
        return _abi.LES_OFF
    if les_form == "local":
        if les_model != "smagorinsky":
            raise ValueError("les_form='local' is Smagorinsky only: WALE needs the velocity gradient's "
                             "rotation part, which the local non-equilibrium stress does not carry.")
        return _abi.LES_SMAGORINSKY_LOCAL
    return _abi.LES_WALE if les_model == "wale" else _abi.LES_SMAGORINSKY


def extra_params_entry(extra: dict) -> dict:
    return {EXTRA_PARAMS_KEY: np.array(json.dumps(extra))}


def checkpoint_params(data) -> dict:
    """Constructor arguments of a checkpoint: the reference's params_json plus,
    when present, this build's own entries (also accepted inside params_json,
    as older checkpoints of this build wrote them)."""
    params = json.loads(str(data["params_json"]))
    if EXTRA_PARAMS_KEY in data:
        params.update(json.loads(str(data[EXTRA_PARAMS_KEY])))
    return params


class _WriteThrough:
    """Shared by a state array handed out by DeviceSolver.f / .g and every
    view of it: the solver, the attribute, the whole array and the state
    generation it was read at."""

    __slots__ = ("owner", "name", "root", "gen")

    def __init__(self, owner, name, root, gen):
        self.owner, self.name, self.root, self.gen = owner, name, root, gen


class LiveField(np.ndarray):
    """The reference keeps self.f / self.g as numpy arrays, so scripts edit
    them in place (``s.f[:, y, x] += 0.05``).  Here the state lives on the
    device and f / g return host copies; an element assignment into such a
    copy, or into any view of it, writes the whole array back while the
    solver's state is still the one the copy was read from.  After a step or
    another write the copy is detached, as an array the reference replaced
    would be.  Arrays computed from it (``s.f * 2``) are plain copies."""

    def __array_finalize__(self, obj):
        # only views share the state array's memory (a ufunc result is a view
        # of a fresh buffer: no overlap)
        wt = getattr(obj, "_wt", None)
        self._wt = wt if wt is not None and np.may_share_memory(self, wt.root) else None

    def __array_wrap__(self, obj, context=None, return_scalar=False):
        # results of ufuncs and reductions are plain arrays (or scalars)
        out = obj.view(np.ndarray) if isinstance(obj, np.ndarray) else obj
        return out[()] if return_scalar else out

    def __setitem__(self, key, value):
        super().__setitem__(key, value)
        wt = getattr(self, "_wt", None)
        if wt is None:
            return
        owner = wt.owner()
        if owner is None or owner._state_gen != wt.gen:
            return
        setattr(owner, wt.name, np.asarray(wt.root))
        wt.gen = owner._state_gen


def _live(owner, name, arr):
    out = arr.view(LiveField)
    out._wt = _WriteThrough(weakref.ref(owner), name, out, owner._state_gen)
    return out


class DeviceSolver:
    """Holds the aero_ctx and the reference-layout views of its state."""

    _Q: int
    _is3d: bool

    # ------------------------------------------------------------- setup
    def _init_device(self, desc: _abi.AeroDesc, solid3: np.ndarray, links4: np.ndarray,
                     q_vals: np.ndarray, damp: Optional[np.ndarray]):
        self._ctx = _abi.Context(desc)
        sol = np.ascontiguousarray(solid3, dtype=np.uint8)
        l4 = np.ascontiguousarray(links4, dtype=np.int32)
        qv = np.ascontiguousarray(q_vals, dtype=np.float32)
        self._ctx.call("aero_set_geometry", _abi.ptr(sol), ctypes.c_int64(len(l4)), _abi.ptr(l4), _abi.ptr(qv))
        if damp is not None:
            self._ctx.call("aero_set_damping", _abi.ptr(np.ascontiguousarray(damp, dtype=np.float64)))
        self._nrow = desc.nz * desc.ny
        self._N = desc.nz * desc.ny * desc.nx

    def _set_ibm(self, phi: Optional[np.ndarray]) -> None:
        arr = None
        if phi is not None:
            arr = np.ascontiguousarray(phi, dtype=np.float64).reshape(-1)
            if arr.size != self._N:
                raise ValueError(f"phi must have {self._N} cells, got {arr.size}")
        self._ctx.call("aero_set_ibm", _abi.ptr(arr))

    # ------------------------------------------------------------- state
    def _ref_shape(self):
        raise NotImplementedError

    @property
    def f(self) -> np.ndarray:
        """The reference's self.f: post-boundary, pre-collision populations."""
        out = np.empty(self._ref_shape(), dtype=np.float64)
        self._ctx.call("aero_get_f", _abi.ptr(out))
        return _live(self, "f", out)

    @f.setter
    def f(self, value) -> None:
        arr = np.ascontiguousarray(value, dtype=np.float64)
        if arr.shape != self._ref_shape():
            raise ValueError(f"f must have shape {self._ref_shape()}, got {arr.shape}")
        self._ctx.call("aero_set_f", _abi.ptr(arr))
        self._bump()

    # generation of the device state: every step and every write bumps it,
    # which detaches host copies handed out before (LiveField)
    _state_gen = 0

    def _bump(self) -> None:
        self._state_gen += 1

    def set_equilibrium(self, rho=None, ux=None, uy=None, uz=None) -> None:
        """self.f = compute_feq(rho, ux, uy[, uz]) evaluated on the device
        (bit-identical to the host formula, without a host field of Q planes).
        Each argument is a full-grid array or None for the constant rho0, u0,
        0, 0.  Not part of the reference API (it sets self.f directly)."""
        shape = self._ref_shape()[1:]
        args = []
        for name, v in (("rho", rho), ("ux", ux), ("uy", uy), ("uz", uz)):
            if v is None:
                args.append(None)
                continue
            if name == "uz" and len(shape) == 2:
                raise ValueError("uz is a 3D field")
            a = np.ascontiguousarray(v, dtype=np.float64)
            if a.shape != shape:
                raise ValueError(f"{name} must have shape {shape}, got {a.shape}")
            args.append(a)
        self._ctx.call("aero_set_f_equilibrium", *[_abi.ptr(a) for a in args])
        self._bump()

    @property
    def g(self) -> Optional[np.ndarray]:
        if not self.thermal:
            return None
        out = np.empty(self._ref_shape(), dtype=np.float64)
        self._ctx.call("aero_get_g", _abi.ptr(out))
        return _live(self, "g", out)

    @g.setter
    def g(self, value) -> None:
        if value is None:
            return
        if not self.thermal:
            raise ValueError("g can only be set when thermal=True")
        arr = np.ascontiguousarray(value, dtype=np.float64)
        if arr.shape != self._ref_shape():
            raise ValueError(f"g must have shape {self._ref_shape()}, got {arr.shape}")
        self._ctx.call("aero_set_g", _abi.ptr(arr))
        self._bump()

    def _prev_shape(self):
        s = self._ref_shape()
        return s[:-1]

    @property
    def _f_outlet_prev(self) -> np.ndarray:
        out = np.empty(self._prev_shape(), dtype=np.float64)
        self._ctx.call("aero_get_outlet_prev", _abi.ptr(out))
        return out

    @_f_outlet_prev.setter
    def _f_outlet_prev(self, value) -> None:
        arr = np.ascontiguousarray(value, dtype=np.float64)
        if arr.shape != self._prev_shape():
            raise ValueError(f"f_outlet_prev must have shape {self._prev_shape()}, got {arr.shape}")
        self._ctx.call("aero_set_outlet_prev", _abi.ptr(arr))
        self._bump()

    def _macroscopic(self):
        n = self._N
        rho, ux, uy, uz = (np.empty(n) for _ in range(4))
        self._ctx.call("aero_macroscopic", _abi.ptr(rho), _abi.ptr(ux), _abi.ptr(uy), _abi.ptr(uz))
        sh = self._ref_shape()[1:]
        return rho.reshape(sh), ux.reshape(sh), uy.reshape(sh), uz.reshape(sh)

    def _device_stats(self) -> np.ndarray:
        """check_stability's statistics of the current state, reduced on the
        device (3D: the z mid-plane, as solver3d.py:608-615)."""
        stats = np.empty(_abi.NCHECK)
        self._ctx.call("aero_check", _abi.ptr(stats))
        return stats

    def _device_check(self, step: int) -> None:
        check_stability_stats(self._device_stats(), step)

    def lift_history(self, window: int = 4096) -> list:
        """The lift history the auto-stop and Strouhal tests read: Cl in 2D;
        in 3D Cly, or Clz when its standard deviation over the last `window`
        steps is larger (aero/lbm/solver3d.py:634-638)."""
        if not self._is3d:
            return self.Cl_history
        w = window
        if np.std(self.Clz_history[-min(len(self.Clz_history), w):]) > np.std(
                self.Cly_history[-min(len(self.Cly_history), w):]):
            return self.Clz_history
        return self.Cly_history

    def _link_forces(self):
        n = ctypes.c_int32(0)
        self._ctx.call("aero_link_forces", None, None, ctypes.byref(n))
        cells = np.empty(n.value, dtype=np.int32)
        forces = np.empty((n.value, 12), dtype=np.float64)
        self._ctx.call("aero_link_forces", _abi.ptr(cells), _abi.ptr(forces), ctypes.byref(n))
        return cells, forces

    def device_bytes(self) -> int:
        b = ctypes.c_int64(0)
        self._ctx.call("aero_device_bytes", ctypes.byref(b))
        return int(b.value)

    # ------------------------------------------------------------- stepping
    def _planes(self, n: int) -> Optional[np.ndarray]:
        """Per-step inlet planes (n, 3, nrow) for field/SEM inlets, or None."""
        return None

    def _advance(self, n: int) -> np.ndarray:
        """Run n steps on the device; returns the (n, 20) coefficient rows."""
        if getattr(self, "_raise_on_step", None):
            raise self._raise_on_step
        planes = self._planes(n)
        hist = np.empty((n, _abi.NCOEF), dtype=np.float64)
        self._ctx.call("aero_step", ctypes.c_int32(n), ctypes.c_int64(self.step_count),
                       _abi.ptr(planes), _abi.ptr(hist))
        self.step_count += n
        self._bump()
        return hist

    def close(self) -> None:
        ctx = getattr(self, "_ctx", None)
        if ctx is not None:
            ctx.close()


def run_loop(solver, steps, check_every, verbose, checkpoint_every, checkpoint_dir, ckpt_name,
             on_chunk, on_check, auto_stop_fn):
    """Chunked version of the reference run loop (solver.py:517-570).

    Chunks end at every multiple of check_every and checkpoint_every so the
    host only synchronises where the reference would look at the state.
    Returns (stop_reason, last_convergence, last_strouhal).
    """
    import pathlib
    t0 = time.perf_counter()
    step = 0
    stop_reason = "max_steps"
    last_conv = last_st = None
    while step < steps:
        nxt = steps
        for every in (check_every, checkpoint_every if (checkpoint_every and checkpoint_dir) else None):
            if every:
                nxt = min(nxt, (step // every + 1) * every)
        n = nxt - step
        rows = solver._advance(n)
        on_chunk(rows)
        step = nxt
        if checkpoint_every and checkpoint_dir and step % checkpoint_every == 0:
            path = pathlib.Path(checkpoint_dir) / ckpt_name.format(step=solver.step_count)
            solver.save_checkpoint(str(path))
        if step % check_every == 0:
            elapsed = time.perf_counter() - t0
            on_check(step, steps, elapsed)
            if auto_stop_fn is not None:
                last_conv, last_st, stop = auto_stop_fn()
                if stop:
                    stop_reason = "auto_converged"
                    break
    return stop_reason, last_conv, last_st


def warn_hdf5(hdf5_path):
    if hdf5_path:
        warnings.warn("HDF5 snapshots are not written by this build (the reference writer is a no-op "
                      "without h5py as well)", RuntimeWarning, stacklevel=3)
