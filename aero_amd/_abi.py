"""ctypes binding of libaero_hip.so (declared in include/aero_hip.h).

The library is loaded from this package directory; it is built in-tree by
``python -m aero_amd.build``.  There is no fallback: if the shared library is
missing or a device call fails, the error surfaces immediately.
"""

from __future__ import annotations

import ctypes
import pathlib

import numpy as np

LIB_PATH = pathlib.Path(__file__).resolve().parent / "libaero_hip.so"

# enums mirrored from include/aero_hip.h
BGK, TRT, MRT = 0, 1, 2
STORE = {"f64": 0, "f32": 1, "f16": 2, "bf16": 3}
IN_NONE, IN_ZOUHE, IN_PRESSURE, IN_FIELD, IN_SEM, IN_RECYCLING = range(6)
OUT_NONE, OUT_CONVECTIVE, OUT_ZEROGRAD, OUT_PRESSURE = range(4)
WALL_NONE, WALL_SLIP, WALL_NOSLIP, WALL_MOVING = range(4)
LES_OFF, LES_SMAGORINSKY, LES_WALE, LES_SMAGORINSKY_LOCAL = range(4)
NCOEF = 20
NCHECK = 5
NSTAT = 15
STAT_MAX_WINDOW = 4096


class AeroDesc(ctypes.Structure):
    _fields_ = [
        ("q", ctypes.c_int32), ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("collision", ctypes.c_int32), ("storage", ctypes.c_int32), ("inlet", ctypes.c_int32),
        ("outlet", ctypes.c_int32), ("wall", ctypes.c_int32), ("bouzidi", ctypes.c_int32),
        ("sponge_thickness", ctypes.c_int32), ("les", ctypes.c_int32), ("thermal", ctypes.c_int32),
        ("buoyancy", ctypes.c_int32), ("body_force_on", ctypes.c_int32), ("trt_variant", ctypes.c_int32),
        ("omega", ctypes.c_double), ("u0", ctypes.c_double), ("rho0", ctypes.c_double),
        ("D", ctypes.c_double), ("trt_lambda", ctypes.c_double),
        ("rho_in", ctypes.c_double), ("rho_out", ctypes.c_double),
        ("inlet_perturbation", ctypes.c_double), ("sponge_strength", ctypes.c_double),
        ("les_cs", ctypes.c_double), ("u_wall_top", ctypes.c_double), ("u_wall_bottom", ctypes.c_double),
        ("body_force", ctypes.c_double * 3),
        ("alpha_T", ctypes.c_double), ("T_hot", ctypes.c_double), ("T_cold", ctypes.c_double),
        ("T_ref", ctypes.c_double), ("g_gravity", ctypes.c_double), ("beta", ctypes.c_double),
        ("center", ctypes.c_double * 3),
        ("mrt_s", ctypes.c_double * 19),
        ("mrt_minv", ctypes.c_double * 361),
    ]


_P = ctypes.c_void_p
_SIGS = {
    "aero_last_error": ([], ctypes.c_char_p),
    "aero_version": ([], ctypes.c_int),
    "aero_device_count": ([_P], ctypes.c_int),
    "aero_create": ([ctypes.POINTER(AeroDesc), ctypes.POINTER(_P)], ctypes.c_int),
    "aero_destroy": ([_P], ctypes.c_int),
    "aero_set_geometry": ([_P, _P, ctypes.c_int64, _P, _P], ctypes.c_int),
    "aero_set_damping": ([_P, _P], ctypes.c_int),
    "aero_set_ibm": ([_P, _P], ctypes.c_int),
    "aero_set_sem": ([_P, ctypes.c_int32, ctypes.c_double, ctypes.c_double], ctypes.c_int),
    "aero_sem_planes": ([_P, ctypes.c_int32, _P, _P], ctypes.c_int),
    "aero_set_f": ([_P, _P], ctypes.c_int),
    "aero_get_f": ([_P, _P], ctypes.c_int),
    "aero_set_f_equilibrium": ([_P, _P, _P, _P, _P], ctypes.c_int),
    "aero_set_g": ([_P, _P], ctypes.c_int),
    "aero_get_g": ([_P, _P], ctypes.c_int),
    "aero_set_outlet_prev": ([_P, _P], ctypes.c_int),
    "aero_get_outlet_prev": ([_P, _P], ctypes.c_int),
    "aero_step": ([_P, ctypes.c_int32, ctypes.c_int64, _P, _P], ctypes.c_int),
    "aero_step_noalloc": ([_P, ctypes.c_int32, ctypes.c_int64, _P], ctypes.c_int),
    "aero_tune_set": ([ctypes.c_int32] * 9, ctypes.c_int),
    "aero_tune_clear": ([], ctypes.c_int),
    "aero_tune_grid_class": ([ctypes.c_int64, _P], ctypes.c_int),
    "aero_step_tiles": ([_P, _P], ctypes.c_int),
    "aero_halo_bytes": ([_P, _P], ctypes.c_int),
    "aero_halo_pack": ([_P, ctypes.c_int32, ctypes.c_int32, _P], ctypes.c_int),
    "aero_halo_unpack": ([_P, ctypes.c_int32, ctypes.c_int32, _P], ctypes.c_int),
    "aero_halo_exchange": ([_P, _P], ctypes.c_int),
    "aero_set_history_ring": ([_P, ctypes.c_int64], ctypes.c_int),
    "aero_stationarity": ([ctypes.c_int32, _P, _P, ctypes.c_int32, ctypes.c_double, ctypes.c_int32, ctypes.c_double,
                           _P], ctypes.c_int),
    "aero_stationarity_arrays": ([ctypes.c_int32, _P, _P, _P, _P, _P, _P, ctypes.c_int32, ctypes.c_double,
                                  ctypes.c_int32, ctypes.c_double, _P], ctypes.c_int),
    "aero_macroscopic": ([_P, _P, _P, _P, _P], ctypes.c_int),
    "aero_check": ([_P, _P], ctypes.c_int),
    "aero_link_forces": ([_P, _P, _P, _P], ctypes.c_int),
    "aero_synchronize": ([_P], ctypes.c_int),
    "aero_device_bytes": ([_P, _P], ctypes.c_int),
    "aero_set_timing": ([_P, ctypes.c_int32], ctypes.c_int),
    "aero_get_timing": ([_P, _P, _P], ctypes.c_int),
    "aero_cell_counts": ([_P, _P, _P], ctypes.c_int),
    "aero_mem_info": ([_P, _P], ctypes.c_int),
    "aero_mesh_inside": ([_P, ctypes.c_int64, _P, _P, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32,
                          ctypes.c_int32, _P], ctypes.c_int),
    "aero_mesh_sqdist": ([_P, ctypes.c_int64, _P, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                          _P], ctypes.c_int),
    "aero_batch_create": ([ctypes.POINTER(_P)], ctypes.c_int),
    "aero_batch_destroy": ([_P], ctypes.c_int),
    "aero_batch_step": ([_P, ctypes.c_int32, _P, ctypes.c_int32, _P, _P, _P], ctypes.c_int),
    "aero_batch_check": ([_P, ctypes.c_int32, _P, _P], ctypes.c_int),
}
EXPORTED = tuple(_SIGS)

_lib = None


def lib() -> ctypes.CDLL:
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m aero_amd.build` "
                "(there is no CPU fallback)")
        L = ctypes.CDLL(str(LIB_PATH))
        for name, (args, res) in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        _lib = L
        # the launch tuning measured on this device class, when a table ships
        # with the package (aero_amd/autotune.py); it changes schedules only
        from . import autotune
        autotune.load_default()
    return _lib


def check(rc: int) -> None:
    if rc != 0:
        msg = lib().aero_last_error().decode(errors="replace")
        raise RuntimeError(f"aero_hip: {msg}")


def ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def device_count() -> int:
    n = ctypes.c_int32(0)
    check(lib().aero_device_count(ctypes.byref(n)))
    return int(n.value)


class Context:
    """Owns one aero_ctx (device state of one solver)."""

    def __init__(self, desc: AeroDesc):
        self._h = _P()
        check(lib().aero_create(ctypes.byref(desc), ctypes.byref(self._h)))
        self.desc = desc

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            lib().aero_destroy(self._h)
            self._h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def call(self, name, *args):
        check(getattr(lib(), name)(self._h, *args))
