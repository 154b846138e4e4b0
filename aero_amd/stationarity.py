"""Convergence statistics on the device (FP64).

The reference's auto-stop test -- drag window convergence plus Strouhal
drift between the halves of the lift window (aero/diagnostics.py:87-120,
153-228), on the lift series it picks in 3D (aero/lbm/solver3d.py:634-638)
-- evaluated on the GPU for many cases in one call, over FP64 rings of each
case's drag and lift histories that the step calls append on the device
(include/aero_hip.h: aero_set_history_ring, aero_stationarity).  The host
restatement in ``diagnostics.detect_statistical_stationarity`` remains the
definition; the device test matches its decisions (tests/test_gpu_stationarity.py).
"""

from __future__ import annotations

import ctypes
from typing import Optional, Sequence

import numpy as np

from . import _abi

FIELDS = ("stop", "cd_converged", "ratio", "mean_abs", "std", "strouhal", "peak_frequency", "peak_amplitude",
          "window", "relative_drift", "stationary", "lift", "amp_window", "amp_first_half", "amp_second_half")


def enable_ring(solver, cap: int) -> None:
    """Give the solver device rings of `cap` drag / lift samples, appended
    by every step from now on (cap 0 removes them)."""
    solver._ctx.call("aero_set_history_ring", ctypes.c_int64(int(cap)))
    solver._ring_base = solver.step_count


def stationarity(solvers: Sequence, counts: Optional[Sequence[int]] = None, convergence_window: int = 5000,
                 convergence_tol: float = 0.01, strouhal_window: int = 4096,
                 strouhal_tol: float = 0.05) -> np.ndarray:
    """(n, NSTAT) results for solvers with rings, over the first counts[i]
    samples each ring received (default: every step since enable_ring)."""
    n = len(solvers)
    out = np.zeros((n, _abi.NSTAT))
    if n == 0:
        return out
    if counts is None:
        counts = [s.step_count - getattr(s, "_ring_base", 0) for s in solvers]
    cnt = np.ascontiguousarray(counts, dtype=np.int64)
    ctxs = (ctypes.c_void_p * n)(*[s._ctx.handle.value for s in solvers])
    _abi.check(_abi.lib().aero_stationarity(ctypes.c_int32(n), ctxs, _abi.ptr(cnt), ctypes.c_int32(convergence_window),
                                            ctypes.c_double(convergence_tol), ctypes.c_int32(strouhal_window),
                                            ctypes.c_double(strouhal_tol), _abi.ptr(out)))
    return out


def stationarity_arrays(cases: Sequence[tuple], convergence_window: int = 5000, convergence_tol: float = 0.01,
                        strouhal_window: int = 4096, strouhal_tol: float = 0.05) -> np.ndarray:
    """The same test on host histories: cases of (Cd, Cl_y, Cl_z or None, D, u0)."""
    n = len(cases)
    out = np.zeros((n, _abi.NSTAT))
    if n == 0:
        return out
    keep = []
    cd, ly, lz = (ctypes.c_void_p * n)(), (ctypes.c_void_p * n)(), (ctypes.c_void_p * n)()
    lens = np.zeros(n, dtype=np.int64)
    D = np.zeros(n)
    u0 = np.zeros(n)
    for i, (a, b, c, d, u) in enumerate(cases):
        arrs = [np.ascontiguousarray(a, dtype=np.float64), np.ascontiguousarray(b, dtype=np.float64),
                None if c is None else np.ascontiguousarray(c, dtype=np.float64)]
        keep.append(arrs)
        cd[i], ly[i] = arrs[0].ctypes.data, arrs[1].ctypes.data
        lz[i] = None if arrs[2] is None else arrs[2].ctypes.data
        lens[i], D[i], u0[i] = len(arrs[0]), d, u
    _abi.check(_abi.lib().aero_stationarity_arrays(
        ctypes.c_int32(n), cd, ly, lz, _abi.ptr(lens), _abi.ptr(D), _abi.ptr(u0), ctypes.c_int32(convergence_window),
        ctypes.c_double(convergence_tol), ctypes.c_int32(strouhal_window), ctypes.c_double(strouhal_tol),
        _abi.ptr(out)))
    return out
