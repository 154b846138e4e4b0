"""Batched stepping of many solvers on one device.

``Batch.advance(solvers, n)`` advances every solver by n steps through
``aero_batch_step``: one kernel launch per step for each group of solvers
sharing lattice, collision, storage and scalar transport, over a packed list
of all their tiles.  Each solver computes the same bits as when stepped alone
(``Solver._advance`` runs the same kernel with a batch of one).  The
reference has no counterpart: it runs one case per process.
"""

from __future__ import annotations

import ctypes
from typing import List, Sequence

import numpy as np

from . import _abi


class Batch:
    """Owns one aero_batch (a stream and the batched launch's buffers)."""

    def __init__(self):
        self._h = _abi._P()
        _abi.check(_abi.lib().aero_batch_create(ctypes.byref(self._h)))

    def advance(self, solvers: Sequence, n: int, record: bool = True) -> List[np.ndarray]:
        """Advance every solver by n steps; returns each one's (n, 20)
        coefficient rows and, with record, appends them to its histories as
        ``run`` does."""
        m = len(solvers)
        if m == 0 or n <= 0:
            return [np.empty((0, _abi.NCOEF)) for _ in solvers]
        for s in solvers:
            if getattr(s, "_raise_on_step", None):
                raise s._raise_on_step
        planes = [s._planes(n) for s in solvers]
        hists = [np.empty((n, _abi.NCOEF), dtype=np.float64) for _ in solvers]
        ctxs = (ctypes.c_void_p * m)(*[s._ctx.handle.value for s in solvers])
        step0s = np.array([s.step_count for s in solvers], dtype=np.int64)
        pptr = (ctypes.c_void_p * m)(*[(p.ctypes.data if p is not None else None) for p in planes])
        hptr = (ctypes.c_void_p * m)(*[h.ctypes.data for h in hists])
        _abi.check(_abi.lib().aero_batch_step(self._h, ctypes.c_int32(m), ctxs, ctypes.c_int32(n),
                                              step0s.ctypes.data, pptr, hptr))
        for s, rows in zip(solvers, hists):
            s.step_count += n
            s._bump()
            if record:
                s._record(rows)
        return hists

    def check_stats(self, solvers: Sequence) -> np.ndarray:
        """check_stability's device statistics of every solver's current
        state (as Solver._device_stats, one row each), computed in one call."""
        m = len(solvers)
        out = np.empty((m, _abi.NCHECK), dtype=np.float64)
        if m:
            ctxs = (ctypes.c_void_p * m)(*[s._ctx.handle.value for s in solvers])
            _abi.check(_abi.lib().aero_batch_check(self._h, ctypes.c_int32(m), ctxs, out.ctypes.data))
        return out

    def close(self):
        if self._h:
            _abi.lib().aero_batch_destroy(self._h)
            self._h = _abi._P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
