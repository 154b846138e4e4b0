"""Launch autotuner for the batched step kernel.

The library keeps a tuning table (include/aero_hip.h, aero_tune_set): per
kernel group -- lattice, collision, storage, scalar transport -- and grid
class (floor(log2) of the group's tile count), how a step launch orders and
walks its tiles:

* ``xcd_run``: 0 the built-in rule, -1 the natural order, k > 0 runs of k
  tiles per XCD (MI355X hands workgroups to its 8 XCDs round-robin, each with
  its own L2);
* ``nt``: non-temporal population stores, -1 the built-in rule (on when the
  launch's populations exceed most of the 256 MB last-level cache), 0 / 1;
* ``graph``: 1 replays each chunk's step launches from a captured HIP graph;
* ``rowgroups``: the wave shape, -1 the built-in one (the 2-byte D2Q9
  kernels step bulk rows four to a wave, from a per-case wave list), 0 one
  row per wave.

None of them changes a result: every cell computes the same bits, and the
per-case force reduction keeps its fixed order.  ``tune`` measures a
workload -- solvers that step together, e.g. one large case or a serving
batch -- knob by knob (coordinate descent from the current entry) and keeps
the fastest value when it beats the incumbent by more than ``min_gain``.
The table persists as JSON (``save`` / ``load``); ``load_default`` reads
``aero_amd/tuning.json`` when present, and the library applies it at load.
The reference has no counterpart: its steps are Numba loops.
"""

from __future__ import annotations

import ctypes
import json
import pathlib
import time
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import _abi

DEFAULT_PATH = pathlib.Path(__file__).resolve().parent / "tuning.json"
KNOBS = (("xcd_run", (0, -1, 8, 32, 128)), ("nt", (-1, 0, 1)), ("graph", (0, 1)), ("rowgroups", (-1, 0)))
DEFAULT = {"xcd_run": 0, "nt": -1, "graph": 0, "rowgroups": -1}
NAMES = tuple(n for n, _ in KNOBS)

# entries set in this process: (q, collision, storage, thermal, grid class) -> knobs
_table: Dict[Tuple[int, int, int, int, int], dict] = {}


def grid_class(ntiles: int) -> int:
    c = ctypes.c_int32(0)
    _abi.check(_abi.lib().aero_tune_grid_class(ctypes.c_int64(int(ntiles)), ctypes.byref(c)))
    return int(c.value)


def set_entry(key, knobs: dict) -> None:
    k = {**DEFAULT, **knobs}
    _abi.check(_abi.lib().aero_tune_set(*[ctypes.c_int32(int(v)) for v in key],
                                        *[ctypes.c_int32(int(k[n])) for n in NAMES]))
    _table[tuple(int(v) for v in key)] = k


def clear() -> None:
    _abi.check(_abi.lib().aero_tune_clear())
    _table.clear()


def table() -> dict:
    return {k: dict(v) for k, v in _table.items()}


def _tiles(solver) -> int:
    n = ctypes.c_int64(0)
    solver._ctx.call("aero_step_tiles", ctypes.byref(n))
    return int(n.value)


def group_keys(solvers: Sequence) -> Dict[tuple, list]:
    """The kernel groups a batch of these solvers forms, keyed as the
    library's tuning table is."""
    groups: Dict[tuple, list] = {}
    for s in solvers:
        d = s._ctx.desc
        groups.setdefault((d.q, d.collision, d.storage, int(d.thermal != 0)), []).append(s)
    return {k + (grid_class(sum(_tiles(s) for s in v)),): v for k, v in groups.items()}


def lists_waves(key) -> bool:
    """Kernel groups whose step walks a per-case wave list with grouped bulk
    rows (the rowgroups knob applies): 2-byte D2Q9 without the scalar."""
    q, _coll, st, th = key[:4]
    return q == 9 and st in (2, 3) and not th


def _time(solvers, steps: int, batch, reps: int) -> float:
    out = []
    for _ in range(reps):
        for s in solvers:
            s._ctx.call("aero_synchronize")
        t0 = time.perf_counter()
        batch.advance(solvers, steps, record=False)
        out.append(time.perf_counter() - t0)
    return float(np.median(out))


def _ab(solvers, steps, batch, key, incumbent: dict, trial: dict, reps: int) -> float:
    """Median time ratio trial / incumbent over `reps` back-to-back pairs, so
    clock or thermal drift during tuning affects both sides alike."""
    ratios = []
    for _ in range(reps):
        set_entry(key, incumbent)
        t_inc = _time(solvers, steps, batch, 1)
        set_entry(key, trial)
        t_new = _time(solvers, steps, batch, 1)
        ratios.append(t_new / t_inc)
    return float(np.median(ratios))


def tune(solvers: Sequence, steps: int = 100, reps: int = 3, min_gain: float = 0.02, batch=None,
         log: Optional[list] = None, warm_seconds: float = 0.5) -> Dict[tuple, dict]:
    """Tune every kernel group of this workload in place (the solvers
    advance by the measured steps).  Each candidate is timed against the
    incumbent in back-to-back pairs and replaces it when the median ratio
    shows a gain above `min_gain`.  Returns {key: knobs}."""
    from .batch import Batch
    own = batch is None
    batch = batch or Batch()
    try:
        # tile table, buffers, kernels and device clocks warm
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < warm_seconds:
            batch.advance(list(solvers), steps, record=False)
        for key in group_keys(solvers):
            best = dict(_table.get(key, DEFAULT))
            for name, values in KNOBS:
                if name == "rowgroups" and not lists_waves(key):
                    continue
                for v in values:
                    if v == best[name]:
                        continue
                    trial = {**best, name: v}
                    r = _ab(solvers, steps, batch, key, best, trial, reps)
                    if log is not None:
                        log.append({"key": list(key), **trial, "ratio_to_incumbent": r})
                    if r < 1.0 - min_gain:
                        best = trial
                set_entry(key, best)
    finally:
        if own:
            batch.close()
    return {k: table()[k] for k in group_keys(solvers)}


def save(path=DEFAULT_PATH, entries: Optional[dict] = None) -> None:
    entries = table() if entries is None else entries
    rows = [{"q": k[0], "collision": k[1], "storage": k[2], "thermal": k[3], "grid_class": k[4], **v}
            for k, v in sorted(entries.items())]
    pathlib.Path(path).write_text(json.dumps({"device": "MI355X (gfx950)", "entries": rows}, indent=1))


def load(path=DEFAULT_PATH) -> int:
    """Set every entry of a saved table; returns how many."""
    data = json.loads(pathlib.Path(path).read_text())
    for r in data["entries"]:
        set_entry((r["q"], r["collision"], r["storage"], r["thermal"], r["grid_class"]),
                  {n: r.get(n, DEFAULT[n]) for n in NAMES})
    return len(data["entries"])


def load_default() -> int:
    return load(DEFAULT_PATH) if DEFAULT_PATH.exists() else 0


def _workloads():
    """The standard workloads the shipped table is tuned on: the D3Q19 TRT
    sphere at 512 x 256 x 256 and 128 x 64 x 64 in FP32 and FP16, the
    64-case batched 2D draw in FP32 and FP16, one case of that draw alone,
    and the 400 x 200 cylinder alone in FP32 and FP16."""
    from . import serve
    out = []
    for st in ("f32", "f16"):
        out.append((f"sphere512_{st}", 40 if st == "f32" else 60,
                    [serve.CaseSpec(dim=3, nx=512, ny=256, nz=256, shape="sphere", size=(32.0,), Re=100.0,
                                    collision="trt", storage=st)]))
    for st in ("f32", "f16"):
        out.append((f"sphere128_{st}", 400, [serve.CaseSpec(dim=3, nx=128, ny=64, nz=64, shape="sphere", size=(8.0,),
                                                            Re=100.0, collision="trt", storage=st)]))
    for st in ("f32", "f16"):
        out.append((f"c4_64_{st}", 200, serve.c4_cases(64, seed=0, storage=st)))
    out.append(("single2d_f32", 500, serve.c4_cases(1, seed=0)))
    # the 400 x 200 cylinder alone (the reference's CPU-runnable configuration's grid)
    for st in ("f32", "f16"):
        out.append((f"cylinder400_{st}", 1000, [serve.CaseSpec(dim=2, nx=400, ny=200, shape="cylinder", size=(20.0,),
                                                             Re=100.0, collision="bgk", storage=st)]))
    return out


def main(argv=None) -> int:
    """python -m aero_amd.autotune [--out PATH]: tune the standard workloads
    on this device and write the table (default aero_amd/tuning.json); print
    one JSON line per workload with the built-in and tuned rates."""
    import argparse
    from .batch import Batch
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=str(DEFAULT_PATH))
    ap.add_argument("--reps", type=int, default=3)
    a = ap.parse_args(argv)
    clear()
    picked_all: Dict[tuple, dict] = {}
    for name, steps, specs in _workloads():
        solvers = [s.build() for s in specs]
        b = Batch()
        for key in group_keys(solvers):
            set_entry(key, DEFAULT)
        b.advance(solvers, 2, record=False)
        t0 = _time(solvers, steps, b, a.reps)
        picked = tune(solvers, steps=steps, reps=a.reps, batch=b)
        t1 = _time(solvers, steps, b, a.reps)
        cells = sum(s._N for s in solvers)
        picked_all.update(picked)
        print(json.dumps({"workload": name, "steps": steps, "builtin_MLUPS": round(cells * steps / t0 / 1e6, 1),
                          "tuned_MLUPS": round(cells * steps / t1 / 1e6, 1),
                          "picked": {str(list(k)): v for k, v in picked.items()}}), flush=True)
        b.close()
        for s in solvers:
            s.close()
    save(a.out, picked_all)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
