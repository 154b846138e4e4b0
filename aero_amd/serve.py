"""Continuous-batching serving loop for wind-tunnel cases.

Requests arrive over time (each one a case: lattice, grid, geometry,
Reynolds number, collision model, storage, optional passive scalar).  The
server keeps every admitted case's populations resident on the device and
advances all live cases together, ``chunk`` steps at a time, through
``Batch.advance``: one kernel launch per step for each group of cases that
share a kernel.  Between chunks it

* runs the reference's stability check on every live case
  (check_stability, aero/diagnostics.py:32-71, on device statistics taken
  for all cases in one call at the start of the next chunk): a case it
  would stop with an error is retired as diverged, and its warnings are
  counted;
* retires every case that met the reference's auto-stop criterion
  (``detect_statistical_stationarity``, aero/diagnostics.py:197-228, on the
  drag and the lift history the reference picks, aero/lbm/solver3d.py:634-638,
  checked every ``chunk`` = check_every steps as in the reference run loop,
  aero/lbm/solver.py:517-570) or reached its step limit;
* admits waiting requests whose arrival time has passed while the live set
  stays under ``max_live`` cases and the device keeps ``mem_reserve`` bytes
  free after the new case's estimated footprint; admitted cases are built
  by ``build_workers`` host threads while the device steps, and join the
  batch when built.

A case's trajectory does not depend on what else is live (the batched
kernel computes each case's bits exactly as alone) and its stop decisions
look only at its own history, so every result is the same alone and in the
mix.  Timestamps give time-to-result (completion minus arrival) per request.
The reference has no counterpart: it runs one case per process.
"""

from __future__ import annotations

import math
import sys
import time
from dataclasses import dataclass, field
from typing import Callable, List, Optional, Sequence

import numpy as np

from .diagnostics import detect_statistical_stationarity, stability_findings


# ------------------------------------------------------------------ cases
@dataclass
class CaseSpec:
    """One wind-tunnel case, built into a solver on admission."""

    dim: int = 2
    nx: int = 256
    ny: int = 128
    nz: int = 1
    shape: str = "cylinder"          # cylinder | rectangle | polygon | sphere | box
    size: tuple = (16.0,)            # radius / (w, h) / circumradius / (w, h, d)
    vertices: Optional[list] = None  # polygon vertices, fractional coordinates
    Re: float = 100.0
    u0: float = 0.05
    collision: str = "bgk"
    storage: str = "f32"
    thermal: bool = False
    max_steps: int = 30000
    seed: int = 0
    les_cs: float = 0.0              # > 0: Smagorinsky LES from the local stress (batchable), this C_s

    @property
    def cells(self) -> int:
        return self.nx * self.ny * (self.nz if self.dim == 3 else 1)

    def geometry(self):
        from . import geometry as G
        if self.dim == 2:
            if self.shape == "cylinder":
                return G.Cylinder(radius=self.size[0], cx_frac=0.25, cy_frac=0.5)
            if self.shape == "rectangle":
                return G.Rectangle(width=self.size[0], height=self.size[1], cx_frac=0.25, cy_frac=0.5)
            if self.shape == "polygon":
                return G.Polygon(self.vertices)
            raise ValueError(f"unknown 2D shape {self.shape!r}")
        if self.shape == "sphere":
            return G.Sphere(radius=self.size[0], cx_frac=0.25)
        if self.shape == "box":
            return G.Box(width=self.size[0], height=self.size[1], depth=self.size[2], cx_frac=0.25)
        raise ValueError(f"unknown 3D shape {self.shape!r}")

    def reference_length(self, solid: np.ndarray) -> float:
        if self.shape == "polygon":
            # frontal height in cells (the reference leaves D for polygons to the caller)
            ys = np.nonzero(solid.any(axis=1))[0]
            return float(ys[-1] - ys[0] + 1) if ys.size else 1.0
        return float(self.geometry().reference_length())

    def device_bytes(self) -> int:
        """Footprint estimate used for admission: two population copies, the
        scalar's two copies, flags, link index and small per-row state."""
        q = 9 if self.dim == 2 else 19
        s = {"f64": 8, "f32": 4, "f16": 2, "bf16": 2}[self.storage]
        per_cell = 2 * q * s * (2 if self.thermal else 1) + 1 + 4
        return self.cells * per_cell + (1 << 20)

    def _les_kw(self) -> dict:
        if self.les_cs > 0.0:
            return dict(les=True, les_cs=float(self.les_cs), les_form="local")
        return {}

    def build(self):
        """The solver for this case, at equilibrium (rho = 1, u = u0) plus
        seeded 1e-4 transverse velocity noise, as the benchmark configs state."""
        import aero_amd
        from .lattice import compute_feq, compute_feq_3d
        g = self.geometry()
        rng = np.random.default_rng(self.seed)
        if self.dim == 2:
            solid = g.mark_solid(self.ny, self.nx)
            D = self.reference_length(solid)
            nu = self.u0 * D / self.Re
            omega = 1.0 / (3.0 * nu + 0.5)
            s = aero_amd.Solver(Ny=self.ny, Nx=self.nx, solid=solid, omega=omega, u0=self.u0, D=D,
                                collision=self.collision, trt_lambda=0.1875, trt_variant="lambda",
                                outlet_bc="zerogradient", sponge_thickness=max(8, self.nx // 20),
                                storage=self.storage, thermal=self.thermal, alpha_T=0.01, **self._les_kw())
            sh = (self.ny, self.nx)
            uy = rng.uniform(-1e-4, 1e-4, sh)
            s.set_equilibrium(uy=uy)  # feq(1, u0, uy) on the device
            s._f_outlet_prev = compute_feq(np.ones(self.ny), np.full(self.ny, self.u0), uy[:, -1])
        else:
            solid = g.mark_solid(self.nz, self.ny, self.nx)
            D = self.reference_length(solid)
            nu = self.u0 * D / self.Re
            omega = 1.0 / (3.0 * nu + 0.5)
            s = aero_amd.Solver3D(Nz=self.nz, Ny=self.ny, Nx=self.nx, solid=solid, omega=omega, u0=self.u0, D=D,
                                  collision=self.collision, trt_lambda=0.1875, trt_variant="lambda",
                                  outlet_bc="zerogradient", sponge_thickness=max(8, self.nx // 20),
                                  storage=self.storage, thermal=self.thermal, alpha_T=0.01, **self._les_kw())
            sh = (self.nz, self.ny, self.nx)
            uy = rng.uniform(-1e-4, 1e-4, sh)
            uz = rng.uniform(-1e-4, 1e-4, sh)
            s.set_equilibrium(uy=uy, uz=uz)  # feq(1, u0, uy, uz) on the device
            ps = (self.nz, self.ny)
            s._f_outlet_prev = compute_feq_3d(np.ones(ps), np.full(ps, self.u0), uy[..., -1], uz[..., -1])
        return s


def c4_cases(n: int = 256, seed: int = 0, max_steps: int = 30000, storage: str = "f32") -> List[CaseSpec]:
    """The batched 2D workload: nx uniform in [128, 1024] step 32, ny = nx/2;
    cylinder r in [8, 32], rectangle w, h in [8, 48] or convex polygon with
    5-9 vertices and circumradius in [10, 30]; Re log-uniform in [20, 200];
    collision uniform over BGK / TRT / MRT; u0 = 0.05."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        nx = int(rng.choice(np.arange(128, 1025, 32)))
        ny = nx // 2
        kind = ("cylinder", "rectangle", "polygon")[int(rng.integers(3))]
        verts = None
        if kind == "cylinder":
            size = (float(rng.uniform(8, 32)),)
        elif kind == "rectangle":
            size = (float(rng.uniform(8, 48)), float(rng.uniform(8, 48)))
        else:
            nv = int(rng.integers(5, 10))
            R = float(rng.uniform(10, 30))
            ang = np.sort(rng.uniform(0, 2 * math.pi, nv))
            cx, cy = 0.25 * nx, 0.5 * ny
            verts = [((cx + R * math.cos(a)) / nx, (cy + R * math.sin(a)) / ny) for a in ang]
            size = (R,)
        Re = float(math.exp(rng.uniform(math.log(20.0), math.log(200.0))))
        coll = ("bgk", "trt", "mrt")[int(rng.integers(3))]
        out.append(CaseSpec(dim=2, nx=nx, ny=ny, shape=kind, size=size, vertices=verts, Re=Re, collision=coll,
                            storage=storage, max_steps=max_steps, seed=seed * 100003 + k))
    return out


def c5_requests(n: int, rate: float, seed: int = 0, max_steps: int = 30000,
                storage: str = "f32") -> List[tuple]:
    """Poisson arrival stream (rate requests/s) of mixed cases: 80% 2D as in
    the batched workload, 20% 3D spheres / boxes on grids from 64^3 to
    256x128x128 with Re uniform in [50, 300]; a passive scalar on 25%."""
    rng = np.random.default_rng(seed)
    t = 0.0
    two_d = iter(c4_cases(n, seed=seed + 1, max_steps=max_steps, storage=storage))
    grids = [(64, 64, 64), (128, 64, 64), (192, 96, 96), (256, 128, 128)]
    out = []
    for k in range(n):
        t += float(rng.exponential(1.0 / rate))
        if rng.random() < 0.8:
            spec = next(two_d)
        else:
            nx, ny, nz = grids[int(rng.integers(len(grids)))]
            if rng.random() < 0.5:
                spec = CaseSpec(dim=3, nx=nx, ny=ny, nz=nz, shape="sphere", size=(ny / 8.0,))
            else:
                h = ny / 6.0
                spec = CaseSpec(dim=3, nx=nx, ny=ny, nz=nz, shape="box", size=(h, h, h))
            spec.Re = float(rng.uniform(50, 300))
            spec.collision = ("bgk", "trt", "mrt")[int(rng.integers(3))]
            spec.storage = storage
            spec.max_steps = max_steps
            spec.seed = seed * 100003 + 50000 + k
        spec.thermal = bool(rng.random() < 0.25)
        out.append((t, spec))
    return out


# ---------------------------------------------------------------- serving
@dataclass
class Result:
    rid: int
    spec: CaseSpec
    arrival: float
    admitted: float
    finished: float
    steps: int
    stop_reason: str
    Cd_mean: float
    Cl_std: float
    strouhal: Optional[float]
    warnings: int = 0  # stability checks that warned (Ma, density drift)
    Cd_history: list = field(repr=False, default_factory=list)

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


@dataclass
class _Live:
    rid: int
    spec: CaseSpec
    solver: object
    arrival: float
    admitted: float
    warnings: int = 0
    error: Optional[str] = None  # set when check_stability would have raised
    steps: int = 0               # steps taken, the last chunk's included
    rows: Optional[np.ndarray] = None  # the last chunk's rows, not yet in the histories


# Thread switch interval (seconds) while Server.serve runs.
GIL_SWITCH = 2e-4


class Server:
    """Continuous batching of wind-tunnel cases on one device."""

    def __init__(self, chunk: int = 500, max_live: int = 512, mem_reserve: int = 8 << 30,
                 convergence_window: int = 5000, convergence_tol: float = 0.01, strouhal_window: int = 4096,
                 strouhal_tol: float = 0.05, batch=None, free_memory: Optional[Callable[[], int]] = None,
                 keep_history: bool = False, build_workers: int = 16, device_stats: bool = True,
                 on_result: Optional[Callable[["Result"], None]] = None):
        """on_result, when given, is called with each Result as it is decided."""
        self.chunk = int(chunk)
        self.on_result = on_result
        self.build_workers = max(1, int(build_workers))
        self.max_live = int(max_live)
        self.mem_reserve = int(mem_reserve)
        self.stop_kw = dict(convergence_window=convergence_window, convergence_tol=convergence_tol,
                            strouhal_window=strouhal_window, strouhal_tol=strouhal_tol)
        self._batch = batch
        self._free_memory = free_memory
        self.keep_history = keep_history
        # the auto-stop test on the device (aero_amd/stationarity.py) over rings
        # of each case's drag and lift history, one call for every live case;
        # the host restatement when the Strouhal window exceeds the device's
        from . import _abi
        self.device_stats = bool(device_stats) and 1 <= strouhal_window <= _abi.STAT_MAX_WINDOW
        self.steps_done = 0      # lattice updates, for aggregate MLUPS
        self.busy_seconds = 0.0  # time inside batched stepping

    def _batch_obj(self):
        if self._batch is None:
            from .batch import Batch
            self._batch = Batch()
        return self._batch

    def _free(self) -> int:
        if self._free_memory is not None:
            return int(self._free_memory())
        from . import _abi
        free = _abi.ctypes.c_int64()
        total = _abi.ctypes.c_int64()
        _abi.check(_abi.lib().aero_mem_info(_abi.ctypes.byref(free), _abi.ctypes.byref(total)))
        return int(free.value)

    def _stops(self, lvs: Sequence[_Live]) -> dict:
        """The auto-stop decision of each case on its history so far: one
        device call for all of them, or the host test per case."""
        lvs = [lv for lv in lvs if lv.solver.Cd_history]
        if not lvs:
            return {}
        if self.device_stats and all(getattr(lv.solver, "_ring_base", None) is not None for lv in lvs):
            from .stationarity import stationarity
            out = stationarity([lv.solver for lv in lvs], counts=[len(lv.solver.Cd_history) for lv in lvs],
                               **self.stop_kw)
            return {id(lv): bool(row[0]) for lv, row in zip(lvs, out)}
        return {id(lv): detect_statistical_stationarity(
            lv.solver.Cd_history, lv.solver.lift_history(self.stop_kw["strouhal_window"]), lv.solver.D,
            lv.solver.u0, **self.stop_kw)[2] for lv in lvs}

    def _done(self, lv: _Live, stop: Optional[bool] = None) -> Optional[str]:
        s = lv.solver
        if lv.error is not None or not np.isfinite(s.Cd_history[-1]):
            return "diverged"
        if stop is None:
            stop = self._stops([lv]).get(id(lv), False)
        if stop:
            return "converged"
        if len(s.Cd_history) >= lv.spec.max_steps:
            return "max_steps"
        return None

    def _result(self, lv: _Live, reason: str, now: float, upto: int) -> Result:
        s = lv.solver
        cd = np.asarray(s.Cd_history[:upto])
        cl = np.asarray(s.lift_history(self.stop_kw["strouhal_window"])[:upto])
        w = max(1, len(cd) // 5)
        from .diagnostics import compute_strouhal
        st = compute_strouhal(cl.tolist(), s.D, s.u0) if reason != "diverged" else None
        res = Result(lv.rid, lv.spec, lv.arrival, lv.admitted, now, upto, reason,
                     float(np.mean(cd[-w:])), float(np.std(cl[-w:])), st, lv.warnings,
                     cd.tolist() if self.keep_history else [])
        if self.on_result is not None:
            self.on_result(res)
        return res

    def _build(self, spec: CaseSpec):
        """A builder thread's work: the case's solver with its history ring."""
        solver = spec.build()
        if self.device_stats and hasattr(solver, "_ctx"):
            from .stationarity import enable_ring
            w = max(self.stop_kw["convergence_window"], self.stop_kw["strouhal_window"])
            enable_ring(solver, w + self.chunk)  # room for the chunk stepped during the check
        return solver

    @staticmethod
    def _join_built(building: list, live: list, clock, t0: float) -> None:
        """Move finished builds into the live set."""
        for b in [b for b in building if b[0].done()]:
            building.remove(b)
            fut, arr, rid, spec, _ = b
            live.append(_Live(rid, spec, fut.result(), arr, clock() - t0))

    @staticmethod
    def _chunk(batch, solvers, n):
        """Device work of one chunk: check_stability's statistics of every
        case's current state (one call for all), then n steps of every case."""
        stats = batch.check_stats(solvers)
        return stats, batch.advance(solvers, n, False)

    @staticmethod
    def _take_rows(lv: "_Live") -> None:
        """Record the case's last chunk of coefficient rows in its histories."""
        if lv.rows is not None:
            lv.solver._record(lv.rows)
            lv.rows = None

    def _stability(self, lvs: Sequence["_Live"], stats) -> None:
        """check_stability on each case's state at a chunk boundary (its
        history so far), as the reference's run loop does every check_every
        steps; stats holds one row of device statistics per case."""
        for lv, st in zip(lvs, stats):
            err, warns = stability_findings(st, len(lv.solver.Cd_history))
            lv.warnings += len(warns) > 0
            if err is not None:
                lv.error = err

    def serve(self, requests: Sequence[tuple], clock: Callable[[], float] = time.perf_counter,
              sleep: Callable[[float], None] = time.sleep) -> List[Result]:
        """Serve (arrival_seconds, CaseSpec) requests; returns one Result per
        request in completion order.

        Admitted requests are built (geometry, links, device state) by a pool
        of builder threads while the device steps the live cases; a built
        case joins the batch at the next chunk.  Each chunk's device work (a
        worker thread holds the call) first takes every case's stability
        statistics at the chunk boundary, then steps; meanwhile the host
        records the previous chunk's rows and runs the stop test on them.  A
        case found done at the end of chunk j (check_stability's error, then
        the auto-stop test, the reference's order) is retired with its
        history cut at the end of chunk j, exactly what it would have been
        alone; its speculative chunk j+1 is discarded.  A case whose history
        has reached max_steps is decided before the next chunk instead (it
        stops whatever the checks find), so it costs no speculative chunk.
        A retired case's Result is written, and its device memory released,
        while the next chunk steps (at once when admission waits for memory).

        While it runs, the interpreter's thread switch interval is cut to
        GIL_SWITCH seconds (restored on return): the thread that drives the
        device needs the GIL briefly around each call, and waiting out the
        default 5 ms behind the host's bookkeeping would leave the device
        idle each time."""
        from concurrent.futures import FIRST_COMPLETED, ThreadPoolExecutor, wait
        if any(r[1].shape == "polygon" for r in requests):
            import matplotlib.path  # noqa: F401  (imported once here, not raced by the builder threads)
        order = sorted(range(len(requests)), key=lambda i: requests[i][0])
        waiting = [(requests[i][0], i, requests[i][1]) for i in order]
        live: List[_Live] = []
        building: list = []  # (future, arrival, rid, spec, bytes)
        done: List[Result] = []
        t0 = clock()
        pool = ThreadPoolExecutor(max_workers=1)
        builders = ThreadPoolExecutor(max_workers=self.build_workers)
        switch = sys.getswitchinterval()
        sys.setswitchinterval(min(switch, GIL_SWITCH))
        retired: list = []  # (live case, reason, history length): decided, not yet written up

        # The cases retired at a decision are written up while the next chunk
        # steps, and their device memory is released last (a free of device
        # memory waits for the device's work in flight).
        def report():
            now = clock() - t0
            for lv, reason, upto in retired:
                done.append(self._result(lv, reason, now, upto))

        def release():
            for lv, _, _ in retired:
                lv.solver.close()
            retired.clear()

        def write_up():
            report()
            release()

        try:
            while waiting or live or building:
                # finished builds join first, so their memory is counted once
                # (in the device's free memory) rather than also as pending
                self._join_built(building, live, clock, t0)
                now = clock() - t0
                free = None
                pending = sum(b[4] for b in building)
                while waiting and waiting[0][0] <= now and len(live) + len(building) < self.max_live:
                    arr, rid, spec = waiting[0]
                    need = spec.device_bytes()
                    if free is None:
                        free = self._free()
                    if (live or building) and free - pending - need < self.mem_reserve:
                        if retired:  # release the retired cases' memory now
                            write_up()
                            free = self._free()
                            continue
                        break  # wait for a live case to retire
                    waiting.pop(0)
                    building.append((builders.submit(self._build, spec), arr, rid, spec, need))
                    pending += need
                self._join_built(building, live, clock, t0)
                # A case whose history has reached max_steps stops at this check
                # whatever it finds, so it is decided now rather than stepped for
                # a speculative chunk: its last chunk's rows are recorded, the
                # stability check of its last state runs, then the stop test.
                at_limit = [lv for lv in live if lv.steps >= lv.spec.max_steps]
                if at_limit:
                    for lv in at_limit:
                        self._take_rows(lv)
                    self._stability(at_limit, self._batch_obj().check_stats([lv.solver for lv in at_limit]))
                    stops = self._stops(at_limit)
                    for lv in at_limit:
                        retired.append((lv, self._done(lv, stops.get(id(lv), False)), len(lv.solver.Cd_history)))
                        live.remove(lv)
                if not live:
                    write_up()
                    if building:
                        wait([b[0] for b in building], return_when=FIRST_COMPLETED)
                    elif waiting:
                        sleep(max(0.0, waiting[0][0] - (clock() - t0)))
                    continue
                n = self.chunk
                batch = self._batch_obj()
                stepping = list(live)
                tb = clock()
                # the device checks every case's state at this boundary (the end
                # of its last chunk), then steps the next chunk
                fut = pool.submit(self._chunk, batch, [lv.solver for lv in stepping], n)
                # while it runs: the cases retired at the boundary are written
                # up, the last chunk's rows join the histories and the stop test
                # reads them
                report()
                for lv in stepping:
                    self._take_rows(lv)
                stops = self._stops(stepping)
                release()
                stats, rows = fut.result()
                self.busy_seconds += clock() - tb
                # decisions at the boundary: check_stability first, as the
                # reference's run loop, then the stop test; a case they retire
                # keeps its history through the boundary and its speculative
                # chunk is discarded
                checked = [k for k, lv in enumerate(stepping) if lv.solver.Cd_history]
                self._stability([stepping[k] for k in checked], stats[checked])
                for lv, r in zip(stepping, rows):
                    reason = self._done(lv, stops.get(id(lv), False)) if lv.solver.Cd_history else None
                    if reason is not None:
                        retired.append((lv, reason, len(lv.solver.Cd_history)))
                        live.remove(lv)
                    else:
                        lv.rows = r
                        lv.steps += n
                        self.steps_done += n * lv.spec.cells
            write_up()
        finally:
            builders.shutdown(wait=True)
            pool.shutdown(wait=True)
            sys.setswitchinterval(switch)
        return done


def summarize(results: Sequence[Result], wall: float, updates: int, stepping_seconds: Optional[float] = None) -> dict:
    """Solves per hour and p50 / p99 time-to-result over the requests that
    finished without diverging (converged or at their step limit); diverged
    requests and requests whose stability checks warned are counted apart.
    Aggregate MLUPS over the whole wall time and over the time spent inside
    batched stepping count every lattice update, diverged cases' included."""
    solved = [r for r in results if r.stop_reason != "diverged"]
    lat = np.array([r.latency for r in solved]) if solved else np.zeros(1)
    reasons = {}
    for r in results:
        reasons[r.stop_reason] = reasons.get(r.stop_reason, 0) + 1
    return {"requests": len(results), "solves": len(solved), "diverged": len(results) - len(solved),
            "warned": sum(1 for r in results if r.warnings), "wall_s": round(wall, 3),
            "solves_per_hour": round(len(solved) / wall * 3600.0, 1) if wall > 0 else None,
            "p50_s": round(float(np.percentile(lat, 50)), 3), "p99_s": round(float(np.percentile(lat, 99)), 3),
            "MLUPS": round(updates / wall / 1e6, 1) if wall > 0 else None,
            "stepping_MLUPS": round(updates / stepping_seconds / 1e6, 1) if stepping_seconds else None,
            "stop_reasons": reasons}
