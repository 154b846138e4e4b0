"""Serving loop logic on the CPU: workload generators and the admission /
retirement policy, with a stand-in for the device batch."""

import math

import numpy as np
import pytest

from aero_amd import serve


def test_c4_workload_matches_its_definition():
    cases = serve.c4_cases(256, seed=0)
    assert len(cases) == 256
    nx = np.array([c.nx for c in cases])
    assert nx.min() >= 128 and nx.max() <= 1024 and np.all(nx % 32 == 0)
    assert all(c.ny == c.nx // 2 for c in cases)
    assert {c.shape for c in cases} == {"cylinder", "rectangle", "polygon"}
    assert {c.collision for c in cases} == {"bgk", "trt", "mrt"}
    Re = np.array([c.Re for c in cases])
    assert Re.min() >= 20 and Re.max() <= 200
    # log-uniform: the median sits near the geometric mean of the range
    assert abs(math.log(np.median(Re)) - 0.5 * (math.log(20) + math.log(200))) < 0.3
    for c in cases:
        if c.shape == "cylinder":
            assert 8 <= c.size[0] <= 32
        elif c.shape == "rectangle":
            assert all(8 <= v <= 48 for v in c.size)
        else:
            assert 5 <= len(c.vertices) <= 9 and 10 <= c.size[0] <= 30
    # the same seed gives the same workload
    again = serve.c4_cases(256, seed=0)
    assert [(c.nx, c.shape, c.Re, c.collision) for c in again] == [(c.nx, c.shape, c.Re, c.collision) for c in cases]


def test_polygon_cases_are_convex_and_inside_the_domain():
    for c in serve.c4_cases(64, seed=3):
        if c.shape != "polygon":
            continue
        v = np.array(c.vertices) * [c.nx, c.ny]
        assert np.all(v[:, 0] > 0) and np.all(v[:, 0] < c.nx) and np.all(v[:, 1] > 0) and np.all(v[:, 1] < c.ny)
        # vertices sorted by angle on a circle: every turn has the same sign
        d = np.roll(v, -1, axis=0) - v
        cross = d[:, 0] * np.roll(d, -1, axis=0)[:, 1] - d[:, 1] * np.roll(d, -1, axis=0)[:, 0]
        assert np.all(cross > 0) or np.all(cross < 0)


def test_c5_stream_is_poisson_with_the_stated_mix():
    reqs = serve.c5_requests(2000, rate=10.0, seed=1)
    t = np.array([r[0] for r in reqs])
    gaps = np.diff(np.r_[0.0, t])
    assert abs(gaps.mean() - 0.1) < 0.01
    dims = np.array([r[1].dim for r in reqs])
    assert abs((dims == 3).mean() - 0.2) < 0.03
    assert abs(np.mean([r[1].thermal for r in reqs]) - 0.25) < 0.03
    for _, s in reqs:
        if s.dim == 3:
            assert 50 <= s.Re <= 300 and s.shape in ("sphere", "box")
            assert (s.nx, s.ny, s.nz) in [(64, 64, 64), (128, 64, 64), (192, 96, 96), (256, 128, 128)]


class _FakeSolver:
    def __init__(self, stop_at, diverge=False):
        self.stop_at, self.diverge = stop_at, diverge
        self.step_count = 0
        self.Cd_history, self.Cl_history = [], []
        self.D, self.u0 = 10.0, 0.05
        self.closed = False

    def _record(self, rows):
        self.Cd_history.extend(rows[:, 0].tolist())
        self.Cl_history.extend(rows[:, 1].tolist())

    def lift_history(self, window=4096):
        return self.Cl_history

    def _device_stats(self):
        # nan flag, non-positive density flag, max |u|, sum and count of density
        if self.diverge:
            return np.array([1.0, 0.0, 0.0, 1.0, 1.0])
        return np.array([0.0, 0.0, 0.4 if self.stop_at == 7 else 0.1, 1.0, 1.0])

    def close(self):
        self.closed = True


class _FakeSpec(serve.CaseSpec):
    def build(self):
        return _FakeSolver(self.seed, diverge=self.Re < 0)


class _FakeBatch:
    def __init__(self, clock):
        self.clock, self.calls = clock, []

    def check_stats(self, solvers):
        return np.array([s._device_stats() for s in solvers]).reshape(len(solvers), 5)

    def advance(self, solvers, n, record=True):
        self.calls.append(len(solvers))
        self.clock[0] += 0.01 * n / 100
        out = []
        for s in solvers:
            s.step_count += n
            val = float("nan") if s.diverge else 1.0
            rows = np.zeros((n, 20))
            rows[:, 0] = val
            out.append(rows)
        return out


def test_server_admits_retires_and_times_requests(monkeypatch):
    clock = [0.0]
    batch = _FakeBatch(clock)
    # "converged" once step_count reaches stop_at: patch the criterion
    monkeypatch.setattr(serve, "detect_statistical_stationarity",
                        lambda cd, cl, D, u0, **kw: (None, None, False))
    srv = serve.Server(chunk=100, max_live=3, mem_reserve=0, batch=batch, free_memory=lambda: 1 << 40)
    reqs = [(0.0, _FakeSpec(seed=0, max_steps=300)), (0.0, _FakeSpec(seed=0, max_steps=500)),
            (0.005, _FakeSpec(seed=0, max_steps=200)), (0.0, _FakeSpec(seed=0, max_steps=100)),
            (0.5, _FakeSpec(seed=0, max_steps=100, Re=-1.0))]
    res = srv.serve(reqs, clock=lambda: clock[0], sleep=lambda dt: clock.__setitem__(0, clock[0] + dt))
    assert sorted(r.rid for r in res) == [0, 1, 2, 3, 4]
    assert max(batch.calls) <= 3
    by = {r.rid: r for r in res}
    assert by[0].steps == 300 and by[1].steps == 500 and by[3].steps == 100 and by[2].steps == 200
    assert by[4].stop_reason == "diverged" and by[4].steps == 100
    assert all(r.stop_reason == "max_steps" for r in res if r.rid != 4)
    assert all(r.latency >= 0 and r.admitted >= r.arrival for r in res)
    # request 2 arrives after the first chunk started: it cannot be admitted before 0.005
    assert by[2].admitted >= 0.005
    s = serve.summarize(res, wall=clock[0], updates=srv.steps_done)
    assert s["requests"] == 5 and s["p99_s"] >= s["p50_s"]
    # the diverged request is not a solve
    assert s["solves"] == 4 and s["diverged"] == 1 and s["warned"] == 0
    assert s["solves_per_hour"] == pytest.approx(4 / clock[0] * 3600.0, rel=1e-3)


def test_server_counts_stability_warnings_and_stops_on_errors():
    """check_stability at every chunk boundary: a warning (Ma above 0.3 here)
    is counted on the result; the error the reference raises retires the
    case as diverged with its history through that chunk."""
    clock = [0.0]
    batch = _FakeBatch(clock)
    srv = serve.Server(chunk=100, max_live=4, mem_reserve=0, batch=batch, free_memory=lambda: 1 << 40)
    reqs = [(0.0, _FakeSpec(seed=7, max_steps=300)), (0.0, _FakeSpec(seed=0, max_steps=300)),
            (0.0, _FakeSpec(seed=0, max_steps=300, Re=-1.0))]
    res = srv.serve(reqs, clock=lambda: clock[0], sleep=lambda dt: clock.__setitem__(0, clock[0] + dt))
    by = {r.rid: r for r in res}
    assert by[0].warnings == 3 and by[0].stop_reason == "max_steps"
    assert by[1].warnings == 0
    assert by[2].stop_reason == "diverged" and by[2].steps == 100
    s = serve.summarize(res, wall=clock[0], updates=srv.steps_done)
    assert (s["solves"], s["diverged"], s["warned"]) == (2, 1, 1)


class _CountedSpec(_FakeSpec):
    live = 0
    peak = 0

    def build(self):
        _CountedSpec.live += 1
        _CountedSpec.peak = max(_CountedSpec.peak, _CountedSpec.live)
        s = super().build()
        s.close = lambda: setattr(_CountedSpec, "live", _CountedSpec.live - 1)
        return s


def test_server_holds_requests_until_memory_frees():
    clock = [0.0]
    batch = _FakeBatch(clock)
    need = serve.CaseSpec(nx=256, ny=128).device_bytes()
    total = int(2.5 * need)  # the device holds two such cases
    srv = serve.Server(chunk=100, max_live=10, mem_reserve=0, batch=batch,
                       free_memory=lambda: total - _CountedSpec.live * need)
    reqs = [(0.0, _CountedSpec(nx=256, ny=128, seed=0, max_steps=200)) for _ in range(5)]
    res = srv.serve(reqs, clock=lambda: clock[0], sleep=lambda dt: clock.__setitem__(0, clock[0] + dt))
    assert len(res) == 5
    # never more than the two cases the memory holds (whether two are ever
    # live together depends on how the builder threads are scheduled: a build
    # in progress counts both in the free memory and as pending)
    assert 1 <= _CountedSpec.peak <= 2 and max(batch.calls) <= 2 and _CountedSpec.live == 0


def test_server_restores_the_thread_switch_interval_and_closes_every_case():
    """serve() lowers the interpreter's switch interval while it runs and
    puts it back; every case it built is closed by the time it returns,
    including those retired at its last decision."""
    import sys
    clock = [0.0]
    batch = _FakeBatch(clock)
    _CountedSpec.live = _CountedSpec.peak = 0
    before = sys.getswitchinterval()
    srv = serve.Server(chunk=100, max_live=8, mem_reserve=0, batch=batch, free_memory=lambda: 1 << 40)
    reqs = [(0.0, _CountedSpec(seed=0, max_steps=100 * (k % 3 + 1))) for k in range(6)]
    res = srv.serve(reqs, clock=lambda: clock[0], sleep=lambda dt: clock.__setitem__(0, clock[0] + dt))
    assert sys.getswitchinterval() == before
    assert len(res) == 6 and _CountedSpec.live == 0
    assert sorted(r.steps for r in res) == [100, 100, 200, 200, 300, 300]
