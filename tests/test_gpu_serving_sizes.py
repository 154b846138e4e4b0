"""The batched and serving workloads at their stated sizes.

* Batched (configs: 256 independent 2D D2Q9 cases, nx in [128, 1024]): 32
  cases of the seed-0 draw -- its three nx = 1024 cases and the next 29 --
  covering BGK, TRT and MRT, stepped together through one Batch for 3000
  steps (chunks of 256) and each alone.  Coefficient rows must be
  bit-identical.  Every case whose obstacle blocks at most 40% of the channel
  height stays finite with its mean drag over the last 1000 steps in
  [0.5, 10] (measured: 1.25-4.84); the draw's one more-blocked case
  (a 35-cell polygon in an 80-cell channel) diverges, identically in both
  runs.
* Serving (Poisson-style mixed stream): 3D spheres / boxes up to
  256 x 128 x 128, FP16-shifted storage, a passive-scalar request and 2D
  cases, admitted against a device memory budget given through the
  free_memory hook that holds two of the large cases at once.  Every
  served result equals the same request served alone.
* A served 3D case retires at the step Solver3D.run(auto_stop=True) stops
  at, on the lift history the reference picks (solver3d.py:634-638).
MI355X only.
"""

import numpy as np
import pytest

from aero_amd import serve

pytestmark = pytest.mark.gpu


def _c4_pick():
    specs = serve.c4_cases(256, seed=0, max_steps=3000)
    big = [s for s in specs if s.nx == 1024]
    return big[:3] + [s for s in specs if s.nx != 1024][:29]


def test_c4_draw_batched_equals_alone_at_stated_sizes():
    from aero_amd.batch import Batch
    pick = _c4_pick()
    assert max(s.nx for s in pick) == 1024 and {s.collision for s in pick} == {"bgk", "trt", "mrt"}
    steps = 3000
    b = Batch()
    mixed = [s.build() for s in pick]
    rows_mix = b.advance(mixed, steps, record=False)
    blocked = []
    for s in mixed:
        blocked.append(s.D > 0.4 * s.Ny)
        s.close()
    b.close()
    for k, spec in enumerate(pick):
        s = spec.build()
        rows = s._advance(steps)
        s.close()
        np.testing.assert_array_equal(rows, rows_mix[k], err_msg=f"case {k}")
        cd = rows[:, 0]
        if not blocked[k]:
            assert np.all(np.isfinite(cd)), k
            cdm = float(np.mean(cd[-1000:]))
            assert 0.5 < cdm < 10.0, (k, cdm)
    assert sum(blocked) <= 2


def _c5_requests():
    reqs = []
    big = serve.CaseSpec(dim=3, nx=256, ny=128, nz=128, shape="sphere", size=(16.0,), Re=120.0, collision="trt",
                         storage="f16", max_steps=1500, seed=41)
    box = serve.CaseSpec(dim=3, nx=192, ny=96, nz=96, shape="box", size=(16.0, 16.0, 16.0), Re=80.0,
                         collision="mrt", storage="f16", thermal=True, max_steps=1500, seed=42)
    big2 = serve.CaseSpec(dim=3, nx=256, ny=128, nz=128, shape="box", size=(21.0, 21.0, 21.0), Re=200.0,
                          collision="bgk", storage="f16", max_steps=1000, seed=43)
    two_d = serve.c4_cases(4, seed=5, max_steps=2000, storage="f16")
    two_d[1].thermal = True
    specs = [big, two_d[0], box, two_d[1], big2, two_d[2], two_d[3]]
    arrivals = [0.0, 0.0, 0.01, 0.02, 0.02, 0.05, 0.08]
    return list(zip(arrivals, specs))


def test_c5_stated_sizes_memory_limited_equals_alone():
    reqs = _c5_requests()
    budget = 2 * max(s.device_bytes() for _, s in reqs) + (256 << 20)
    live_bytes = {"n": 0}
    # the hook reports the budget minus what the live cases hold
    srv = serve.Server(chunk=500, keep_history=True, mem_reserve=0,
                       free_memory=lambda: budget - live_bytes["n"])
    orig_build = serve.CaseSpec.build

    def build(spec):
        s = orig_build(spec)
        live_bytes["n"] += spec.device_bytes()
        close = s.close

        def _close():
            live_bytes["n"] -= spec.device_bytes()
            close()
        s.close = _close
        return s

    serve.CaseSpec.build = build
    try:
        mixed = srv.serve(reqs)
    finally:
        serve.CaseSpec.build = orig_build
    alone = serve.Server(chunk=500, max_live=1, keep_history=True).serve([(0.0, s) for _, s in reqs])
    assert len(mixed) == len(alone) == len(reqs)
    m = {r.rid: r for r in mixed}
    a = {r.rid: r for r in alone}
    for rid in m:
        assert m[rid].steps == a[rid].steps and m[rid].stop_reason == a[rid].stop_reason, rid
        np.testing.assert_array_equal(np.array(m[rid].Cd_history), np.array(a[rid].Cd_history), err_msg=str(rid))
        assert m[rid].warnings == a[rid].warnings, rid
    assert live_bytes["n"] == 0


def test_served_3d_case_stops_where_run_auto_stop_stops():
    kw = dict(convergence_window=1000, strouhal_window=512)
    spec = serve.CaseSpec(dim=3, nx=128, ny=64, nz=64, shape="sphere", size=(8.0,), Re=50.0, collision="trt",
                          max_steps=8000, seed=3)
    s = spec.build()
    r = s.run(steps=spec.max_steps, check_every=500, verbose=False, auto_stop=True, **kw)
    s.close()
    assert r["stop_reason"] == "auto_converged"
    res = serve.Server(chunk=500, keep_history=True, **kw).serve([(0.0, spec)])
    assert res[0].stop_reason == "converged" and res[0].steps == r["steps_completed"]
    np.testing.assert_array_equal(np.array(res[0].Cd_history), np.array(r["Cd_history"]))
