"""Launch tuning never changes a result (MI355X only).

The same mix of cases -- 2D BGK / TRT / MRT cylinders in FP32, FP16 and
BF16 and a 3D sphere -- stepped with the built-in schedule and with every
knob of the tuning table moved (natural tile order, other XCD run lengths,
non-temporal stores forced, chunks replayed from captured HIP graphs, one
row per wave instead of the grouped bulk rows of the 2-byte D2Q9 kernels),
and after the autotuner has picked its own entries: coefficient rows and
final states must be bit-identical.
"""

import numpy as np
import pytest

from tests import _golden as G
from tests.test_gpu_batch import _cylinder
from tests.test_gpu_parity import make_solver

pytestmark = pytest.mark.gpu


def _walled(wall_bc, collision, seed):
    """A 2-byte 2D cylinder between no-slip or moving walls, pressure inlet
    and outlet (the grouped rows' wall and pressure-outlet handling)."""
    import aero_amd
    from aero_amd.geometry import Cylinder
    from aero_amd.lattice import compute_feq
    ny, nx = 72, 320
    solid = Cylinder(radius=7.0, cx_frac=0.3).mark_solid(ny, nx)
    s = aero_amd.Solver(Ny=ny, Nx=nx, solid=solid, omega=1.6, u0=0.05, D=14.0, collision=collision, storage="f16",
                        wall_bc=wall_bc, wall_velocity_top=0.03, wall_velocity_bottom=-0.01,
                        inlet_bc="pressure", outlet_bc="pressure", rho_in=1.002, rho_out=1.0)
    rng = np.random.default_rng(seed)
    s.f = compute_feq(np.ones((ny, nx)), np.full((ny, nx), 0.05), rng.uniform(-1e-4, 1e-4, (ny, nx)))
    return s


def _mix():
    return [_cylinder(160, 80, 6.0, "bgk", "f32", 1), _cylinder(96, 48, 5.0, "trt", "f16", 2),
            _cylinder(128, 64, 6.0, "mrt", "f32", 3), make_solver(3, G.load(3, "trt_bouzidi_conv_slip"), "f32"),
            _cylinder(640, 96, 8.0, "bgk", "f16", 4), _cylinder(512, 128, 10.0, "trt", "bf16", 5),
            _cylinder(384, 96, 8.0, "mrt", "f16", 6), _walled("moving", "trt", 7), _walled("noslip", "mrt", 8)]



def _run(solvers, steps=300):
    from aero_amd.batch import Batch
    b = Batch()
    rows = b.advance(solvers, steps, record=False)
    b.close()
    return rows, [s.f for s in solvers]


def test_tuned_schedules_are_bit_identical():
    from aero_amd import autotune
    saved = autotune.table()
    try:
        autotune.clear()
        ref_rows, ref_f = _run(_mix())
        mix = _mix()
        for key in autotune.group_keys(mix):
            autotune.set_entry(key, {"xcd_run": -1, "nt": 1, "graph": 1, "rowgroups": 0})
        rows, fs = _run(mix)
        for a, b in zip(ref_rows, rows):
            np.testing.assert_array_equal(a, b)
        for a, b in zip(ref_f, fs):
            np.testing.assert_array_equal(a, b)
        # single-case stepping under the same table (a batch of one)
        for k in (0, 4, 6, 7):
            s, t = _mix()[k], _mix()[k]
            autotune.clear()
            r0 = s._advance(300)
            for key in autotune.group_keys([t]):
                autotune.set_entry(key, {"xcd_run": 8, "nt": 0, "graph": 1, "rowgroups": 0})
            np.testing.assert_array_equal(r0, t._advance(300))
            np.testing.assert_array_equal(s.f, t.f)
    finally:
        autotune.clear()
        for k, v in saved.items():
            autotune.set_entry(k, v)


def test_autotuner_picks_entries_and_keeps_results(tmp_path):
    from aero_amd import autotune
    saved = autotune.table()
    try:
        autotune.clear()
        ref_rows, ref_f = _run(_mix())
        work = _mix()
        picked = autotune.tune(work, steps=20, reps=1)
        assert set(picked) == set(autotune.group_keys(work))
        path = tmp_path / "t.json"
        autotune.save(path, picked)
        autotune.clear()
        assert autotune.load(path) == len(picked)
        rows, fs = _run(_mix())
        for a, b in zip(ref_rows, rows):
            np.testing.assert_array_equal(a, b)
        for a, b in zip(ref_f, fs):
            np.testing.assert_array_equal(a, b)
    finally:
        autotune.clear()
        for k, v in saved.items():
            autotune.set_entry(k, v)
