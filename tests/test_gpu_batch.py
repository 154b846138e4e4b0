"""Batched stepping: a case computes the same bits alone and in a mix.

Cases of different lattices, grid sizes, collisions, storage formats,
boundary sets and inlets are advanced together through aero_batch_step (one
launch per step for each group sharing a kernel) and compared with the same
cases stepped alone: populations, outlet state and every coefficient row
must be bit-identical.  MI355X only.
"""

import numpy as np
import pytest

from tests import _golden as G
from tests.test_gpu_parity import make_solver

pytestmark = pytest.mark.gpu

MIX = [(2, "bgk_bb_conv_slip"), (2, "trt_bouzidi_conv_slip"), (2, "mrt_bouzidi_zg_sponge"),
       (2, "bgk_thermal_buoy"), (2, "bgk_pressure_io"), (2, "bgk_synth_inflow"),
       (3, "trt_bouzidi_conv_slip"), (3, "bgk_sem"), (3, "mrt_bb_conv_sponge"), (3, "trt_recycling"),
       (3, "bgk_periodic_bodyforce")]


def _cylinder(nx, ny, r, collision, storage, seed):
    import aero_amd
    from aero_amd.geometry import Cylinder
    g = Cylinder(radius=r, cx_frac=0.3)
    solid = g.mark_solid(ny, nx)
    u0, D = 0.05, 2 * r
    omega = 1.0 / (3.0 * u0 * D / 60.0 + 0.5)
    s = aero_amd.Solver(Ny=ny, Nx=nx, solid=solid, omega=omega, u0=u0, D=D, collision=collision,
                        storage=storage, outlet_bc="zerogradient", sponge_thickness=8, inlet_perturbation=0.1)
    rng = np.random.default_rng(seed)
    from aero_amd.lattice import compute_feq
    f0 = compute_feq(np.ones((ny, nx)), np.full((ny, nx), u0), rng.uniform(-1e-4, 1e-4, (ny, nx)))
    s.f = f0
    return s


def _make_all(storage):
    out = []
    for dim, name in MIX:
        out.append(make_solver(dim, G.load(dim, name), storage))
    for k, (nx, ny, coll) in enumerate([(96, 48, "bgk"), (160, 80, "trt"), (128, 64, "mrt")]):
        out.append(_cylinder(nx, ny, 6.0, coll, storage, k))
    return out


@pytest.mark.parametrize("storage", ["f64", "f32", "f16"])
def test_batch_matches_alone_bit_for_bit(storage):
    from aero_amd.batch import Batch
    steps = 300  # crosses the 256-step (and 64-step plane) chunk boundaries
    alone = _make_all(storage)
    rows_alone = [s._advance(steps) for s in alone]
    mixed = _make_all(storage)
    b = Batch()
    rows_mix = b.advance(mixed, steps, record=False)
    for k, (a, m) in enumerate(zip(alone, mixed)):
        np.testing.assert_array_equal(rows_alone[k], rows_mix[k], err_msg=f"case {k}")
        np.testing.assert_array_equal(a.f, m.f, err_msg=f"case {k}")
        np.testing.assert_array_equal(a._f_outlet_prev, m._f_outlet_prev, err_msg=f"case {k}")
        if a.g is not None:
            np.testing.assert_array_equal(a.g, m.g, err_msg=f"case {k}")
        assert a.step_count == m.step_count
    b.close()


def test_batch_membership_changes_nothing():
    """The same case advanced in batches of changing composition (cases
    joining and leaving between calls) ends bit-identical to it alone."""
    from aero_amd.batch import Batch
    ref = _cylinder(160, 80, 6.0, "trt", "f32", 1)
    ref_rows = ref._advance(200)
    probe = _cylinder(160, 80, 6.0, "trt", "f32", 1)
    others = [_cylinder(96, 48, 5.0, "bgk", "f32", 7), _cylinder(128, 64, 6.0, "mrt", "f32", 8),
              make_solver(3, G.load(3, "trt_bouzidi_conv_slip"), "f32")]
    b = Batch()
    rows = [b.advance([probe], 50)[0], b.advance([others[0], probe, others[1]], 50)[1],
            b.advance([others[2], probe], 50)[1], b.advance(others + [probe], 50)[3]]
    np.testing.assert_array_equal(np.concatenate(rows), ref_rows)
    np.testing.assert_array_equal(probe.f, ref.f)
    b.close()


def test_batch_rejects_les_and_duplicates():
    from aero_amd.batch import Batch
    les = make_solver(2, G.load(2, "bgk_les_smag"), "f32")
    plain = make_solver(2, G.load(2, "bgk_bb_conv_slip"), "f32")
    b = Batch()
    with pytest.raises(RuntimeError, match="LES"):
        b.advance([plain, les], 2)
    with pytest.raises(RuntimeError, match="twice"):
        b.advance([plain, plain], 2)
    b.close()


def _host_stats(s):
    """check_stability's statistics from the materialised f64 field (the
    reference's own check reads self.f; 3D on the z mid-plane)."""
    f = np.asarray(s.f, dtype=np.float64)
    if f.ndim == 4:
        f = f[:, f.shape[1] // 2]
    from aero_amd.lattice import E, E3
    e = E if f.shape[0] == 9 else E3
    rho = f.sum(axis=0)
    mx = np.tensordot(e[:, 0].astype(np.float64), f, axes=1)
    my = np.tensordot(e[:, 1].astype(np.float64), f, axes=1)
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = np.where(rho > 0, 1.0 / rho, 0.0)
    with np.errstate(invalid="ignore", over="ignore"):
        um = np.sqrt((inv * mx) ** 2 + (inv * my) ** 2)
    um = np.where(np.isnan(um), 0.0, um)  # a NaN speed is flagged by the NaN test instead
    return np.array([float(np.isnan(f).any()), float((rho <= 0).any()), float(um.max()), float(rho.sum()),
                     float(rho.size)])


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_batched_stability_stats_match_each_case(storage):
    """aero_batch_check (one call for every live case) gives each case's
    aero_check statistics bit for bit, and both agree with the statistics of
    the materialised float64 field."""
    from aero_amd.batch import Batch
    cases = _make_all(storage)
    b = Batch()
    b.advance(cases, 40, record=False)
    many = b.check_stats(cases)
    for k, s in enumerate(cases):
        one = s._device_stats()
        np.testing.assert_array_equal(many[k], one, err_msg=f"case {k}")
        ref = _host_stats(s)
        assert many[k][0] == ref[0] and many[k][1] == ref[1] and many[k][4] == ref[4], (k, many[k], ref)
        np.testing.assert_allclose(many[k][2:4], ref[2:4], rtol=1e-12, err_msg=f"case {k}")
    b.close()
