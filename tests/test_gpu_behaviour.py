"""What scripts written for the reference solver rely on, on the device path.

Each test asserts what the reference's own solver-level test asserts
(cited per test), on the same small case and through the same public API:
the constructor keywords (``backend="numpy"`` included), ``run()`` result
keys, ``_step()``, in-place edits of ``f``, histories and checkpoints."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _macro2d(f):
    from aero_amd.lattice import compute_macroscopic
    return compute_macroscopic(np.asarray(f))


def _cylinder_case(Ny=60, Nx=120, radius=None, Re=100.0, u0=0.05, **kw):
    """The case builder of tests/test_solver.py:20-34: nu from u0 D / Re with
    D the diameter, or nu = 0.02 without an obstacle."""
    import aero_amd
    from aero_amd.geometry import Cylinder
    if radius is None:
        solid, D = np.zeros((Ny, Nx), dtype=bool), 1.0
    else:
        solid, D = Cylinder(radius=radius, cx_frac=1 / 3, cy_frac=0.5).mark_solid(Ny, Nx), 2.0 * radius
    nu = u0 * D / Re if D > 1.0 else 0.02
    return aero_amd.Solver(Ny=Ny, Nx=Nx, solid=solid, omega=1.0 / (3.0 * nu + 0.5), u0=u0, D=D, **kw)


def test_mass_is_conserved_without_an_obstacle():
    """tests/test_solver.py:41-57: total mass within 0.1% over 200 steps."""
    s = _cylinder_case()
    m0 = _macro2d(s.f)[0].sum()
    s.run(steps=200, check_every=200, verbose=False)
    m1 = _macro2d(s.f)[0].sum()
    assert abs(m1 - m0) / m0 < 1e-3


def test_uniform_flow_develops_without_an_obstacle():
    """tests/test_solver.py:63-75: interior ux within 20% of u0 after 300 steps."""
    s = _cylinder_case(Ny=40, Nx=80, u0=0.05)
    s.run(steps=300, check_every=300, verbose=False)
    ux = _macro2d(s.f)[1]
    assert ux[1:-1, 5:-5].mean() == pytest.approx(0.05, rel=0.20)


def test_no_nan_after_steps():
    """tests/test_solver.py:81-86."""
    s = _cylinder_case(Ny=40, Nx=80, radius=8, Re=100)
    s.run(steps=100, check_every=100, verbose=False)
    assert not np.isnan(s.f).any()


def test_cylinder_drag_in_range_and_lift_near_zero():
    """tests/test_solver.py:92-115: Cd_mean in [0.8, 3.0] and |Cl_mean| < 0.5
    after 3000 steps at Re 100 (radius 20 on 100 x 200)."""
    s = _cylinder_case(Ny=100, Nx=200, radius=20, Re=100, u0=0.05)
    r = s.run(steps=3000, check_every=3000, verbose=False)
    print(f"[gpu] cylinder Cd_mean {r['Cd_mean']:.4f} Cl_mean {r['Cl_mean']:.4f}", flush=True)
    assert 0.8 <= r["Cd_mean"] <= 3.0
    assert abs(r["Cl_mean"]) < 0.5


def test_sponge_damps_an_outlet_disturbance_set_in_place():
    """tests/test_sponge.py:28-47: a density bump added in place to f near
    the outlet (``s.f[:, y, x] += 0.05``) is damped more with the sponge."""
    import aero_amd
    ny, nx = 24, 120
    common = dict(Ny=ny, Nx=nx, solid=np.zeros((ny, nx), dtype=bool), omega=1.0, u0=0.05, D=1.0, backend="numpy")
    plain = aero_amd.Solver(**common, sponge_thickness=0)
    sponge = aero_amd.Solver(**common, sponge_thickness=8, sponge_strength=0.3)
    for s in (plain, sponge):
        before = s.f[:, ny // 2, -10].copy()
        s.f[:, ny // 2, -10] += 0.05
        np.testing.assert_allclose(s.f[:, ny // 2, -10], before + 0.05, rtol=0, atol=1e-12)
    for _ in range(30):
        plain._step()
        sponge._step()
    assert float(np.std(sponge.f[0, :, -5:])) < float(np.std(plain.f[0, :, -5:]))


def test_trt_runs_and_its_drag_split_sums_to_the_total():
    """tests/test_trt.py:9-21 (2D) and 23-36 (3D)."""
    import aero_amd
    from aero_amd.geometry import Cylinder, Sphere
    solid = Cylinder(radius=6, cx_frac=1 / 3, cy_frac=0.5).mark_solid(40, 80)
    r = aero_amd.Solver(Ny=40, Nx=80, solid=solid, omega=1.2, u0=0.05, D=12.0, backend="numpy",
                        collision="trt").run(steps=200, check_every=200, verbose=False)
    assert not np.isnan(r["Cd_mean"]) and not np.isnan(r["Cd_p_mean"])
    assert abs(r["Cd_p_mean"] + r["Cd_v_mean"] - r["Cd_mean"]) < 1e-10
    solid3 = Sphere(radius=4, cx_frac=0.25, cy_frac=0.5, cz_frac=0.5).mark_solid(12, 16, 32)
    r3 = aero_amd.Solver3D(Nz=12, Ny=16, Nx=32, solid=solid3, omega=1.1, u0=0.05, D=8.0, backend="numpy",
                           collision="trt").run(steps=30, check_every=30, verbose=False)
    assert not np.isnan(r3["Cd_mean"])


def test_drag_split_sums_to_the_total_and_both_parts_are_nonzero():
    """tests/test_stress.py:24-50."""
    import aero_amd
    from aero_amd.geometry import Cylinder
    solid = Cylinder(radius=8, cx_frac=1 / 3, cy_frac=0.5).mark_solid(60, 120)
    r = aero_amd.Solver(Ny=60, Nx=120, solid=solid, omega=1.0, u0=0.05, D=16.0, backend="numpy",
                        collision="bgk").run(steps=200, check_every=200, verbose=False)
    assert abs(r["Cd_p_mean"] + r["Cd_v_mean"] - r["Cd_mean"]) / max(abs(r["Cd_mean"]), 1e-6) < 0.05
    r2 = aero_amd.Solver(Ny=60, Nx=120, solid=solid, omega=0.8, u0=0.05, D=16.0, backend="numpy",
                         collision="bgk").run(steps=300, check_every=300, verbose=False)
    assert abs(r2["Cd_p_mean"]) > 0.01 and abs(r2["Cd_v_mean"]) > 0.01


def test_les_runs_with_both_models():
    """tests/test_les.py:24-33 and 57-65."""
    import aero_amd
    from aero_amd.geometry import Cylinder
    solid = Cylinder(radius=6, cx_frac=1 / 3, cy_frac=0.5).mark_solid(40, 80)
    r = aero_amd.Solver(Ny=40, Nx=80, solid=solid, omega=1.0, u0=0.05, D=12.0, backend="numpy", les=True,
                        les_cs=0.16).run(steps=30, check_every=30, verbose=False)
    assert not np.isnan(r["Cd_mean"])
    r2 = aero_amd.Solver(Ny=24, Nx=48, solid=np.zeros((24, 48), dtype=bool), omega=1.0, u0=0.05, D=12.0,
                         backend="numpy", les=True, les_model="wale").run(steps=10, check_every=10, verbose=False)
    assert not np.isnan(r2["Cd_mean"])


def test_streamwise_periodic_and_recycling_modes_run():
    """tests/test_streamwise_bc3d.py:26-61."""
    import aero_amd
    from aero_amd.geometry import Sphere
    s = aero_amd.Solver3D(Nz=8, Ny=12, Nx=20, solid=np.zeros((8, 12, 20), dtype=bool), omega=1.0, u0=0.05,
                          D=10.0, backend="numpy", wall_bc="noslip", streamwise_bc="periodic", body_force_x=1.0e-5)
    r = s.run(steps=40, check_every=20, verbose=False)
    assert len(r["Cd_history"]) == 40 and r["steps_completed"] == 40
    solid = Sphere(radius=3, cx_frac=0.35, cy_frac=0.5, cz_frac=0.5).mark_solid(10, 12, 24)
    r2 = aero_amd.Solver3D(Nz=10, Ny=12, Nx=24, solid=solid, omega=1.1, u0=0.05, D=6.0, backend="numpy",
                           streamwise_bc="recycling").run(steps=20, check_every=10, verbose=False)
    assert len(r2["Cd_history"]) == 20 and not np.isnan(r2["Cd_mean"])


def test_inlet_perturbation_raises_the_lift_variance():
    """tests/test_inlet3d.py:9-21."""
    import aero_amd
    from aero_amd.geometry import Sphere
    solid = Sphere(radius=4, cx_frac=0.25, cy_frac=0.5, cz_frac=0.5).mark_solid(12, 16, 32)
    common = dict(Nz=12, Ny=16, Nx=32, solid=solid, omega=1.0, u0=0.05, D=8.0, backend="numpy")
    r0 = aero_amd.Solver3D(**common, inlet_perturbation=0.0).run(steps=80, check_every=80, verbose=False)
    r1 = aero_amd.Solver3D(**common, inlet_perturbation=0.02).run(steps=80, check_every=80, verbose=False)
    var = [float(np.var(r["Cly_history"][-40:])) + float(np.var(r["Clz_history"][-40:])) for r in (r0, r1)]
    assert var[1] > var[0]


def test_observables_and_synthetic_inflow_in_the_run_result():
    """tests/test_wall_and_inflow.py:34-70."""
    import aero_amd
    r = aero_amd.Solver(Ny=20, Nx=40, solid=np.zeros((20, 40), dtype=bool), omega=1.0, u0=0.05, D=10.0,
                        backend="numpy", synthetic_inflow=True, synthetic_inflow_intensity=0.02).run(
        steps=40, check_every=20, verbose=False)
    assert "force_profile_last" in r["observables"] and len(r["Cm_history"]) == 40
    r3 = aero_amd.Solver3D(Nz=6, Ny=10, Nx=20, solid=np.zeros((6, 10, 20), dtype=bool), omega=1.0, u0=0.05,
                           D=10.0, backend="numpy", synthetic_inflow=True, synthetic_inflow_intensity=0.02).run(
        steps=20, check_every=10, verbose=False)
    assert "spanwise_force_profile_last" in r3["observables"] and len(r3["Cmx_history"]) == 20


def test_thermal_runs_bounded_and_buoyancy_drives_vertical_flow():
    """tests/test_thermal.py:105-145: T stays in [-0.1, 1.1] over 10 steps;
    with buoyancy and pressure inlet / outlet, uy becomes nonzero."""
    import aero_amd
    s = aero_amd.Solver(Ny=24, Nx=32, solid=np.zeros((24, 32), dtype=bool), omega=1.5, u0=0.0, D=10.0,
                        backend="numpy", thermal=True, T_hot=1.0, T_cold=0.0, alpha_T=0.1)
    for _ in range(10):
        s._step()
    T = np.asarray(s.g).sum(axis=0)
    assert T.min() >= -0.1 and T.max() <= 1.1
    b = aero_amd.Solver(Ny=24, Nx=24, solid=np.zeros((24, 24), dtype=bool), omega=1.5, u0=0.0, D=10.0,
                        backend="numpy", inlet_bc="pressure", outlet_bc="pressure", thermal=True, T_hot=1.0,
                        T_cold=0.0, alpha_T=0.01, buoyancy=True, g_gravity=1e-4, beta=0.1, T_ref=0.5)
    for _ in range(200):
        b._step()
    assert np.abs(_macro2d(b.f)[2]).max() > 1e-10


@pytest.mark.parametrize("dim", [2, 3])
def test_scalar_outputs_and_checkpoint_round_trip(tmp_path, dim):
    """tests/test_scalar_transport.py:33-88."""
    import aero_amd
    if dim == 2:
        shape = (16, 24)
        s = aero_amd.Solver(Ny=16, Nx=24, solid=np.zeros(shape, dtype=bool), omega=1.5, u0=0.01, D=8.0,
                            backend="numpy", thermal=True, T_hot=1.0, T_cold=0.0, alpha_T=0.05)
        r = s.run(steps=8, check_every=100, verbose=False)
    else:
        shape = (6, 8, 12)
        s = aero_amd.Solver3D(Nz=6, Ny=8, Nx=12, solid=np.zeros(shape, dtype=bool), omega=1.4, u0=0.01, D=6.0,
                              backend="numpy", thermal=True, T_hot=1.0, T_cold=0.0, alpha_T=0.05)
        r = s.run(steps=4, check_every=100, verbose=False)
    assert r["scalar"] is not None and r["scalar_stats"]["enabled"] is True and r["scalar"].shape == shape
    path = str(tmp_path / "scalar.npz")
    s.save_checkpoint(path)
    loaded = (aero_amd.Solver if dim == 2 else aero_amd.Solver3D).from_checkpoint(path)
    assert loaded.thermal is True and loaded.g is not None and loaded.alpha_T == 0.05


def _small(**kw):
    """The case of tests/test_tier3.py:17-22."""
    import aero_amd
    from aero_amd.geometry import Cylinder
    args = dict(Ny=30, Nx=60, solid=Cylinder(radius=6, cx_frac=1 / 3, cy_frac=0.5).mark_solid(30, 60), omega=1.2,
                u0=0.05, D=12.0, backend="numpy")
    args.update(kw)
    return aero_amd.Solver(**args)


def test_hdf5_path_is_a_no_op_without_h5py(tmp_path):
    """tests/test_tier3.py:91-100: passing hdf5_path without h5py must not
    crash (this build writes no HDF5, as the reference without h5py)."""
    import importlib.util
    import warnings
    if importlib.util.find_spec("h5py") is not None:
        pytest.skip("h5py installed")
    s = _small()
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", RuntimeWarning)
        s.run(steps=50, check_every=50, verbose=False, hdf5_path=str(tmp_path / "out.h5"), hdf5_every=25)
    assert s.step_count == 50


def test_checkpoint_reproduces_the_state_and_continues_the_run(tmp_path):
    """tests/test_tier3.py:130-176: from_checkpoint rebuilds the same state
    and history; a restart continues as the uninterrupted run; the file holds
    solid and params_json."""
    import aero_amd
    s = _small()
    s.run(steps=100, check_every=100, verbose=False)
    path = str(tmp_path / "ckpt.npz")
    s.save_checkpoint(path)
    with np.load(path, allow_pickle=False) as data:
        assert "solid" in data and "params_json" in data
    r = aero_amd.Solver.from_checkpoint(path)
    assert (r.step_count, r.Ny, r.Nx) == (s.step_count, s.Ny, s.Nx)
    np.testing.assert_allclose(r.f, s.f, rtol=0, atol=1e-14)
    assert r.Cd_history[-1] == pytest.approx(s.Cd_history[-1], rel=1e-12)
    full = _small().run(steps=200, check_every=200, verbose=False)
    r.run(steps=100, check_every=100, verbose=False)
    assert r.Cd_history[-1] == pytest.approx(full["Cd_history"][-1], rel=1e-10)
