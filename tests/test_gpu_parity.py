"""GPU parity: the HIP step path against the reference's own outputs.

Every golden case (tests/golden/, written from the reference's NumPy path)
is replayed through the product API -- Solver / Solver3D backed by
libaero_hip.so -- with float64 storage and compared at 1e-10 relative on the
full fields f, g, f_outlet_prev and on every per-step coefficient.  The same
cases in float32 storage are compared with the float64 oracle at a stated
float32 tolerance.  These run on an MI355X only.
"""

import numpy as np
import pytest

from tests import _golden as G

pytestmark = pytest.mark.gpu

F64_TOL = 1e-10      # relative, fields and coefficients (float64 storage)
F32_FIELD_TOL = 1e-5  # max |f - f_ref| / max |f_ref| after the last snapshot (float32 storage);
                      # float32 rounding grows fastest near tau = 1/2 (LES case: tau = 0.513)
F32_COEF_TOL = 2e-3   # max |C - C_ref| / max |C_ref| over the history (float32 storage)


def _aero():
    import aero_amd
    return aero_amd


def make_solver(dim, case, storage="f64"):
    A = _aero()
    kw = dict(case["kw"])
    use_phi = kw.get("bouzidi", False) or kw.get("van_driest", False) or kw.get("ibm_enabled", False)
    phi = case.get("phi") if use_phi else None
    solid = case["solid"]
    args = dict(omega=float(case["omega"]), u0=float(case["u0"]), D=float(case["D"]), phi=phi,
                storage=storage)
    if dim == 2:
        s = A.Solver(Ny=solid.shape[0], Nx=solid.shape[1], solid=solid, **args, **kw)
    else:
        s = A.Solver3D(Nz=solid.shape[0], Ny=solid.shape[1], Nx=solid.shape[2], solid=solid, **args, **kw)
    f0 = G.initial_f(case)
    s.f = f0
    s._f_outlet_prev = f0[..., -1].copy()
    return s


def replay(dim, name, storage="f64"):
    case = G.load(dim, name)
    s = make_solver(dim, case, storage)
    np.testing.assert_array_equal(s.surface_links, case["links"])
    np.testing.assert_array_equal(s.q_vals, case["q_vals"])
    snaps = G.SNAP2 if dim == 2 else G.SNAP3
    rows = []
    fields = {}
    done = 0
    for snap in snaps:
        rows.append(s._advance(snap - done))
        done = snap
        fields[snap] = (s.f, s._f_outlet_prev, s.g)
    return case, s, np.concatenate(rows), fields


def _cols(dim):
    if dim == 2:
        return dict(Cd=0, Cl=1, Cm=5, Cd_p=6, Cd_v=7)
    return dict(Cd=0, Cly=1, Clz=2, Cmx=3, Cmy=4, Cmz=5, Cd_p=6, Cd_v=7)


@pytest.mark.parametrize("dim,name", [(2, n) for n in G.CASES_2D] + [(3, n) for n in G.CASES_3D])
def test_f64_matches_reference(dim, name):
    print(f"[gpu] f64 parity {dim}d {name}", flush=True)
    case, s, rows, fields = replay(dim, name)
    for snap, (f, prev, g) in fields.items():
        err = G.rel_err(f, case[f"f_{snap}"])
        assert err < F64_TOL, (name, snap, err)
        assert G.rel_err(prev, case[f"fprev_{snap}"]) < F64_TOL, (name, snap)
        if g is not None:
            assert G.rel_err(g, case[f"g_{snap}"]) < F64_TOL, (name, snap)
    for key, col in _cols(dim).items():
        ref = case[f"{key}_hist"]
        scale = max(np.max(np.abs(ref)), 1e-3)
        assert np.max(np.abs(rows[:, col] - ref)) / scale < F64_TOL * 10, (name, key)
    prof = s._last_force_profile if dim == 2 else s._last_spanwise_profile
    for k in ("fx", "fy") + (("fz",) if dim == 3 else ()):
        np.testing.assert_allclose(prof[k], case[f"profile_{k}"], rtol=1e-9, atol=1e-12)


F32_CASES = [(2, "bgk_bb_conv_slip"), (2, "trt_bouzidi_conv_slip"), (2, "mrt_bouzidi_zg_sponge"),
             (2, "bgk_thermal_buoy"), (2, "bgk_les_smag"), (3, "trt_bouzidi_conv_slip"),
             (3, "mrt_bouzidi_zg_noslip"), (3, "bgk_thermal_buoy"), (3, "trt_recycling")]


@pytest.mark.parametrize("dim,name", F32_CASES)
def test_f32_matches_reference_within_tolerance(dim, name):
    print(f"[gpu] f32 parity {dim}d {name}", flush=True)
    case, s, rows, fields = replay(dim, name, storage="f32")
    last = max(fields)
    f, _, g = fields[last]
    err = G.rel_err(f, case[f"f_{last}"])
    assert err < F32_FIELD_TOL, (name, err)
    if g is not None:
        err_g = G.rel_err(g, case[f"g_{last}"])
        assert err_g < F32_FIELD_TOL, (name, err_g)
    print(f"[gpu] f32 {dim}d {name}: field rel err {err:.2e}", flush=True)
    ref = case["Cd_hist"]
    assert np.max(np.abs(rows[:, 0] - ref)) / np.max(np.abs(ref)) < F32_COEF_TOL, name


def test_set_get_roundtrip_and_checkpoint(tmp_path):
    """f written through the ABI reads back bit-exact; a checkpoint restart
    continues the run (tests/test_tier3.py:148-163 of the reference)."""
    A = _aero()
    case = G.load(2, "bgk_bouzidi_zg_sponge")
    s = make_solver(2, case)
    np.testing.assert_array_equal(s.f, G.initial_f(case))
    s.run(steps=20, check_every=10, verbose=False)
    path = str(tmp_path / "c.npz")
    s.save_checkpoint(path)
    s.run(steps=20, check_every=10, verbose=False)
    r = A.Solver.from_checkpoint(path)
    assert r.step_count == 20
    r.run(steps=20, check_every=10, verbose=False)
    np.testing.assert_allclose(r.f, s.f, rtol=0, atol=1e-13)
    assert r.Cd_history[-1] == pytest.approx(s.Cd_history[-1], rel=1e-10)


@pytest.mark.parametrize("dim,name", [(3, "bgk_sem"), (2, "bgk_synth_inflow"), (3, "bgk_synth_inflow")])
def test_inflow_checkpoint_resume_continues_the_same_run(tmp_path, dim, name):
    """A case with a stateful inlet checkpointed mid-run and rebuilt with
    from_checkpoint continues exactly as the uninterrupted run: the checkpoint
    carries the SEM eddies and random state (sem_state_json) or the synthetic
    inflow's filtered field and random state (inflow_state_json) besides the
    reference's keys."""
    A = _aero()
    case = G.load(dim, name)
    s = make_solver(dim, case)
    s.run(steps=12, check_every=6, verbose=False)
    path = str(tmp_path / "c.npz")
    s.save_checkpoint(path)
    s.run(steps=18, check_every=6, verbose=False)
    r = (A.Solver3D if dim == 3 else A.Solver).from_checkpoint(path)
    if name == "bgk_sem":
        assert r.sem_inlet and r._sem_device
    r.run(steps=18, check_every=6, verbose=False)
    np.testing.assert_allclose(r.f, s.f, rtol=0, atol=1e-13)
    assert r.Cd_history[-1] == pytest.approx(s.Cd_history[-1], rel=1e-10)


def test_divergence_raises_runtime_error():
    """check_stability semantics (aero/diagnostics.py:43-52) on device statistics."""
    case = G.load(2, "bgk_bb_conv_slip")
    s = make_solver(2, case)
    f = s.f
    f[3, 5, 7] = np.nan
    s.f = f
    with pytest.raises(RuntimeError, match="NaN"):
        s.run(steps=2, check_every=1, verbose=False)


def _sphere_case(storage, steps, variant_kw, nz=12, ny=14, nx=28, radius=3.4):
    """Small D3Q19 TRT sphere; product and float64 oracle from the same state."""
    import aero_amd
    from aero_amd.geometry import Sphere
    from oracle import oracle as orc
    g = Sphere(radius=radius, cx_frac=0.3)
    solid, phi = g.mark_solid(nz, ny, nx), g.sdf_field(nz, ny, nx, offset=0.0)
    u0, D = 0.05, 2 * radius
    omega = 1.0 / (3.0 * u0 * D / 40.0 + 0.5)
    s = aero_amd.Solver3D(Nz=nz, Ny=ny, Nx=nx, solid=solid, omega=omega, u0=u0, D=D, phi=phi, bouzidi=True,
                          collision="trt", trt_lambda=0.1875, storage=storage, **variant_kw)
    o = orc.OracleSolver(solid, omega=omega, u0=u0, D=D, phi=phi, bouzidi=True, collision="trt",
                         trt_lambda=0.1875, **variant_kw)
    rng = np.random.default_rng(7)
    sh = (nz, ny, nx)
    f0 = orc.feq(orc.E19, orc.W19, np.ones(sh), np.full(sh, u0), rng.uniform(-1e-4, 1e-4, sh),
                 rng.uniform(-1e-4, 1e-4, sh))
    s.f = f0
    s._f_outlet_prev = f0[..., -1].copy()
    o.set_state(f0)
    np.testing.assert_array_equal(s.q_vals, o.q_vals)
    rows = s._advance(steps)
    cd = np.array([o.coefficients(o.step())[0] for _ in range(steps)])
    return s, o, rows, cd


def test_lambda_trt_linear_bouzidi_f64_matches_oracle():
    """The labelled non-reference operators (standard Lambda-TRT, linear
    Bouzidi with link fractions) against the float64 oracle's restatement."""
    s, o, rows, cd = _sphere_case("f64", 30, dict(trt_variant="lambda", bouzidi_variant="linear"))
    assert G.rel_err(s.f, o.f_ref()) < F64_TOL
    assert np.max(np.abs(rows[:, 0] - cd)) / np.max(np.abs(cd)) < F64_TOL * 10


HALF_FIELD_TOL = {"f16": 2e-4, "bf16": 2e-3}  # max |f - f64| / max |f| after 60 steps
HALF_CD_TOL = {"f16": 5e-3, "bf16": 5e-2}     # max |Cd - Cd_f64| / max |Cd_f64| over the run


@pytest.mark.parametrize("storage", ["f16", "bf16"])
def test_shifted_half_storage_accuracy(storage):
    """Shifted half-precision storage (f - w_i), float32 arithmetic, against
    the float64 oracle: stated field and drag tolerances."""
    s, o, rows, cd = _sphere_case(storage, 60, dict(trt_variant="lambda", bouzidi_variant="linear"))
    err = G.rel_err(s.f, o.f_ref())
    cd_err = float(np.max(np.abs(rows[:, 0] - cd)) / np.max(np.abs(cd)))
    print(f"[gpu] {storage}: field rel err {err:.2e}, Cd rel err {cd_err:.2e}", flush=True)
    assert err < HALF_FIELD_TOL[storage]
    assert cd_err < HALF_CD_TOL[storage]


@pytest.mark.parametrize("shape", [(4, 6, 96), (3, 3, 66), (5, 65)])
@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
def test_convective_outlet_with_one_outlet_cell_per_wave(storage, shape):
    """Rows wider than a 64-lane wavefront put exactly one outlet cell in
    each wave holding one, at varying lane positions; there the 2-byte
    storage forms fetch the row's previous outlet state with one load spread
    over the wave's lanes 0..Q-1.  The 3 x 3 x 66 and 5 x 65 grids end in a
    partial wave with fewer than Q valid lanes (594 = 9 * 64 + 18 cells,
    325 = 5 * 64 + 5), where that fetch must not be used.  The state and the
    outlet state vary smoothly by 1-2% from row to row (tau = 1: the run is
    insensitive to rounding, 2e-8 after 20 steps from a float32-rounded start
    on the 4 x 6 x 96 grid), while the outlet state of the neighbouring row in
    place of the right one moves that result by 6e-3 of max |f|: far above
    the stated tolerances.  Float64 oracle from the same state."""
    import aero_amd
    from oracle import oracle as orc
    u0 = 0.05
    solid = np.zeros(shape, dtype=bool)
    if len(shape) == 3:
        nz, ny, nx = shape
        z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        rho = 1.0 + 0.01 * np.sin(2 * np.pi * (y / ny + 2 * z / nz))
        f0 = orc.feq(orc.E19, orc.W19, rho, u0 + 0.02 * np.cos(2 * np.pi * (y / ny + z / nz)),
                     0.01 * np.sin(2 * np.pi * y / ny), 0.01 * np.cos(2 * np.pi * z / nz))
        prev = np.roll(np.roll(f0[..., -1], 1, axis=2), 1, axis=1)  # valid, and not each row's own
        s = aero_amd.Solver3D(Nz=nz, Ny=ny, Nx=nx, solid=solid, omega=1.0, u0=u0, D=8.0, collision="trt",
                              trt_lambda=0.1875, outlet_bc="convective", storage=storage)
    else:
        ny, nx = shape
        y, x = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
        rho = 1.0 + 0.01 * np.sin(2 * np.pi * y / ny)
        f0 = orc.feq(orc.E9, orc.W9, rho, u0 + 0.02 * np.cos(2 * np.pi * y / ny), 0.01 * np.sin(4 * np.pi * y / ny))
        prev = np.roll(f0[..., -1], 1, axis=1)
        s = aero_amd.Solver(Ny=ny, Nx=nx, solid=solid, omega=1.0, u0=u0, D=8.0, collision="trt",
                            trt_lambda=0.1875, outlet_bc="convective", storage=storage)
    o = orc.OracleSolver(solid, omega=1.0, u0=u0, D=8.0, collision="trt", trt_lambda=0.1875)
    s.f = f0
    s._f_outlet_prev = prev
    o.set_state(f0, prev=prev)
    s._advance(20)
    for _ in range(20):
        o.step()
    tol = {"f32": F32_FIELD_TOL, "f16": HALF_FIELD_TOL["f16"], "bf16": HALF_FIELD_TOL["bf16"]}[storage]
    err_f, err_p = G.rel_err(s.f, o.f_ref()), G.rel_err(s._f_outlet_prev, o.prev_ref())
    print(f"[gpu] outlet rows {shape} {storage}: field {err_f:.2e} outlet state {err_p:.2e}", flush=True)
    assert err_f < tol and err_p < tol
    s.close()


@pytest.mark.parametrize("dim,name,total,splits", [
    (2, "bgk_thermal_buoy", 600, (1, 255, 344)),   # crosses 256-step chunks
    (3, "bgk_sem", 150, (1, 63, 86)),              # per-step inlet planes, 64-step chunks
    (3, "trt_bouzidi_conv_slip", 300, (300,)),
])
def test_chunked_stepping_is_invariant(dim, name, total, splits):
    """aero_step works in fixed-size chunks; one call of `total` steps and the
    same steps split over several calls give bit-identical state, outlet state
    and coefficient rows."""
    case = G.load(dim, name)
    a = make_solver(dim, case)
    ra = a._advance(total)
    b = make_solver(dim, case)
    rb = np.concatenate([b._advance(n) for n in splits] if sum(splits) == total else [b._advance(total)])
    np.testing.assert_array_equal(ra, rb)
    np.testing.assert_array_equal(a.f, b.f)
    np.testing.assert_array_equal(a._f_outlet_prev, b._f_outlet_prev)
    if dim == 3 and splits == (300,):
        # 300 steps in one call vs 300 single-step calls on the first 40 rows
        c = make_solver(dim, case)
        rc = np.concatenate([c._advance(1) for _ in range(40)])
        np.testing.assert_array_equal(ra[:40], rc)
