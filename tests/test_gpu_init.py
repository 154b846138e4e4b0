"""Device-side equilibrium initialisation (aero_set_f_equilibrium).

The device evaluates compute_feq / compute_feq_3d (aero/lbm/d2q9.py:66-89,
aero/lbm/d3q19.py:105-129) in the host formula's order of operations, so
the state it sets is bit-identical to uploading the host field.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _solver2d(storage="f64"):
    import aero_amd
    Ny, Nx = 24, 40
    solid = np.zeros((Ny, Nx), dtype=bool)
    solid[10:14, 8:12] = True
    return aero_amd.Solver(Ny=Ny, Nx=Nx, solid=solid, omega=1.2, u0=0.05, D=4.0, storage=storage)


def _solver3d(storage="f64"):
    import aero_amd
    Nz, Ny, Nx = 8, 12, 20
    solid = np.zeros((Nz, Ny, Nx), dtype=bool)
    solid[3:5, 5:7, 6:8] = True
    return aero_amd.Solver3D(Nz=Nz, Ny=Ny, Nx=Nx, solid=solid, omega=1.2, u0=0.05, D=2.0, storage=storage)


def test_constructor_state_is_the_host_equilibrium_bitwise():
    from aero_amd.lattice import compute_feq, compute_feq_3d
    s = _solver2d()
    sh = (s.Ny, s.Nx)
    print(f"[gpu] init 2d {sh}", flush=True)
    assert np.array_equal(s.f, compute_feq(np.full(sh, 1.0), np.full(sh, 0.05), np.zeros(sh)))
    assert np.array_equal(s._f_outlet_prev, s.f[:, :, -1])
    t = _solver3d()
    sh3 = (t.Nz, t.Ny, t.Nx)
    print(f"[gpu] init 3d {sh3}", flush=True)
    assert np.array_equal(t.f, compute_feq_3d(np.full(sh3, 1.0), np.full(sh3, 0.05), np.zeros(sh3), np.zeros(sh3)))
    assert np.array_equal(t._f_outlet_prev, t.f[..., -1])


@pytest.mark.parametrize("dim", [2, 3])
def test_device_equilibrium_of_fields_matches_host_bitwise(dim):
    from aero_amd.lattice import compute_feq, compute_feq_3d
    rng = np.random.default_rng(11 + dim)
    s = _solver2d() if dim == 2 else _solver3d()
    sh = s._ref_shape()[1:]
    rho = 1.0 + rng.uniform(-0.02, 0.02, sh)
    ux, uy = rng.uniform(-0.1, 0.1, sh), rng.uniform(-0.1, 0.1, sh)
    print(f"[gpu] set_equilibrium dim {dim} {sh}", flush=True)
    if dim == 2:
        s.set_equilibrium(rho, ux, uy)
        assert np.array_equal(s.f, compute_feq(rho, ux, uy))
        s.set_equilibrium(uy=uy)  # rho0 = 1, ux = u0
        assert np.array_equal(s.f, compute_feq(np.ones(sh), np.full(sh, 0.05), uy))
    else:
        uz = rng.uniform(-0.1, 0.1, sh)
        s.set_equilibrium(rho, ux, uy, uz)
        assert np.array_equal(s.f, compute_feq_3d(rho, ux, uy, uz))
        s.set_equilibrium(uz=uz)
        assert np.array_equal(s.f, compute_feq_3d(np.ones(sh), np.full(sh, 0.05), np.zeros(sh), uz))


def test_device_equilibrium_steps_like_the_uploaded_field():
    from aero_amd.lattice import compute_feq
    rng = np.random.default_rng(5)
    a, b = _solver2d("f32"), _solver2d("f32")
    uy = rng.uniform(-1e-4, 1e-4, (a.Ny, a.Nx))
    a.f = compute_feq(np.ones_like(uy), np.full_like(uy, 0.05), uy)
    b.set_equilibrium(uy=uy)
    print("[gpu] equilibrium trajectories", flush=True)
    a.run(steps=60, check_every=30, verbose=False)
    b.run(steps=60, check_every=30, verbose=False)
    assert np.array_equal(np.asarray(a.Cd_history), np.asarray(b.Cd_history))
    assert np.array_equal(a.f, b.f)


def test_equilibrium_rejects_bad_fields():
    s = _solver2d()
    with pytest.raises(ValueError):
        s.set_equilibrium(uy=np.zeros((3, 3)))
    with pytest.raises(ValueError):
        s.set_equilibrium(uz=np.zeros((s.Ny, s.Nx)))
