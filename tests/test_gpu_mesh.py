"""STL geometry on the device against the mesh oracle (oracle/mesh_oracle.py).

The device kernels and the oracle evaluate the same products and sums in the
same order without fused multiply-adds, so masks must agree exactly and
squared distances to the last bit."""

import pathlib

import numpy as np
import pytest

from oracle import mesh_oracle as mo

pytestmark = pytest.mark.gpu

STL = pathlib.Path(__file__).resolve().parent / "golden" / "stl"

CASES = [  # (file, fit_frac, orient, grid nz, ny, nx, oracle sample: cells or None for all)
    ("unit_sphere.stl", 0.25, "none", (16, 16, 32), None),
    ("unit_cube.stl", 0.5, "auto", (24, 24, 48), None),
    ("cylinder.stl", 0.3, "auto", (20, 20, 40), None),
    ("sphere.stl", 0.4, "none", (18, 20, 36), None),
    ("cube_20mm.stl", 0.4, "auto", (20, 20, 40), None),
    ("simple_plane.stl", 0.35, "auto", (48, 48, 96), 40000),  # the airplane needs cells across its thin body
]


def _placed(name, fit, orient, grid):
    from aero_amd import mesh
    nz, ny, nx = grid
    tris, L = mesh.prepare_mesh_triangles(mesh.load_stl_triangles(str(STL / name)), nz, ny, nx, fit_frac=fit,
                                          mesh_orient=orient)
    return tris, L


@pytest.mark.parametrize("name,fit,orient,grid,sample", CASES)
def test_voxelisation_matches_the_oracle_exactly(name, fit, orient, grid, sample):
    from aero_amd import mesh
    nz, ny, nx = grid
    tris, _ = _placed(name, fit, orient, grid)
    print(f"[gpu] voxelise {name} {grid} ({len(tris)} triangles)", flush=True)
    got = mesh.grid_inside_mesh(tris, nz, ny, nx)
    assert 0 < got.sum() < got.size
    cells = np.arange(got.size) if sample is None else np.random.default_rng(1).choice(got.size, sample, False)
    want = mo.inside(mo.grid_centres(nz, ny, nx)[cells], tris)
    np.testing.assert_array_equal(got.reshape(-1)[cells], want)
    m = mesh.MeshMask(str(STL / name), fit_frac=fit, mesh_orient=orient)
    np.testing.assert_array_equal(m.mark_solid(nz, ny, nx), got)
    assert m.reference_length() >= 1.0


def test_points_inside_match_the_oracle_exactly():
    from aero_amd import mesh
    rng = np.random.default_rng(2)
    tris, _ = _placed("simple_plane.stl", 0.35, "auto", (24, 24, 48))
    lo, hi = mesh.triangle_bounds(tris)
    P = rng.uniform(lo - 1.0, hi + 1.0, (20000, 3))
    print("[gpu] points inside simple_plane", flush=True)
    np.testing.assert_array_equal(mesh.points_inside_mesh(P, tris), mo.inside(P, tris))
    assert mesh.points_inside_mesh(np.zeros((0, 3)), tris).shape == (0,)


@pytest.mark.parametrize("name", ["unit_sphere.stl", "cube_20mm.stl", "simple_plane.stl"])
def test_squared_distances_match_the_oracle(name):
    from aero_amd import mesh
    rng = np.random.default_rng(9)
    nz, ny, nx = 20, 20, 40
    tris, _ = _placed(name, 0.35, "auto", (nz, ny, nx))
    cells = rng.choice(nz * ny * nx, 3000, replace=False)
    print(f"[gpu] distances {name}", flush=True)
    got = mesh.cell_sq_distance(tris, cells, nz, ny, nx)
    want = mo.min_sq_dist(mo.grid_centres(nz, ny, nx)[cells], tris)
    np.testing.assert_array_equal(got, want)


def test_phi_field_matches_the_oracle_and_the_analytic_sphere():
    """The reference's accuracy bands (tests/test_signed_distance.py:106-165)
    and the oracle's field."""
    from aero_amd import mesh
    nz, ny, nx = 24, 24, 48
    tris_raw = mesh.load_stl_triangles(str(STL / "unit_sphere.stl"))
    print("[gpu] phi field unit_sphere", flush=True)
    phi = mesh.compute_phi_field(tris_raw, nz, ny, nx, fit_frac=0.5, mesh_orient="none", band_width=4, far_value=16.0)
    placed, _ = mesh.prepare_mesh_triangles(tris_raw, nz, ny, nx, fit_frac=0.5, mesh_orient="none")
    np.testing.assert_array_equal(phi, mo.phi_field(placed, nz, ny, nx, band_width=4, far_value=16.0))
    z, y, x = np.meshgrid(np.arange(nz) + 0.5, np.arange(ny) + 0.5, np.arange(nx) + 0.5, indexing="ij")
    ana = np.sqrt((x - nx / 3) ** 2 + (y - 12.0) ** 2 + (z - 12.0) ** 2) - 6.0
    band = np.abs(phi) < 14.0
    err = phi[band] - ana[band]
    assert np.sqrt(np.mean(err ** 2)) < 0.1 and np.abs(err).max() < 0.5
    assert (phi[ana < -1.0] < 0).all() and (phi[ana > 12.0] == 16.0).all()


def test_mesh_mask_known_counts():
    """The reference's own mask bands (tests/test_mesh_mask.py:51-58,
    tests/test_sample_stl.py:31-35)."""
    from aero_amd import mesh
    solid = mesh.MeshMask(str(STL / "unit_cube.stl"), fit_frac=0.5).mark_solid(Nz=24, Ny=24, Nx=48)
    assert solid.dtype == bool and solid.shape == (24, 24, 48) and 50 < solid.sum() < 5000
    s2 = mesh.MeshMask(str(STL / "unit_sphere.stl"), fit_frac=0.25).mark_solid(16, 16, 32)
    assert 0 < s2.sum() < s2.size
    _, _, blk = mesh.compute_frontal_blockage(
        mesh.MeshMask(str(STL / "simple_plane.stl"), fit_frac=0.4).mark_solid(48, 48, 96))
    assert 0.05 < blk < 0.95


def test_stl_obstacle_runs_in_the_solver():
    """A voxelised STL with Bouzidi links from its SDF steps with a finite drag."""
    import aero_amd
    from aero_amd import mesh
    nz, ny, nx = 24, 24, 48
    m = mesh.MeshMask(str(STL / "unit_sphere.stl"), fit_frac=0.3, mesh_orient="none")
    solid = m.mark_solid(nz, ny, nx)
    phi = mesh.compute_phi_field(m._triangles, nz, ny, nx, fit_frac=0.3, mesh_orient="none")
    D = m.reference_length()
    u0 = 0.05
    s = aero_amd.Solver3D(Nz=nz, Ny=ny, Nx=nx, solid=solid, omega=1.0 / (3 * u0 * D / 20 + 0.5), u0=u0, D=D,
                          phi=phi, bouzidi=True, collision="trt", trt_lambda=0.1875, trt_variant="lambda",
                          storage="f32")
    print("[gpu] stl solver run", flush=True)
    r = s.run(steps=200, check_every=100, verbose=False)
    assert np.isfinite(r["Cd_mean"]) and r["Cd_mean"] > 0
    s.close()


def test_ibm_short_run_on_the_cube():
    """tests/test_ibm.py:46-59: Guo IBM with an STL SDF runs without NaN."""
    import aero_amd
    from aero_amd import mesh
    phi = mesh.compute_phi_field(mesh.load_stl_triangles(str(STL / "unit_cube.stl")), 12, 12, 24)
    s = aero_amd.Solver3D(Nz=12, Ny=12, Nx=24, solid=phi <= 0, omega=1.0, u0=0.05, D=4.0, ibm_enabled=True,
                          phi=phi)
    print("[gpu] ibm cube", flush=True)
    r = s.run(steps=20, check_every=20, verbose=False)
    assert not np.isnan(r["Cd_mean"])


def test_ibm_and_voxel_drag_agree_at_re20():
    """tests/test_signed_distance.py:180-264: on the unit-sphere STL at Re 20
    both the voxel (bounce-back) and the IBM run give Cd in [1, 8], within
    40% of each other."""
    import aero_amd
    from aero_amd import mesh
    nz, ny, nx, fit, u0 = 16, 16, 32, 0.4, 0.05
    phi = mesh.compute_phi_field(mesh.load_stl_triangles(str(STL / "unit_sphere.stl")), nz, ny, nx, fit_frac=fit,
                                 mesh_orient="none", band_width=3, far_value=16.0)
    solid = phi <= 0.0
    D = 2.0 * fit * min(ny, nz) / 2.0
    omega = 1.0 / (3.0 * u0 * D / 20.0 + 0.5)
    cds = []
    for ibm in (False, True):
        s = aero_amd.Solver3D(Nz=nz, Ny=ny, Nx=nx, solid=solid, omega=omega, u0=u0, D=D, collision="bgk",
                              ibm_enabled=ibm, phi=phi if ibm else None)
        print(f"[gpu] re20 ibm={ibm}", flush=True)
        cds.append(s.run(steps=400, check_every=400, verbose=False)["Cd_mean"])
    print(f"[gpu] re20 Cd voxel {cds[0]:.3f} ibm {cds[1]:.3f}", flush=True)
    assert all(1.0 <= c <= 8.0 for c in cds)
    assert abs(cds[1] - cds[0]) / max(abs(cds[0]), 1e-6) < 0.40
