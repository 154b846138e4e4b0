"""STL path on the CPU: the loader and tunnel placement of aero_amd.mesh, and
the pinning of the mesh oracle (oracle/mesh_oracle.py) by the reference's
known answers and by closed forms."""

import pathlib
import struct

import numpy as np
import pytest

from aero_amd import mesh
from oracle import mesh_oracle as mo

STL = pathlib.Path(__file__).resolve().parent / "golden" / "stl"

CUBE = [  # unit cube [0, 1]^3, two triangles per face (tests/test_mesh_mask.py:13-26)
    [(0, 0, 0), (1, 0, 0), (1, 1, 0)], [(0, 0, 0), (1, 1, 0), (0, 1, 0)],
    [(0, 0, 1), (1, 1, 1), (1, 0, 1)], [(0, 0, 1), (0, 1, 1), (1, 1, 1)],
    [(0, 0, 0), (0, 1, 0), (0, 1, 1)], [(0, 0, 0), (0, 1, 1), (0, 0, 1)],
    [(1, 0, 0), (1, 0, 1), (1, 1, 1)], [(1, 0, 0), (1, 1, 1), (1, 1, 0)],
    [(0, 0, 0), (0, 0, 1), (1, 0, 1)], [(0, 0, 0), (1, 0, 1), (1, 0, 0)],
    [(0, 1, 0), (1, 1, 0), (1, 1, 1)], [(0, 1, 0), (1, 1, 1), (0, 1, 1)],
]


def _cube():
    return np.array(CUBE, dtype=np.float64)


def _octahedron(r=1.0, c=(0.0, 0.0, 0.0)):
    v = np.array([[r, 0, 0], [-r, 0, 0], [0, r, 0], [0, -r, 0], [0, 0, r], [0, 0, -r]], float) + c
    faces = [(0, 2, 4), (2, 1, 4), (1, 3, 4), (3, 0, 4), (2, 0, 5), (1, 2, 5), (3, 1, 5), (0, 3, 5)]
    return np.array([[v[a], v[b], v[c]] for a, b, c in faces])


def _write_binary(path, tris):
    with open(path, "wb") as fh:
        fh.write(b"\0" * 80 + struct.pack("<I", len(tris)))
        for t in tris:
            fh.write(struct.pack("<12f", 0, 0, 0, *np.ravel(t)) + struct.pack("<H", 0))


def _write_ascii(path, tris):
    lines = ["solid t"]
    for t in tris:
        lines += ["facet normal 0 0 0", " outer loop"] + [f"  vertex {a} {b} {c}" for a, b, c in t]
        lines += [" endloop", "endfacet"]
    pathlib.Path(path).write_text("\n".join(lines + ["endsolid t"]))


# ------------------------------------------------------------------ loader
@pytest.mark.parametrize("name", ["unit_sphere.stl", "unit_cube.stl", "cylinder.stl", "sphere.stl",
                                  "simple_plane.stl", "cube_20mm.stl"])
def test_sample_stl_loads(name):
    tris = mesh.load_stl_triangles(str(STL / name))
    assert tris.shape[1:] == (3, 3) and len(tris) > 0 and tris.dtype == np.float64
    lo, hi = mesh.triangle_bounds(tris)
    assert (hi > lo).all()


def test_binary_and_ascii_give_the_same_triangles(tmp_path):
    tris = _cube() * 0.7 + 0.1
    _write_binary(tmp_path / "b.stl", tris)
    _write_ascii(tmp_path / "a.stl", tris.astype(np.float32).astype(np.float64))
    b = mesh.load_stl_triangles(str(tmp_path / "b.stl"))
    a = mesh.load_stl_triangles(str(tmp_path / "a.stl"))
    assert b.shape == (12, 3, 3)
    np.testing.assert_array_equal(b, tris.astype(np.float32).astype(np.float64))
    np.testing.assert_array_equal(a, b)


def test_ascii_without_vertices_is_rejected(tmp_path):
    (tmp_path / "e.stl").write_text("solid empty\nendsolid empty\n")
    with pytest.raises(ValueError):
        mesh.load_stl_triangles(str(tmp_path / "e.stl"))


# --------------------------------------------------------------- placement
def _box(dx, dy, dz):
    return _cube() * np.array([dx, dy, dz])


def test_orientation_puts_the_longest_axis_on_y_and_the_thinnest_on_z():
    ext = np.ptp(mesh.orient_mesh_for_tunnel(_box(5.0, 20.0, 2.0)).reshape(-1, 3), axis=0)
    order = np.argsort(ext)
    assert order[2] == 1 and order[0] == 2


def test_rotation_about_z_swaps_the_x_and_y_extents():
    v = _box(10.0, 30.0, 4.0).reshape(-1, 3)
    c = (v - v.mean(axis=0)).reshape(-1, 3, 3)
    r = mesh.apply_mesh_rotation(c, 0.0, 0.0, 90.0)
    e0, e1 = np.ptp(c.reshape(-1, 3), axis=0), np.ptp(r.reshape(-1, 3), axis=0)
    assert abs(e1[0] - e0[1]) < 1e-9 and abs(e1[1] - e0[0]) < 1e-9
    R = mesh.rotation_matrix_xyz(10.0, -20.0, 30.0)
    np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-14)


def test_scale_and_place_fit_the_cross_section_and_centre():
    tris, L = mesh.prepare_mesh_triangles(_box(4.0, 2.0, 1.0), 32, 40, 80, fit_frac=0.5, mesh_orient="none")
    lo, hi = mesh.triangle_bounds(tris)
    assert abs(max(hi[1] - lo[1], hi[2] - lo[2]) - 0.5 * 32) < 1e-9 and abs(L - 16.0) < 1e-9
    np.testing.assert_allclose((lo + hi) / 2, [80 / 3, 20.0, 16.0], atol=1e-9)
    with pytest.raises(ValueError):
        mesh.prepare_mesh_triangles(_box(1, 1, 1), 8, 8, 8, mesh_orient="sideways")


def test_frontal_blockage_of_a_block():
    s = np.zeros((20, 30, 40), bool)
    s[5:9, 10:22, 12:20] = True
    by, bz, blk = mesh.compute_frontal_blockage(s)
    assert (by, bz, blk) == (12 / 30, 4 / 20, 12 / 30)


# ------------------------------------------------------ oracle: known answers
def test_oracle_unit_cube_inside_and_outside_points():
    ins = mo.inside(np.array([[0.5, 0.5, 0.5], [1.5, 0.5, 0.5]]), _cube())
    assert ins.tolist() == [True, False]


def test_oracle_point_to_triangle_distances():
    A, B, C = np.array([0.0, 0, 0]), np.array([1.0, 0, 0]), np.array([0.0, 1, 0])
    sq = mo.sq_dist(np.array([[1 / 3, 1 / 3, 0.5], [0.0, 0.0, 0.0], [-1.0, 0.5, 0.0]]), (A, B, C))
    assert abs(sq[0] - 0.25) < 1e-10 and sq[1] < 1e-12 and abs(sq[2] - 1.0) < 1e-9
    eq = (np.array([1.0, 0, 0]), np.array([-0.5, np.sqrt(3) / 2, 0]), np.array([-0.5, -np.sqrt(3) / 2, 0]))
    assert mo.sq_dist(np.zeros((1, 3)), eq)[0] < 1e-10


def test_oracle_sphere_sdf_against_the_analytic_sphere():
    """The reference's own accuracy bands for unit_sphere.stl
    (tests/test_signed_distance.py:106-165)."""
    nz, ny, nx = 24, 24, 48
    tris, _ = mesh.prepare_mesh_triangles(mesh.load_stl_triangles(str(STL / "unit_sphere.stl")), nz, ny, nx,
                                          fit_frac=0.5, mesh_orient="none")
    phi = mo.phi_field(tris, nz, ny, nx, band_width=4, far_value=16.0)
    z, y, x = np.meshgrid(np.arange(nz) + 0.5, np.arange(ny) + 0.5, np.arange(nx) + 0.5, indexing="ij")
    ana = np.sqrt((x - nx / 3) ** 2 + (y - 12.0) ** 2 + (z - 12.0) ** 2) - 6.0
    band = np.abs(phi) < 14.0
    err = phi[band] - ana[band]
    assert np.sqrt(np.mean(err ** 2)) < 0.1 and np.abs(err).max() < 0.5
    assert (phi[ana < -1.0] < 0).all() and (phi[(ana > 1.0) & (ana < 14.0)] > 0).all()
    assert (phi[ana > 12.0] == 16.0).all()


# ------------------------------------------------------- oracle: closed forms
def _halfspace_inside(P, tris):
    """Convex mesh: inside iff behind every face plane (outward normals from
    the centroid); returns (inside, distance to the nearest plane)."""
    cen = tris.reshape(-1, 3).mean(axis=0)
    n = np.cross(tris[:, 1] - tris[:, 0], tris[:, 2] - tris[:, 0])
    n /= np.linalg.norm(n, axis=1)[:, None]
    flip = np.einsum("ij,ij->i", n, tris[:, 0] - cen) < 0
    n[flip] *= -1
    s = np.einsum("tj,mj->mt", n, P) - np.einsum("tj,tj->t", n, tris[:, 0])[None, :]
    return (s < 0).all(axis=1), np.abs(s).min(axis=1)


@pytest.mark.parametrize("shape", ["cube", "octahedron", "unit_cube.stl", "cylinder.stl", "unit_sphere.stl"])
def test_oracle_parity_test_equals_the_half_space_test_on_convex_meshes(shape):
    rng = np.random.default_rng(3)
    if shape == "cube":
        tris = _cube()
    elif shape == "octahedron":
        tris = _octahedron(0.8, (0.5, 0.5, 0.5))
    else:  # convex sample meshes, scaled into the unit box
        tris = mesh.load_stl_triangles(str(STL / shape))
        lo, hi = mesh.triangle_bounds(tris)
        tris = (tris - lo) / (hi - lo)
    P = rng.uniform(-0.3, 1.3, (4000, 3))
    want, margin = _halfspace_inside(P, tris)
    keep = margin > 1e-6
    got = mo.inside(P, tris)
    assert keep.sum() > 3900
    np.testing.assert_array_equal(got[keep], want[keep])
    assert 0 < got.sum() < len(P)


def test_oracle_distance_is_attained_by_a_point_of_the_triangle():
    """The rule returns the distance to a point of the triangle, so it is never
    below the minimum over a dense barycentric sampling minus one sampling
    step.  (It may be above the true minimum: in the three corner regions the
    reference takes a fixed candidate, vertex A or edge AB / AC.)"""
    rng = np.random.default_rng(4)
    k = np.linspace(0.0, 1.0, 161)
    s, t = np.meshgrid(k, k)
    m = (s + t) <= 1.0
    s, t = s[m], t[m]
    for _ in range(20):
        tri = rng.normal(size=(3, 3))
        P = rng.normal(scale=2.0, size=(50, 3))
        d = np.sqrt(mo.sq_dist(P, tri))
        samp = tri[0] + s[:, None] * (tri[1] - tri[0]) + t[:, None] * (tri[2] - tri[0])
        dmin = np.sqrt(((P[:, None, :] - samp[None]) ** 2).sum(-1)).min(axis=1)
        h = max(np.linalg.norm(tri[1] - tri[0]), np.linalg.norm(tri[2] - tri[0]),
                np.linalg.norm(tri[2] - tri[1])) / 160
        assert (d >= dmin - h).all()


def test_oracle_distance_is_exact_across_each_edge_and_at_each_vertex():
    """On acute triangles (where every edge's and vertex's Voronoi region lies
    in the matching region of the minimiser), points whose closest point is
    inside an edge (offset perpendicular to it, in plane and along the
    normal) or a vertex (inside its Voronoi cone) have the closed-form
    distance."""
    rng = np.random.default_rng(7)
    done = 0
    while done < 10:
        tri = rng.normal(size=(3, 3))
        ang_ok = all(np.dot(tri[(i + 1) % 3] - tri[i], tri[(i + 2) % 3] - tri[i]) > 0 for i in range(3))
        if not ang_ok:
            continue
        done += 1
        n = np.cross(tri[1] - tri[0], tri[2] - tri[0])
        n /= np.linalg.norm(n)
        for i in range(3):
            a, b, c = tri[i], tri[(i + 1) % 3], tri[(i + 2) % 3]
            e = (b - a) / np.linalg.norm(b - a)
            out = np.cross(e, n)
            if np.dot(out, c - a) > 0:
                out = -out  # in-plane, away from the third vertex
            lam = rng.uniform(0.1, 0.9, 20)
            r, z = rng.uniform(0.05, 1.0, 20), rng.uniform(-1.0, 1.0, 20)
            P = a + lam[:, None] * (b - a) + r[:, None] * out + z[:, None] * n
            np.testing.assert_allclose(mo.sq_dist(P, tri), r * r + z * z, rtol=1e-11, atol=1e-13)
            # vertex a: directions with non-positive projection on both edges from a
            u, w = (b - a) / np.linalg.norm(b - a), (c - a) / np.linalg.norm(c - a)
            D = rng.normal(size=(200, 3))
            D = D[(D @ u < -0.05) & (D @ w < -0.05)]
            np.testing.assert_allclose(mo.sq_dist(a + D, tri), (D * D).sum(axis=1), rtol=1e-11, atol=1e-13)


def test_oracle_distance_equals_the_plane_distance_over_the_face():
    rng = np.random.default_rng(5)
    tri = np.array([[0.0, 0, 0], [2.0, 0, 0], [0.0, 3.0, 0]])
    uv = rng.uniform(0, 1, (200, 2))
    uv = uv[uv.sum(axis=1) < 0.95]
    h = rng.uniform(-2, 2, len(uv))
    P = np.column_stack([2 * uv[:, 0], 3 * uv[:, 1], h])
    np.testing.assert_allclose(mo.sq_dist(P, tri), h * h, rtol=0, atol=1e-13)


def test_a_planted_mistake_fails_the_convex_check():
    """The half-space check has the power to catch a plausible slip: the ray
    test with the wrong edge in q = tvec x edge1 (edge2 instead) disagrees."""
    P = np.random.default_rng(6).uniform(-0.3, 1.3, (2000, 3))
    tris = _octahedron(0.8, (0.5, 0.5, 0.5))
    want, margin = _halfspace_inside(P, tris)
    keep = margin > 1e-6

    def slipped(points, triangles):
        ray = mo.reference_ray()
        par = np.zeros(len(points), dtype=bool)
        for tri in triangles:
            e1, e2 = tri[1] - tri[0], tri[2] - tri[0]
            pv = mo._cross(ray, e2)
            det = mo._dot(e1, pv)
            if abs(det) < mo.MT_EPS:
                continue
            tv = points - tri[0]
            u = mo._dot(tv, pv) / det
            q = mo._cross(tv, e2)  # the slip
            v, t = mo._dot(q, ray) / det, mo._dot(q, e2) / det
            par ^= (u >= 0) & (v >= 0) & (u + v <= 1) & (t > mo.MT_EPS)
        return par

    assert np.array_equal(mo.inside(P, tris)[keep], want[keep])
    assert not np.array_equal(slipped(P, tris)[keep], want[keep])
