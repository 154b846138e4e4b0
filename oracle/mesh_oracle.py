"""CPU oracle for the STL geometry path -- TEST INFRASTRUCTURE ONLY.

A float64 numpy restatement of the two O(triangles x points) loops of the
reference's STL pipeline, vectorised over points and looped over triangles:

* ``inside``: parity of forward Moller-Trumbore hits along a fixed ray,
  eps = 1e-9 (aero/geometry3d/mesh_mask.py:14-45);
* ``sq_dist``: squared distance from points to one triangle by the region of
  the minimiser of |A + s e0 + t e1 - P|^2, with the reference's degenerate
  rule (aero/geometry3d/signed_distance.py:38-137), and ``min_sq_dist`` over
  all triangles (signed_distance.py:222-231).

Every dot product is written as (a0 b0 + a1 b1) + a2 b2 with separate
roundings, the order the device kernels use.  Only ``tests/`` and
``bench.py``'s CPU baseline use it, as the checker.

Pinning (tests/test_mesh.py):
* the reference's own known answers: the unit-cube inside / outside points
  (tests/test_mesh_mask.py:41-48), the point-to-triangle distances 0.25, 0,
  1 and the equilateral centroid (tests/test_signed_distance.py:56-99), the
  unit-sphere STL's narrow-band SDF within 0.1 cells RMS / 0.5 max of the
  analytic sphere (tests/test_signed_distance.py:106-165);
* closed forms: for convex meshes (cube, octahedron and the sample cube,
  cylinder and sphere) the parity test equals the half-space test for every point farther than 1e-6
  from the surface; distances to a triangle equal the closed form over the
  face, across each edge and in each vertex's Voronoi cone, and are never
  below a dense barycentric sampling.  (The reference's rule takes a fixed
  candidate -- vertex A, edge AB or edge AC -- in the three corner regions of
  the minimiser, so there it can exceed the true distance; restated as is.)
"""

from __future__ import annotations

import numpy as np

MT_EPS = 1e-9


def _dot(a, b):
    return (a[..., 0] * b[..., 0] + a[..., 1] * b[..., 1]) + a[..., 2] * b[..., 2]


def _cross(a, b):
    return np.stack([a[..., 1] * b[..., 2] - a[..., 2] * b[..., 1],
                     a[..., 2] * b[..., 0] - a[..., 0] * b[..., 2],
                     a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]], axis=-1)


def reference_ray() -> np.ndarray:
    r = np.array([1.0, 0.123456789, 0.987654321], dtype=np.float64)
    return r / np.linalg.norm(r)


def inside(points: np.ndarray, triangles: np.ndarray, ray=None) -> np.ndarray:
    """bool (M,): odd number of forward hits of the ray from each point."""
    ray = reference_ray() if ray is None else np.asarray(ray, dtype=np.float64)
    P = np.asarray(points, dtype=np.float64).reshape(-1, 3)
    parity = np.zeros(len(P), dtype=bool)
    for tri in np.asarray(triangles, dtype=np.float64):
        v0 = tri[0]
        e1 = tri[1] - v0
        e2 = tri[2] - v0
        pv = _cross(ray, e2)
        det = _dot(e1, pv)
        if abs(det) < MT_EPS:
            continue
        inv = 1.0 / det
        tv = P - v0
        u = _dot(tv, pv) * inv
        q = _cross(tv, e1)
        v = _dot(q, ray) * inv
        t = _dot(q, e2) * inv
        parity ^= (u >= 0.0) & (v >= 0.0) & (u + v <= 1.0) & (t > MT_EPS)
    return parity


def grid_centres(nz: int, ny: int, nx: int) -> np.ndarray:
    z, y, x = np.meshgrid(np.arange(nz) + 0.5, np.arange(ny) + 0.5, np.arange(nx) + 0.5, indexing="ij")
    return np.column_stack([x.ravel(), y.ravel(), z.ravel()])


def sq_dist(P: np.ndarray, tri: np.ndarray) -> np.ndarray:
    """(M,) squared distances from points P to triangle tri = (A, B, C)."""
    P = np.asarray(P, dtype=np.float64).reshape(-1, 3)
    A, B, C = (np.asarray(v, dtype=np.float64) for v in tri)
    e0, e1 = B - A, C - A
    df = A - P
    a, b, c = _dot(e0, e0), _dot(e0, e1), _dot(e1, e1)
    d, e = _dot(df, e0), _dot(df, e1)
    det = a * c - b * b
    if det < 1e-14:
        if a >= c:
            r = np.clip(-d / max(a, 1e-30), 0.0, 1.0)
            Q = A + r[:, None] * e0
        else:
            r = np.clip(-e / max(c, 1e-30), 0.0, 1.0)
            Q = A + r[:, None] * e1
    else:
        s = b * e - c * d
        t = b * d - a * e
        beyond, neg_s, neg_t = (s + t) > det, s < 0.0, t < 0.0
        Q = np.empty_like(P)  # the five regions below partition the points
        inv = 1.0 / det
        interior = ~neg_s & ~neg_t & ~beyond
        Q[interior] = A + (s[interior] * inv)[:, None] * e0 + (t[interior] * inv)[:, None] * e1
        vert = neg_s & neg_t & ~beyond
        Q[vert] = A
        ac = neg_s & ~vert
        Q[ac] = A + np.clip(-e[ac] / max(c, 1e-30), 0.0, 1.0)[:, None] * e1
        ab = neg_t & ~neg_s
        Q[ab] = A + np.clip(-d[ab] / max(a, 1e-30), 0.0, 1.0)[:, None] * e0
        bc = beyond & ~neg_s & ~neg_t
        if bc.any():
            BCv = C - B
            bc2 = _dot(BCv, BCv)
            if bc2 > 1e-14:
                r = np.clip(_dot(P[bc] - B, BCv) / bc2, 0.0, 1.0)
                Q[bc] = B + r[:, None] * BCv
            else:
                Q[bc] = B
    dv = P - Q
    return _dot(dv, dv)


def min_sq_dist(P: np.ndarray, triangles: np.ndarray) -> np.ndarray:
    P = np.asarray(P, dtype=np.float64).reshape(-1, 3)
    best = np.full(len(P), np.inf)
    for tri in np.asarray(triangles, dtype=np.float64):
        np.minimum(best, sq_dist(P, tri), out=best)
    return best


def dilate(mask: np.ndarray, steps: int) -> np.ndarray:
    out = mask.copy()
    for _ in range(steps):
        g = out.copy()
        for ax in range(3):
            sl_a = [slice(None)] * 3
            sl_b = [slice(None)] * 3
            sl_a[ax], sl_b[ax] = slice(1, None), slice(None, -1)
            g[tuple(sl_a)] |= out[tuple(sl_b)]
            g[tuple(sl_b)] |= out[tuple(sl_a)]
        out = g
    return out


def phi_field(placed_triangles: np.ndarray, nz: int, ny: int, nx: int, band_width: int = 3,
              far_value: float = 16.0) -> np.ndarray:
    """Narrow-band signed distance of already placed triangles."""
    ins = inside(grid_centres(nz, ny, nx), placed_triangles).reshape(nz, ny, nx)
    band = dilate(ins, band_width) & dilate(~ins, band_width)
    phi = np.where(ins, -far_value, far_value).astype(np.float64)
    idx = np.flatnonzero(band)
    if idx.size:
        pts = grid_centres(nz, ny, nx)[idx]
        dist = np.sqrt(np.maximum(min_sq_dist(pts, placed_triangles), 0.0))
        flat = phi.reshape(-1)
        flat[idx] = np.where(ins.reshape(-1)[idx], -dist, dist)
    return phi
