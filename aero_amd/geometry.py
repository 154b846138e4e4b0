"""Obstacle geometry: solid masks and signed-distance fields (host, setup only).

Restates the reference's analytic shapes with the same conventions so masks
and q values are bit-identical:
* masks test ``<=`` on integer cell coordinates,
* SDFs sample cell centres offset by +0.5 (which the reference's Bouzidi q
  takes as the wall fraction, aero/lbm/boundary.py:273-278).

2D: aero/geometry/{cylinder,rectangle,polygon}.py
3D: aero/geometry3d/{sphere,box,cylinder3d}.py
"""

from __future__ import annotations

import threading
from typing import List, Optional, Tuple

import numpy as np


class Geometry:
    """2D obstacle (aero/geometry/base.py:8-39)."""

    def mark_solid(self, Ny: int, Nx: int) -> np.ndarray:
        raise NotImplementedError

    def center(self, Ny: int, Nx: int) -> Tuple[float, float]:
        raise NotImplementedError

    def reference_length(self) -> float:
        raise NotImplementedError

    def sdf_field(self, Ny: int, Nx: int, offset: float = 0.5) -> Optional[np.ndarray]:
        """SDF at cell coordinates + offset (the reference samples +0.5)."""
        return None


class Cylinder(Geometry):
    """aero/geometry/cylinder.py:8-42"""

    def __init__(self, radius: float = 20.0, cx_frac: float = 1 / 3, cy_frac: float = 0.5):
        self.radius, self.cx_frac, self.cy_frac = radius, cx_frac, cy_frac

    def center(self, Ny, Nx):
        return self.cx_frac * Nx, self.cy_frac * Ny

    def mark_solid(self, Ny, Nx):
        cx, cy = self.center(Ny, Nx)
        xx, yy = np.meshgrid(np.arange(Nx, dtype=np.float64), np.arange(Ny, dtype=np.float64))
        return (xx - cx) ** 2 + (yy - cy) ** 2 <= self.radius ** 2

    def reference_length(self):
        return 2.0 * self.radius

    def sdf_field(self, Ny, Nx, offset=0.5):
        cx, cy = self.center(Ny, Nx)
        xx, yy = np.meshgrid(np.arange(Nx, dtype=np.float64) + offset, np.arange(Ny, dtype=np.float64) + offset)
        return np.sqrt((xx - cx) ** 2 + (yy - cy) ** 2) - self.radius


class Rectangle(Geometry):
    """aero/geometry/rectangle.py:8-55"""

    def __init__(self, width: float = 40.0, height: float = 20.0, cx_frac: float = 1 / 3, cy_frac: float = 0.5):
        self.width, self.height, self.cx_frac, self.cy_frac = width, height, cx_frac, cy_frac

    def center(self, Ny, Nx):
        return self.cx_frac * Nx, self.cy_frac * Ny

    def mark_solid(self, Ny, Nx):
        cx, cy = self.center(Ny, Nx)
        xx, yy = np.meshgrid(np.arange(Nx, dtype=np.float64), np.arange(Ny, dtype=np.float64))
        return (np.abs(xx - cx) <= self.width / 2) & (np.abs(yy - cy) <= self.height / 2)

    def reference_length(self):
        return self.height

    def sdf_field(self, Ny, Nx, offset=0.5):
        cx, cy = self.center(Ny, Nx)
        xx, yy = np.meshgrid(np.arange(Nx, dtype=np.float64) + offset, np.arange(Ny, dtype=np.float64) + offset)
        dx = np.abs(xx - cx) - self.width / 2.0
        dy = np.abs(yy - cy) - self.height / 2.0
        return np.sqrt(np.maximum(dx, 0.0) ** 2 + np.maximum(dy, 0.0) ** 2) + np.minimum(np.maximum(dx, dy), 0.0)


_MPL_IMPORT = threading.Lock()


class Polygon(Geometry):
    """Polygon with vertices in fractional coordinates (aero/geometry/polygon.py:16-62).

    The mask uses matplotlib's ``Path.contains_points`` on integer cell
    coordinates with explicit MOVETO/LINETO/CLOSEPOLY codes, as the reference
    does, so boundary cells classify identically.
    """

    def __init__(self, vertices: List[Tuple[float, float]]):
        if len(vertices) < 3:
            raise ValueError("Polygon requires at least 3 vertices.")
        self.vertices = list(vertices)

    def mark_solid(self, Ny, Nx):
        with _MPL_IMPORT:  # case builders run on several threads; matplotlib's first import is not thread-safe
            from matplotlib.path import Path
        verts = [(x * Nx, y * Ny) for x, y in self.vertices]
        n = len(verts)
        path = Path(list(verts) + [verts[0]], [Path.MOVETO] + [Path.LINETO] * (n - 1) + [Path.CLOSEPOLY])
        xx, yy = np.meshgrid(np.arange(Nx), np.arange(Ny))
        pts = np.column_stack([xx.ravel().astype(float), yy.ravel().astype(float)])
        return path.contains_points(pts).reshape(Ny, Nx)

    def center(self, Ny, Nx):
        return (float(np.mean([v[0] * Nx for v in self.vertices])),
                float(np.mean([v[1] * Ny for v in self.vertices])))

    def reference_length(self):
        xs = [v[0] for v in self.vertices]
        ys = [v[1] for v in self.vertices]
        return float(max(max(xs) - min(xs), max(ys) - min(ys)))


class Geometry3D:
    """3D obstacle (aero/geometry3d/base.py:7-25)."""

    def mark_solid(self, Nz: int, Ny: int, Nx: int) -> np.ndarray:
        raise NotImplementedError

    def center(self, Nz: int, Ny: int, Nx: int) -> Tuple[float, float, float]:
        raise NotImplementedError

    def reference_length(self) -> float:
        raise NotImplementedError

    def sdf_field(self, Nz: int, Ny: int, Nx: int, offset: float = 0.5) -> Optional[np.ndarray]:
        """SDF at cell coordinates + offset (the reference samples +0.5)."""
        return None


def _grid3(Nz, Ny, Nx, off=0.0):
    z = np.arange(Nz, dtype=np.float64) + off
    y = np.arange(Ny, dtype=np.float64) + off
    x = np.arange(Nx, dtype=np.float64) + off
    return np.meshgrid(z, y, x, indexing="ij")


class Sphere(Geometry3D):
    """aero/geometry3d/sphere.py:7-51"""

    def __init__(self, radius: float = 10.0, cx_frac: float = 1 / 3, cy_frac: float = 0.5, cz_frac: float = 0.5):
        self.radius, self.cx_frac, self.cy_frac, self.cz_frac = radius, cx_frac, cy_frac, cz_frac

    def center(self, Nz, Ny, Nx):
        return self.cx_frac * Nx, self.cy_frac * Ny, self.cz_frac * Nz

    def mark_solid(self, Nz, Ny, Nx):
        cx, cy, cz = self.center(Nz, Ny, Nx)
        zz, yy, xx = _grid3(Nz, Ny, Nx)
        return (xx - cx) ** 2 + (yy - cy) ** 2 + (zz - cz) ** 2 <= self.radius ** 2

    def reference_length(self):
        return 2.0 * self.radius

    def sdf_field(self, Nz, Ny, Nx, offset=0.5):
        cx, cy, cz = self.center(Nz, Ny, Nx)
        zz, yy, xx = _grid3(Nz, Ny, Nx, offset)
        return np.sqrt((xx - cx) ** 2 + (yy - cy) ** 2 + (zz - cz) ** 2) - self.radius


class Box(Geometry3D):
    """aero/geometry3d/box.py:7-69"""

    def __init__(self, width: float = 10.0, height: float = 10.0, depth: float = 10.0,
                 cx_frac: float = 1 / 3, cy_frac: float = 0.5, cz_frac: float = 0.5):
        self.width, self.height, self.depth = width, height, depth
        self.cx_frac, self.cy_frac, self.cz_frac = cx_frac, cy_frac, cz_frac

    def center(self, Nz, Ny, Nx):
        return self.cx_frac * Nx, self.cy_frac * Ny, self.cz_frac * Nz

    def mark_solid(self, Nz, Ny, Nx):
        cx, cy, cz = self.center(Nz, Ny, Nx)
        zz, yy, xx = _grid3(Nz, Ny, Nx)
        return ((np.abs(xx - cx) <= self.width / 2.0) & (np.abs(yy - cy) <= self.height / 2.0)
                & (np.abs(zz - cz) <= self.depth / 2.0))

    def reference_length(self):
        return self.height

    def sdf_field(self, Nz, Ny, Nx, offset=0.5):
        cx, cy, cz = self.center(Nz, Ny, Nx)
        zz, yy, xx = _grid3(Nz, Ny, Nx, offset)
        dx = np.abs(xx - cx) - self.width / 2.0
        dy = np.abs(yy - cy) - self.height / 2.0
        dz = np.abs(zz - cz) - self.depth / 2.0
        return (np.sqrt(np.maximum(dx, 0.0) ** 2 + np.maximum(dy, 0.0) ** 2 + np.maximum(dz, 0.0) ** 2)
                + np.minimum(np.maximum(np.maximum(dx, dy), dz), 0.0))


class Cylinder3D(Geometry3D):
    """Finite z-aligned cylinder (aero/geometry3d/cylinder3d.py:7-63)."""

    def __init__(self, radius: float = 10.0, length: float = 20.0, cx_frac: float = 1 / 3,
                 cy_frac: float = 0.5, cz_frac: float = 0.5):
        self.radius, self.length = radius, length
        self.cx_frac, self.cy_frac, self.cz_frac = cx_frac, cy_frac, cz_frac

    def center(self, Nz, Ny, Nx):
        return self.cx_frac * Nx, self.cy_frac * Ny, self.cz_frac * Nz

    def mark_solid(self, Nz, Ny, Nx):
        cx, cy, cz = self.center(Nz, Ny, Nx)
        zz, yy, xx = _grid3(Nz, Ny, Nx)
        return ((xx - cx) ** 2 + (yy - cy) ** 2 <= self.radius ** 2) & (np.abs(zz - cz) <= self.length / 2.0)

    def reference_length(self):
        return 2.0 * self.radius

    def sdf_field(self, Nz, Ny, Nx, offset=0.5):
        cx, cy, cz = self.center(Nz, Ny, Nx)
        zz, yy, xx = _grid3(Nz, Ny, Nx, offset)
        dr = np.sqrt((xx - cx) ** 2 + (yy - cy) ** 2) - self.radius
        dz = np.abs(zz - cz) - self.length / 2.0
        return np.sqrt(np.maximum(dr, 0.0) ** 2 + np.maximum(dz, 0.0) ** 2) + np.minimum(np.maximum(dr, dz), 0.0)
