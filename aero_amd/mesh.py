"""STL obstacles: loading, tunnel placement, voxelisation and the
narrow-band signed-distance field.

Mirrors the reference's STL geometry API (aero/geometry3d/stl_io.py,
stl_prep.py, mesh_mask.py, signed_distance.py).  The host keeps the
one-off, O(vertices) steps in numpy (parsing, PCA orientation, rotation,
scaling) with the reference's arithmetic, so the placed triangles are the
same doubles.  The O(triangles x cells) parts -- the parity ray cast of every
cell centre and the exact point-to-triangle distance of the narrow band --
run on the device (``aero_mesh_inside``, ``aero_mesh_sqdist``); there is no
host fallback.
"""

from __future__ import annotations

import ctypes
import pathlib
import re
from typing import Optional, Tuple

import numpy as np

from . import _abi

# Parity-test ray: fixed and oblique so no cell centre sees a mesh edge
# symmetrically (mesh_mask.py:40-42).
_RAY = np.array([1.0, 0.123456789, 0.987654321], dtype=np.float64)
_RAY /= np.linalg.norm(_RAY)


# ------------------------------------------------------------------ STL I/O
def load_stl_triangles(path: str) -> np.ndarray:
    """(T, 3, 3) float64 vertices of an ASCII or binary STL (stl_io.py:12-27):
    binary when the byte count is exactly 84 + 50 T, else ASCII when the file
    starts with ``solid``, else binary."""
    data = pathlib.Path(path).read_bytes()
    if _binary_layout(data):
        return _binary(data)
    if data[:5].lower().startswith(b"solid"):
        return _ascii(data)
    return _binary(data)


def _binary_layout(data: bytes) -> bool:
    return len(data) >= 84 and len(data) == 84 + 50 * int(np.frombuffer(data, "<u4", 1, 80)[0])


_FACET = np.dtype([("normal", "<f4", (3,)), ("v", "<f4", (3, 3)), ("attr", "<u2")])


def _binary(data: bytes) -> np.ndarray:
    if len(data) < 84:
        raise ValueError("STL file too small to be valid binary STL")
    n = int(np.frombuffer(data, "<u4", 1, 80)[0])
    if len(data) < 84 + 50 * n:
        raise ValueError("Binary STL truncated")
    facets = np.frombuffer(data, _FACET, n, 84)
    return facets["v"].astype(np.float64)


_VERTEX = re.compile(rb"^\s*vertex\s+(\S+)\s+(\S+)\s+(\S+)", re.IGNORECASE | re.MULTILINE)


def _ascii(data: bytes) -> np.ndarray:
    text = data.replace(b"\r\n", b"\n").replace(b"\r", b"\n")
    vals = [float(c.decode("utf-8", "replace")) for m in _VERTEX.finditer(text) for c in m.groups()]
    n = len(vals) // 9
    if n == 0:
        raise ValueError("No triangles found in ASCII STL")
    return np.array(vals[: 9 * n], dtype=np.float64).reshape(n, 3, 3)


def triangle_bounds(triangles: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    """(min_xyz, max_xyz) over all vertices (stl_io.py:74-79)."""
    v = triangles.reshape(-1, 3)
    return v.min(axis=0), v.max(axis=0)


# ------------------------------------------------------- tunnel placement
def orient_mesh_for_tunnel(triangles: np.ndarray) -> np.ndarray:
    """Centre at the vertex mean and rotate the principal axes onto the
    tunnel (stl_prep.py:8-28): middle variance -> stream x, largest -> span y,
    smallest -> thin z, right-handed."""
    v = triangles.reshape(-1, 3)
    xc = v - v.mean(axis=0)
    if len(xc) < 3:
        return triangles - v.mean(axis=0)
    axes = np.linalg.eigh(np.cov(xc.T))[1]  # columns by ascending variance
    thin, stream, span = axes[:, 0].copy(), axes[:, 1].copy(), axes[:, 2].copy()
    if np.dot(np.cross(stream, span), thin) < 0.0:
        thin = -thin
    return (xc @ np.stack([stream, span, thin], axis=1)).reshape(-1, 3, 3)


def rotation_matrix_xyz(rx_deg: float, ry_deg: float, rz_deg: float) -> np.ndarray:
    """R = Rx Ry Rz for angles in degrees (stl_prep.py:31-41)."""
    a = np.deg2rad([rx_deg, ry_deg, rz_deg])
    # scalar cos / sin per angle (an array call may take another code path)
    c = [np.cos(a[k]) for k in range(3)]
    s = [np.sin(a[k]) for k in range(3)]
    rx = np.array([[1.0, 0.0, 0.0], [0.0, c[0], -s[0]], [0.0, s[0], c[0]]], dtype=np.float64)
    ry = np.array([[c[1], 0.0, s[1]], [0.0, 1.0, 0.0], [-s[1], 0.0, c[1]]], dtype=np.float64)
    rz = np.array([[c[2], -s[2], 0.0], [s[2], c[2], 0.0], [0.0, 0.0, 1.0]], dtype=np.float64)
    return rx @ ry @ rz


def apply_mesh_rotation(triangles: np.ndarray, rx_deg: float, ry_deg: float, rz_deg: float) -> np.ndarray:
    """Rotate about the vertex mean (stl_prep.py:44-57); no-op for zero angles."""
    if max(abs(rx_deg), abs(ry_deg), abs(rz_deg)) < 1e-12:
        return triangles
    v = triangles.reshape(-1, 3)
    return ((v - v.mean(axis=0)) @ rotation_matrix_xyz(rx_deg, ry_deg, rz_deg)).reshape(-1, 3, 3)


def scale_and_place_mesh(triangles: np.ndarray, Nz: int, Ny: int, Nx: int, cx_frac: float, cy_frac: float,
                         cz_frac: float, fit_frac: float,
                         manual_scale: Optional[float] = None) -> Tuple[np.ndarray, float]:
    """Scale the larger cross-stream extent to fit_frac * min(Ny, Nz) (times
    manual_scale) and add (cx_frac Nx, cy_frac Ny, cz_frac Nz); returns the
    placed triangles and the reference length, the scaled cross-stream
    extent, at least 1 (stl_prep.py:60-92)."""
    lo, hi = triangle_bounds(triangles)
    extent = np.maximum(hi - lo, 1e-6)
    cross = max(float(extent[1]), float(extent[2]))
    scale = fit_frac * min(Ny, Nz) / cross
    if manual_scale is not None:
        scale = scale * manual_scale
    tris = triangles * scale
    tris[:, :, 0] += cx_frac * Nx
    tris[:, :, 1] += cy_frac * Ny
    tris[:, :, 2] += cz_frac * Nz
    return tris, max(cross * scale, 1.0)


def prepare_mesh_triangles(triangles: np.ndarray, Nz: int, Ny: int, Nx: int, cx_frac: float = 1.0 / 3.0,
                           cy_frac: float = 0.5, cz_frac: float = 0.5, fit_frac: float = 0.35,
                           mesh_orient: str = "auto", manual_scale: Optional[float] = None,
                           mesh_rot_x: float = 0.0, mesh_rot_y: float = 0.0,
                           mesh_rot_z: float = 0.0) -> Tuple[np.ndarray, float]:
    """Orient ("auto": PCA, "none": centre only), rotate, scale and place on
    the grid (stl_prep.py:95-121)."""
    if mesh_orient == "auto":
        tris = orient_mesh_for_tunnel(triangles)
    elif mesh_orient in ("none", ""):
        v = triangles.reshape(-1, 3)
        tris = (v - v.mean(axis=0)).reshape(-1, 3, 3)
    else:
        raise ValueError(f"Unknown mesh_orient '{mesh_orient}'. Use 'auto' or 'none'.")
    tris = apply_mesh_rotation(tris, mesh_rot_x, mesh_rot_y, mesh_rot_z)
    return scale_and_place_mesh(tris, Nz, Ny, Nx, cx_frac, cy_frac, cz_frac, fit_frac, manual_scale)


def compute_frontal_blockage(solid: np.ndarray) -> Tuple[float, float, float]:
    """Largest y and z extents of the solid over x stations, as fractions of
    Ny and Nz, and their maximum (stl_prep.py:124-143)."""
    nz, ny, nx = solid.shape
    ys_any = solid.any(axis=0)  # (ny, nx)
    zs_any = solid.any(axis=1)  # (nz, nx)
    by = bz = 0.0
    for xi in np.nonzero(ys_any.any(axis=0))[0]:
        ys = np.nonzero(ys_any[:, xi])[0]
        zs = np.nonzero(zs_any[:, xi])[0]
        by = max(by, (ys[-1] - ys[0] + 1) / max(ny, 1))
        bz = max(bz, (zs[-1] - zs[0] + 1) / max(nz, 1))
    return float(by), float(bz), float(max(by, bz))


def mesh_midplane_slice(solid: np.ndarray, axis: str = "y") -> np.ndarray:
    """Mid-plane slice of the mask for previews (stl_prep.py:146-157)."""
    nz, ny, nx = solid.shape
    if axis == "y":
        return solid[:, ny // 2, :].astype(np.float64)
    if axis == "z":
        return solid[nz // 2, :, :].astype(np.float64)
    return solid[:, :, nx // 2].astype(np.float64)


# --------------------------------------------------------- device queries
def _tris(triangles: np.ndarray) -> np.ndarray:
    t = np.ascontiguousarray(triangles, dtype=np.float64)
    if t.ndim != 3 or t.shape[1:] != (3, 3):
        raise ValueError(f"triangles must have shape (T, 3, 3), got {t.shape}")
    return t


def points_inside_mesh(points: np.ndarray, triangles: np.ndarray) -> np.ndarray:
    """Parity ray cast of arbitrary (M, 3) points (mesh_mask.py:40-45), on the
    device."""
    t = _tris(triangles)
    p = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
    out = np.zeros(len(p), dtype=np.uint8)
    _abi.check(_abi.lib().aero_mesh_inside(_abi.ptr(t), ctypes.c_int64(len(t)), _abi.ptr(_RAY), _abi.ptr(p),
                                           ctypes.c_int64(len(p)), 0, 0, 0, _abi.ptr(out)))
    return out.astype(bool)


def grid_inside_mesh(triangles: np.ndarray, Nz: int, Ny: int, Nx: int) -> np.ndarray:
    """(Nz, Ny, Nx) parity ray cast of every cell centre (x + 0.5, ...)."""
    t = _tris(triangles)
    out = np.zeros((Nz, Ny, Nx), dtype=np.uint8)
    _abi.check(_abi.lib().aero_mesh_inside(_abi.ptr(t), ctypes.c_int64(len(t)), _abi.ptr(_RAY), None,
                                           ctypes.c_int64(0), int(Nx), int(Ny), int(Nz), _abi.ptr(out)))
    return out.astype(bool)


def cell_sq_distance(triangles: np.ndarray, cells: np.ndarray, Nz: int, Ny: int, Nx: int) -> np.ndarray:
    """Squared distance of the listed cell centres (linear indices) to the
    nearest triangle, on the device."""
    t = _tris(triangles)
    c = np.ascontiguousarray(cells, dtype=np.int64).ravel()
    out = np.empty(len(c), dtype=np.float64)
    _abi.check(_abi.lib().aero_mesh_sqdist(_abi.ptr(t), ctypes.c_int64(len(t)), _abi.ptr(c), ctypes.c_int64(len(c)),
                                           int(Nx), int(Ny), int(Nz), _abi.ptr(out)))
    return out


# -------------------------------------------------------------- geometry
class MeshMask:
    """A watertight STL obstacle voxelised onto the grid (mesh_mask.py:48-133):
    PCA-oriented (or centred only), rotated, scaled so its larger
    cross-stream extent is fit_frac * min(Ny, Nz), and centred at
    (cx_frac Nx, cy_frac Ny, cz_frac Nz)."""

    def __init__(self, path: str, cx_frac: float = 1.0 / 3.0, cy_frac: float = 0.5, cz_frac: float = 0.5,
                 fit_frac: float = 0.35, mesh_orient: str = "auto", scale: Optional[float] = None,
                 mesh_rot_x: float = 0.0, mesh_rot_y: float = 0.0, mesh_rot_z: float = 0.0):
        self.path = path
        self.cx_frac, self.cy_frac, self.cz_frac = cx_frac, cy_frac, cz_frac
        self.fit_frac = max(float(fit_frac), 0.05)
        self.mesh_orient = mesh_orient
        self.mesh_rot_x, self.mesh_rot_y, self.mesh_rot_z = float(mesh_rot_x), float(mesh_rot_y), float(mesh_rot_z)
        self.manual_scale = float(scale) if scale is not None else None
        self._triangles = load_stl_triangles(path)
        self._reference_length: Optional[float] = None

    def _transform_to_grid(self, Nz: int, Ny: int, Nx: int) -> Tuple[np.ndarray, float]:
        return prepare_mesh_triangles(self._triangles, Nz, Ny, Nx, self.cx_frac, self.cy_frac, self.cz_frac,
                                      self.fit_frac, self.mesh_orient, self.manual_scale, self.mesh_rot_x,
                                      self.mesh_rot_y, self.mesh_rot_z)

    def mark_solid(self, Nz: int, Ny: int, Nx: int) -> np.ndarray:
        tris, self._reference_length = self._transform_to_grid(Nz, Ny, Nx)
        return grid_inside_mesh(tris, Nz, Ny, Nx)

    def center(self, Nz: int, Ny: int, Nx: int) -> Tuple[float, float, float]:
        return self.cx_frac * Nx, self.cy_frac * Ny, self.cz_frac * Nz

    def reference_length(self) -> float:
        if self._reference_length is None:
            raise RuntimeError("Call mark_solid() before reference_length()")
        return self._reference_length


def _dilate(mask: np.ndarray, steps: int) -> np.ndarray:
    """Grow a 3D mask by `steps` face-neighbour layers (signed_distance.py:144-156)."""
    out = mask.copy()
    for _ in range(steps):
        g = out.copy()
        g[1:] |= out[:-1]
        g[:-1] |= out[1:]
        g[:, 1:] |= out[:, :-1]
        g[:, :-1] |= out[:, 1:]
        g[:, :, 1:] |= out[:, :, :-1]
        g[:, :, :-1] |= out[:, :, 1:]
        out = g
    return out


def compute_phi_field(triangles: np.ndarray, nz: int, ny: int, nx: int, cx_frac: float = 1.0 / 3.0,
                      cy_frac: float = 0.5, cz_frac: float = 0.5, fit_frac: float = 0.35,
                      mesh_orient: str = "auto", mesh_rot_x: float = 0.0, mesh_rot_y: float = 0.0,
                      mesh_rot_z: float = 0.0, band_width: int = 3, far_value: float = 16.0) -> np.ndarray:
    """Narrow-band signed distance on cell centres (signed_distance.py:163-239):
    phi < 0 inside, |phi| the exact distance to the nearest triangle within
    band_width face-neighbour layers of the surface, +-far_value beyond."""
    tris, _ = prepare_mesh_triangles(triangles, nz, ny, nx, cx_frac, cy_frac, cz_frac, fit_frac, mesh_orient,
                                     mesh_rot_x=mesh_rot_x, mesh_rot_y=mesh_rot_y, mesh_rot_z=mesh_rot_z)
    inside = grid_inside_mesh(tris, nz, ny, nx)
    band = _dilate(inside, band_width) & _dilate(~inside, band_width)
    phi = np.where(inside, -far_value, far_value).astype(np.float64)
    cells = np.flatnonzero(band)
    if cells.size == 0:
        return phi
    dist = np.sqrt(np.maximum(cell_sq_distance(tris, cells, nz, ny, nx), 0.0))
    flat = phi.reshape(-1)
    flat[cells] = np.where(inside.reshape(-1)[cells], -dist, dist)
    return phi
