"""Surface-link lists for bounce-back (host, setup only).

Same definition and order as the reference (aero/lbm/boundary.py:232-280,
aero/lbm/boundary3d.py:279-330): a link (i, cell) exists when the cell is
fluid and its neighbour along e_i -- with the index CLIPPED to the domain,
not wrapped -- is solid; links are ordered by direction i = 1..Q-1 and then
row-major by cell.  q = clip(float32(phi[cell]), 1e-4, 1) when an SDF is
given, else 0.5.  Vectorised: no per-link Python loop.
"""

from __future__ import annotations

from typing import Optional, Tuple

import numpy as np

from .lattice import E, E3


def _links(solid3: np.ndarray, E_: np.ndarray) -> np.ndarray:
    if not solid3.any():
        return np.empty((0, 4), dtype=np.int32)
    # A link cell lies within one cell of a solid cell on every axis, so only
    # the solid's bounding box dilated by one needs scanning.  Cells on the
    # box faces are fluid, so clipping inside the box equals clipping inside
    # the domain wherever the box stops short of the domain edge.
    lo, hi = [], []
    for ax, n in enumerate(solid3.shape):
        idx = np.nonzero(solid3.any(axis=tuple(a for a in range(3) if a != ax)))[0]
        lo.append(max(int(idx[0]) - 1, 0))
        hi.append(min(int(idx[-1]) + 1, n - 1))
    sub = solid3[lo[0]:hi[0] + 1, lo[1]:hi[1] + 1, lo[2]:hi[2] + 1]
    nz, ny, nx = sub.shape
    fluid = ~sub
    zz, yy, xx = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    out = []
    for i in range(1, len(E_)):
        ex = int(E_[i, 0])
        ey = int(E_[i, 1])
        ez = int(E_[i, 2]) if E_.shape[1] > 2 else 0
        zn = np.clip(zz + ez, 0, nz - 1)
        yn = np.clip(yy + ey, 0, ny - 1)
        xn = np.clip(xx + ex, 0, nx - 1)
        mask = fluid & sub[zn, yn, xn]
        z, y, x = np.nonzero(mask)
        if z.size:
            out.append(np.stack([np.full(z.size, i), z + lo[0], y + lo[1], x + lo[2]], axis=1))
    if not out:
        return np.empty((0, 4), dtype=np.int32)
    return np.concatenate(out).astype(np.int32)


def _q(links_cells: Tuple[np.ndarray, ...], n: int, phi: Optional[np.ndarray]) -> np.ndarray:
    if phi is None or n == 0:
        return np.full(n, 0.5, dtype=np.float32)
    return np.clip(phi[links_cells].astype(np.float32), 1e-4, 1.0)


def build_surface_links(solid: np.ndarray, phi: Optional[np.ndarray] = None):
    """2D links (L, 3) = (i, y, x) and q (L,) float32."""
    solid = np.asarray(solid, dtype=bool)
    l4 = _links(solid[None], E)
    links = np.ascontiguousarray(l4[:, [0, 2, 3]])
    return links, _q((links[:, 1], links[:, 2]), len(links), phi)


def build_surface_links_3d(solid: np.ndarray, phi: Optional[np.ndarray] = None):
    """3D links (L, 4) = (i, z, y, x) and q (L,) float32."""
    solid = np.asarray(solid, dtype=bool)
    links = _links(solid, E3)
    return links, _q((links[:, 1], links[:, 2], links[:, 3]), len(links), phi)


def link_fractions(links: np.ndarray, phi: np.ndarray) -> np.ndarray:
    """Wall position along each link for the linear Bouzidi rule.

    ``phi`` is a signed distance sampled at the integer cell coordinates the
    solid mask is tested on (negative inside).  Along link (i, cell) the wall
    sits where the linear interpolant of phi between the fluid cell and its
    solid neighbour crosses zero: q = phi(x) / (phi(x) - phi(x + e_i)),
    clipped to [1e-4, 1].  (The reference instead takes q = phi at the fluid
    cell of a +0.5-offset SDF, aero/lbm/boundary.py:273-278.)
    """
    E_ = E3 if links.shape[1] == 4 else E
    ph = phi if phi.ndim == 3 else phi[None]
    nz, ny, nx = ph.shape
    if links.shape[1] == 4:
        i, z, y, x = (links[:, k] for k in range(4))
        ez = E_[i, 2].astype(int)
    else:
        i, y, x = (links[:, k] for k in range(3))
        z = np.zeros_like(i)
        ez = np.zeros(len(i), int)
    zn = np.clip(z + ez, 0, nz - 1)
    yn = np.clip(y + E_[i, 1].astype(int), 0, ny - 1)
    xn = np.clip(x + E_[i, 0].astype(int), 0, nx - 1)
    a = ph[z, y, x]
    b = ph[zn, yn, xn]
    den = np.where(a - b != 0.0, a - b, 1.0)
    q = np.where(a - b != 0.0, a / den, 0.5)
    return np.clip(q, 1e-4, 1.0).astype(np.float32)
