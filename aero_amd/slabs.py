"""z-slab decomposition of one D3Q19 domain: the one-device part of a
multi-GPU run.

The reference distributes Solver3D over MPI ranks in z slabs on the host,
each padded with one ghost plane on either side and exchanging whole
boundary planes (all 19 populations) with MPI Sendrecv after every step
(aero/lbm/mpi_solver3d.py:116-150).  Here a slab is an ordinary Solver3D context
whose grid holds its own planes plus one halo plane below and one above; the
row-interleaved population storage keeps a slab contiguous.  Each step a
slab needs, in its halo planes, the 5 populations its neighbours' boundary
planes send across the face (e_z = +1 upward, -1 downward); the library
packs and unpacks exactly those (aero_halo_pack / aero_halo_unpack, a
[5][ny][nx] buffer per face) and, for slabs on one device,
aero_halo_exchange does both faces of a pair.  Across GPUs the same buffers
would travel by RCCL send / recv between neighbours.

Cells of the halo planes are stepped with the rest (their results are
overwritten by the next exchange before anything reads them), and carry no
surface links, so the force sums of the slabs add up to the whole domain's.
The domain is periodic in z, as the reference's np.roll streaming is.
"""

from __future__ import annotations

import ctypes
from typing import List, Sequence, Tuple

import numpy as np

from . import _abi


def slab_bounds(nz: int, nslabs: int) -> List[Tuple[int, int]]:
    """[z0, z1) of each slab, as even as the planes allow."""
    if nslabs < 1 or nslabs > nz:
        raise ValueError("need 1 <= nslabs <= nz")
    edges = np.linspace(0, nz, nslabs + 1).round().astype(int)
    return [(int(edges[k]), int(edges[k + 1])) for k in range(nslabs)]


def with_halos(a: np.ndarray, z0: int, z1: int, axis: int) -> np.ndarray:
    """Planes z0-1 .. z1 of `a` along `axis` (periodic): a slab and its halos."""
    nz = a.shape[axis]
    idx = np.arange(z0 - 1, z1 + 1) % nz
    return np.ascontiguousarray(np.take(a, idx, axis=axis))


def make_slab(solid: np.ndarray, z0: int, z1: int, **kw):
    """A Solver3D for planes [z0, z1) of the (nz, ny, nx) mask `solid`, with
    halo planes; surface links only on its own planes (global geometry, so a
    cell next to a solid across the face keeps its link)."""
    from .solver3d import Solver3D
    phi = kw.pop("phi", None)
    sl = with_halos(solid, z0, z1, 0)
    s = Solver3D(Nz=sl.shape[0], Ny=sl.shape[1], Nx=sl.shape[2], solid=sl,
                 phi=None if phi is None else with_halos(phi, z0, z1, 0), **kw)
    keep = (s.surface_links[:, 1] >= 1) & (s.surface_links[:, 1] <= sl.shape[0] - 2)
    s.surface_links, s.q_vals = s.surface_links[keep], s.q_vals[keep]
    sol = np.ascontiguousarray(sl, dtype=np.uint8)
    l4 = np.ascontiguousarray(s.surface_links, dtype=np.int32)
    qv = np.ascontiguousarray(s.q_vals, dtype=np.float32)
    s._ctx.call("aero_set_geometry", _abi.ptr(sol), ctypes.c_int64(len(l4)), _abi.ptr(l4), _abi.ptr(qv))
    s.slab = (z0, z1)
    return s


def exchange(slabs: Sequence, periodic: bool = True) -> None:
    """Halo faces between consecutive slabs (lowest first), and between the
    last and the first when the domain is periodic in z."""
    pairs = list(zip(slabs[:-1], slabs[1:]))
    if periodic and len(slabs) > 1:
        pairs.append((slabs[-1], slabs[0]))
    for lo, up in pairs:
        _abi.check(_abi.lib().aero_halo_exchange(lo._ctx.handle, up._ctx.handle))


def advance(slabs: Sequence, n: int, periodic: bool = True) -> np.ndarray:
    """n steps of the slab stack, exchanging halos before each step; returns
    the (n, 20) coefficient rows of the whole domain (forces and
    coefficients summed over the slabs; moments are each slab's own)."""
    rows = np.zeros((n, _abi.NCOEF))
    for k in range(n):
        exchange(slabs, periodic)
        for s in slabs:
            rows[k] += s._advance(1)[0]
    return rows
