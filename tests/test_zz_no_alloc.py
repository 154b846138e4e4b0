"""The allocation-free step entry point against aero_step (MI355X only).

aero_step_noalloc runs the same fused step kernels with its case table, tile
list, coefficient rows and reduction record in a file-scope device array
instead of allocated buffers.  Two calls in a row through the C ABI, the
second on a smaller grid, each checked bit for bit against aero_step on a
twin context from the same state (same kernels and tile order, so the bits
must agree); the scratch array keeps the first call's contents when the
second starts.  Contexts the entry point does not serve are rejected with an
error.
"""

import ctypes

import numpy as np
import pytest

from tests import _golden as G
from tests.test_gpu_parity import make_solver

pytestmark = pytest.mark.gpu


def _noalloc(solver, n):
    from aero_amd import _abi
    hist = np.empty((n, _abi.NCOEF), dtype=np.float64)
    _abi.check(_abi.lib().aero_step_noalloc(solver._ctx.handle, ctypes.c_int32(n), ctypes.c_int64(solver.step_count),
                                            _abi.ptr(hist)))
    solver.step_count += n
    solver._bump()
    return hist


def _sphere(storage):
    import aero_amd
    from aero_amd.geometry import Sphere
    from aero_amd.lattice import compute_feq_3d
    nz, ny, nx, r = 64, 64, 128, 8.0
    g = Sphere(radius=r, cx_frac=0.25)
    solid, phi = g.mark_solid(nz, ny, nx), g.sdf_field(nz, ny, nx, offset=0.0)
    s = aero_amd.Solver3D(Nz=nz, Ny=ny, Nx=nx, solid=solid, omega=1.0 / (3.0 * 0.05 * 16 / 100 + 0.5), u0=0.05,
                          D=16.0, phi=phi, bouzidi=True, collision="trt", trt_lambda=0.1875, trt_variant="lambda",
                          bouzidi_variant="linear", storage=storage)
    rng = np.random.default_rng(1)
    sh = (nz, ny, nx)
    s.f = compute_feq_3d(np.ones(sh), np.full(sh, 0.05), rng.uniform(-1e-4, 1e-4, sh), rng.uniform(-1e-4, 1e-4, sh))
    return s


def test_noalloc_step_matches_aero_step_twice_in_a_row():
    # call 1: a 128 x 64 x 64 sphere, 300 steps (two chunks of the 256-step loop)
    a, b = _sphere("f32"), _sphere("f32")
    ra = a._advance(300)
    rb = _noalloc(b, 300)
    np.testing.assert_array_equal(ra, rb)
    np.testing.assert_array_equal(a.f, b.f)
    np.testing.assert_array_equal(a._f_outlet_prev, b._f_outlet_prev)
    a.close()
    b.close()
    # call 2: a smaller 2D grid with a convective outlet, the scalar and a
    # 40-step run, on the same scratch
    case = G.load(2, "bgk_thermal_buoy")
    c, d = make_solver(2, case, "f32"), make_solver(2, case, "f32")
    rc = c._advance(40)
    rd = _noalloc(d, 40)
    np.testing.assert_array_equal(rc, rd)
    np.testing.assert_array_equal(c.f, d.f)
    np.testing.assert_array_equal(c.g, d.g)
    c.close()
    d.close()


def test_noalloc_step_rejects_what_it_does_not_serve():
    from aero_amd import _abi
    sem = make_solver(3, G.load(3, "bgk_sem"), "f32")
    with pytest.raises(RuntimeError, match="per-step planes"):
        _noalloc(sem, 2)
    les = make_solver(2, G.load(2, "bgk_les_smag"), "f32")
    with pytest.raises(RuntimeError, match="LES"):
        _noalloc(les, 2)
    sem.close()
    les.close()
    assert _abi.lib() is not None


def test_noalloc_step_serves_local_les():
    """The fused local-stress LES form (no split step) steps allocation-free,
    with the same bits as aero_step."""
    case = G.load(3, "trt_ibm_les")
    kw = dict(case["kw"])
    kw["les_form"] = "local"
    case = dict(case, kw=kw)
    a, b = make_solver(3, case, "f32"), make_solver(3, case, "f32")
    ra = a._advance(30)
    rb = _noalloc(b, 30)
    np.testing.assert_array_equal(ra, rb)
    np.testing.assert_array_equal(a.f, b.f)
    a.close()
    b.close()
