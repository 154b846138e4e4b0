"""Serving loop on the device: results are the same alone and in the mix.

A stream of mixed requests (2D cylinders / rectangles / polygons under BGK,
TRT and MRT, a 3D sphere, one case with a passive scalar, 2D and 3D cases
with local-stress LES) is served with
continuous batching; every request is then run alone (a server with one
slot).  Stop step, stop reason and the whole drag history must agree
bit-for-bit.  MI355X only.
"""

import numpy as np
import pytest

from aero_amd import serve

pytestmark = pytest.mark.gpu


def _requests():
    specs = serve.c4_cases(6, seed=11, max_steps=1500)
    for s in specs:  # keep the grids small
        s.nx, s.ny = 128 + 32 * (s.seed % 3), (128 + 32 * (s.seed % 3)) // 2
        if s.shape == "polygon":
            s.vertices = [(0.25 + 10 * np.cos(a) / s.nx, 0.5 + 10 * np.sin(a) / s.ny)
                          for a in np.linspace(0, 2 * np.pi, 7)[:-1]]
    specs[1].thermal = True
    specs.append(serve.CaseSpec(dim=3, nx=64, ny=32, nz=32, shape="sphere", size=(4.0,), Re=80.0,
                                collision="trt", max_steps=1000, seed=5))
    # local-stress LES cases batch with the rest (their own kernel group)
    specs.append(serve.CaseSpec(dim=2, nx=160, ny=80, shape="cylinder", size=(8.0,), Re=400.0, collision="trt",
                                les_cs=0.17, max_steps=1500, seed=6))
    specs.append(serve.CaseSpec(dim=3, nx=64, ny=32, nz=32, shape="sphere", size=(4.0,), Re=300.0,
                                collision="bgk", les_cs=0.17, storage="f16", max_steps=1000, seed=7))
    arrivals = [0.0, 0.0, 0.02, 0.05, 0.05, 0.3, 0.1, 0.0, 0.05]
    return list(zip(arrivals, specs))


def test_served_results_equal_results_alone():
    reqs = _requests()
    mixed = serve.Server(chunk=500, keep_history=True).serve(reqs)
    alone = serve.Server(chunk=500, max_live=1, keep_history=True).serve([(0.0, s) for _, s in reqs])
    assert len(mixed) == len(alone) == len(reqs)
    m = {r.rid: r for r in mixed}
    a = {r.rid: r for r in alone}
    for rid in m:
        assert m[rid].steps == a[rid].steps, rid
        assert m[rid].stop_reason == a[rid].stop_reason, rid
        np.testing.assert_array_equal(np.array(m[rid].Cd_history), np.array(a[rid].Cd_history), err_msg=str(rid))
        assert np.isfinite(m[rid].Cd_mean), rid
