"""Checkpoints across the two implementations, and restarts of IBM runs.

* A checkpoint written by the reference's own save_checkpoint (fixtures made
  by tests/golden/make_golden.py: the reference ran 12 steps, saved, then ran
  9 more) is resumed here with from_checkpoint; 9 steps later the state and
  the drag match the reference's continued run (float64, 1e-10).
* An immersed-boundary run checkpointed mid-way and resumed with
  from_checkpoint continues the uninterrupted run (a new geometry must not
  drop the IBM band, whose forcing moves f by ~1e-3).  The restart collides
  its first state in the field-collision kernel, whose compiled FMA
  contraction can differ from the fused step kernel's in the last bit, hence
  1e-13 rather than bit equality, as for the other restart tests.
* A batch whose case list is rebuilt with a context of another collision at
  the same list position steps that context exactly as it steps alone.
MI355X only.
"""

import numpy as np
import pytest

from tests import _golden as G
from tests.test_gpu_parity import make_solver

pytestmark = pytest.mark.gpu

CKPTS = sorted(p.name[len("ckpt_"):-len(".npz")] for p in G.GOLDEN.glob("ckpt_ref*d_*.npz") if "_cont" not in p.name)


@pytest.mark.parametrize("name", CKPTS)
def test_resume_reference_written_checkpoint(name):
    import aero_amd
    cls = aero_amd.Solver3D if name.startswith("ref3d") else aero_amd.Solver
    s = cls.from_checkpoint(str(G.GOLDEN / f"ckpt_{name}.npz"))
    with np.load(G.GOLDEN / f"ckpt_{name}_cont.npz", allow_pickle=False) as d:
        f_ref, cd_ref, n = d["f"], d["Cd_hist"], int(d["steps"])
    cd = [s._step()[0] for _ in range(n)]
    assert G.rel_err(s.f, f_ref) < 1e-10, name
    np.testing.assert_allclose(cd, cd_ref, rtol=1e-9, atol=1e-12)
    s.close()


@pytest.mark.parametrize("dim,name", [(2, "bgk_ibm"), (2, "mrt_ibm_zg_sponge"), (3, "bgk_ibm"), (3, "trt_ibm_les")])
def test_ibm_restart_continues_the_run(dim, name, tmp_path):
    import aero_amd
    case = G.load(dim, name)
    ref = make_solver(dim, case)
    ref_rows = ref._advance(14)
    s = make_solver(dim, case)
    first = s._advance(6)
    path = tmp_path / "ck.npz"
    s.save_checkpoint(str(path))
    s.close()
    cls = aero_amd.Solver3D if dim == 3 else aero_amd.Solver
    r = cls.from_checkpoint(str(path))
    rest = r._advance(8)
    np.testing.assert_allclose(np.concatenate([first, rest]), ref_rows, rtol=1e-10, atol=1e-13)
    np.testing.assert_allclose(r.f, ref.f, rtol=0, atol=1e-13)
    ref.close()
    r.close()


def test_batch_retire_and_admit_other_collision_same_slot():
    """Case lists [A, P] then [B, P] with A retired (destroyed) before B of
    another collision and the same grid is created: B and P step as alone."""
    from aero_amd.batch import Batch
    from tests.test_gpu_batch import _cylinder
    b = Batch()
    probe = _cylinder(128, 64, 6.0, "trt", "f32", 3)
    a = _cylinder(160, 80, 6.0, "bgk", "f32", 4)
    b.advance([a, probe], 40)
    a.close()
    del a
    new = _cylinder(160, 80, 6.0, "mrt", "f32", 5)
    rows = b.advance([new, probe], 60)
    alone = _cylinder(160, 80, 6.0, "mrt", "f32", 5)
    np.testing.assert_array_equal(rows[0], alone._advance(60))
    np.testing.assert_array_equal(new.f, alone.f)
    ref = _cylinder(128, 64, 6.0, "trt", "f32", 3)
    ref._advance(100)
    np.testing.assert_array_equal(probe.f, ref.f)
    b.close()
