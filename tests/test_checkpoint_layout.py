"""Checkpoint parameters stay interchangeable with the reference (CPU only).

``params_json`` must hold only the reference constructors' arguments
(aero/lbm/solver.py:75-116, aero/lbm/solver3d.py:71-119, names recorded from
the reference in tests/golden/ref_signatures.json) with a backend the
reference accepts, so the reference's from_checkpoint (solver.py:711-727,
solver3d.py:816-838) can rebuild a solver from a checkpoint written here;
this build's own arguments go under a key of their own.  Checkpoints the
reference wrote must map onto this build's constructors.
"""

import inspect
import json

import numpy as np
import pytest

from tests import _golden as G

SIG = json.loads((G.GOLDEN / "ref_signatures.json").read_text())
REF_BACKENDS = {"auto", "numpy", "numba"}


class _Any:
    """Stands in for a solver: every attribute reads as 0."""

    def __getattr__(self, name):
        return 0


@pytest.mark.parametrize("cls_name", ["Solver", "Solver3D"])
def test_params_json_keys_are_reference_constructor_arguments(cls_name):
    import aero_amd
    cls = getattr(aero_amd, cls_name)
    params = cls._params(_Any())
    assert set(params) <= set(SIG[cls_name]), set(params) - set(SIG[cls_name])
    assert params["backend"] in REF_BACKENDS
    extra = cls._extra_params(_Any())
    assert not set(extra) & set(SIG[cls_name])
    # json round trip (what np.savez stores)
    json.loads(json.dumps(params))


@pytest.mark.parametrize("path", sorted(p.name for p in G.GOLDEN.glob("ckpt_ref*d_*.npz") if "_cont" not in p.name))
def test_reference_checkpoint_params_fit_our_constructors(path):
    import aero_amd
    from aero_amd._base import checkpoint_params
    with np.load(G.GOLDEN / path, allow_pickle=False) as d:
        params = checkpoint_params(d)
        is3d = "Cly_history" in d.files
    cls = aero_amd.Solver3D if is3d else aero_amd.Solver
    names = set(inspect.signature(cls.__init__).parameters)
    assert set(params) <= names, set(params) - names


def test_checkpoint_params_merges_this_builds_entries():
    from aero_amd._base import EXTRA_PARAMS_KEY, checkpoint_params, extra_params_entry
    data = {"params_json": np.array(json.dumps({"Ny": 4, "backend": "auto"})),
            **extra_params_entry({"storage": "f16", "trt_variant": "lambda"})}
    assert EXTRA_PARAMS_KEY in data
    assert checkpoint_params(data) == {"Ny": 4, "backend": "auto", "storage": "f16", "trt_variant": "lambda"}
