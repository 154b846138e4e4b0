"""Local Smagorinsky LES (les_form="local") on the GPU.

The fused local form has no counterpart in the reference, so its parity is
anchored on the oracle's restatement of the same closure (pinned in
tests/test_oracle.py by the Chapman-Enskog identity) composed into the
reference's step order, on the golden cases' geometries and states:

* FP64 storage: fields and every coefficient within 1e-10 of the oracle;
  FP32 storage within the FP32 parity tolerances of the golden replay;
* C_s = 0: the local kernels reduce to the plain step (1e-13 in FP64);
* batching: a local-LES case computes the same bits alone and in a mix;
* physics: on a 2D Taylor-Green vortex the local form decays the kinetic
  energy as the reference's finite-difference Smagorinsky does (same model,
  different strain estimate) and clearly faster than no LES.

MI355X only.
"""

import numpy as np
import pytest

from tests import _golden as G
from tests.test_gpu_parity import make_solver

pytestmark = pytest.mark.gpu

F64_TOL = 1e-10
F32_FIELD_TOL = 1e-5
F32_COEF_TOL = 2e-3

# golden geometries / states with the local closure (WALE cases become
# Smagorinsky: WALE has no local form); TRT + thermal + SEM covers the
# scalar-carrying kernels
CASES = [
    (2, "bgk_les_smag", {}),
    (2, "bgk_les_smag_vandriest", {}),
    (2, "trt_les_wale", {"les_model": "smagorinsky"}),
    (3, "bgk_les_smag", {}),
    (3, "trt_ibm_les", {}),
    (3, "trt_les_wale", {"les_model": "smagorinsky"}),
    (3, "mrt_les_wale_thermal_sem", {"collision": "trt", "les_model": "smagorinsky"}),
]


def _local_case(dim, name, over):
    case = dict(G.load(dim, name))
    kw = dict(case["kw"])
    kw.update(over)
    kw["les_form"] = "local"
    case["kw"] = kw
    return case


def _oracle_run(dim, case, steps):
    from tests.test_oracle import make_oracle
    o = make_oracle(dim, case)
    rows = []
    for step in range(1, steps + 1):
        plane = case["sem_planes"][step - 1] if o.inlet_mode == "sem" else None
        rows.append(o.coefficients(o.step(plane)))
    return o, np.array(rows)


@pytest.mark.parametrize("dim,name,over", CASES, ids=[f"{d}d-{n}" for d, n, _ in CASES])
@pytest.mark.parametrize("storage", ["f64", "f32"])
def test_local_les_matches_the_oracle(dim, name, over, storage):
    print(f"[gpu] local LES {dim}d {name} {storage}", flush=True)
    case = _local_case(dim, name, over)
    steps = G.SNAP2[-1] if dim == 2 else G.SNAP3[-1]
    o, ref = _oracle_run(dim, case, steps)
    s = make_solver(dim, case, storage)
    rows = s._advance(steps)
    f_tol, c_tol = (F64_TOL, F64_TOL) if storage == "f64" else (F32_FIELD_TOL, F32_COEF_TOL)
    assert G.rel_err(s.f, o.f_ref()) < f_tol, G.rel_err(s.f, o.f_ref())
    if o.g is not None:
        assert G.rel_err(s.g, o.g_ref()) < f_tol
    scale = max(np.max(np.abs(ref[:, 0])), 1e-3)
    assert np.max(np.abs(rows[:, 0] - ref[:, 0])) / scale < c_tol
    # the closure changed the state: the same case without LES differs
    if storage == "f64":
        kw = dict(case["kw"])
        kw["les"] = False
        plain = make_solver(dim, dict(case, kw=kw), storage)
        plain._advance(steps)
        assert G.rel_err(plain.f, s.f) > 1e-9


@pytest.mark.parametrize("dim,name", [(2, "bgk_les_smag"), (3, "trt_ibm_les")])
def test_zero_constant_reduces_to_the_plain_step(dim, name):
    print(f"[gpu] local LES C_s = 0 {dim}d {name}", flush=True)
    case = _local_case(dim, name, {"les_cs": 0.0})
    steps = G.SNAP2[-1] if dim == 2 else G.SNAP3[-1]
    s = make_solver(dim, case, "f64")
    rows = s._advance(steps)
    kw = dict(case["kw"])
    kw["les"] = False
    p = make_solver(dim, dict(case, kw=kw), "f64")
    rows_p = p._advance(steps)
    assert G.rel_err(s.f, p.f) < 1e-13
    # drag sums link forces with cancellation: compare against its scale
    np.testing.assert_allclose(rows[:, 0], rows_p[:, 0], rtol=0, atol=1e-12 * np.max(np.abs(rows_p[:, 0])))


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_local_les_batches_bit_for_bit(storage):
    from aero_amd.batch import Batch
    print(f"[gpu] local LES batch {storage}", flush=True)
    mix = [(2, "bgk_les_smag", {}, True), (2, "bgk_bb_conv_slip", {}, False),
           (2, "trt_les_wale", {"les_model": "smagorinsky"}, True), (3, "trt_ibm_les", {}, True),
           (3, "trt_bouzidi_conv_slip", {}, False), (2, "bgk_les_smag_vandriest", {}, True)]

    def make_all():
        out = []
        for dim, name, over, local in mix:
            case = _local_case(dim, name, over) if local else G.load(dim, name)
            out.append(make_solver(dim, case, storage))
        return out

    steps = 300
    alone = make_all()
    rows_alone = [s._advance(steps) for s in alone]
    mixed = make_all()
    rows_mix = Batch().advance(mixed, steps, record=False)
    for k, (a, m) in enumerate(zip(alone, mixed)):
        np.testing.assert_array_equal(rows_alone[k], rows_mix[k], err_msg=f"case {k}")
        np.testing.assert_array_equal(a.f, m.f, err_msg=f"case {k}")


def taylor_green_velocity(n=64, U=0.04):
    """A 2D Taylor-Green vortex phased so that its strain vanishes on the
    rows and columns x = 0 and y = 0 (and is small on the last ones): the
    reference's finite-difference model leaves those edge cells without eddy
    viscosity, which would otherwise dominate the comparison."""
    k = 2.0 * np.pi / n
    y, x = np.mgrid[0:n, 0:n].astype(np.float64)
    return U * np.cos(k * x) * np.sin(k * y), -U * np.sin(k * x) * np.cos(k * y)


def _taylor_green(les, form, n=64, U=0.04, omega=1.99, cs=0.3, steps=600):
    import aero_amd
    from aero_amd.lattice import compute_feq, compute_macroscopic
    ux, uy = taylor_green_velocity(n, U)
    s = aero_amd.Solver(Ny=n, Nx=n, solid=np.zeros((n, n), bool), omega=omega, u0=U, D=1.0, inlet_bc="none",
                        outlet_bc="none", wall_bc="none", les=les, les_cs=cs, les_form=form)
    s.f = compute_feq(np.ones((n, n)), ux, uy)
    s._advance(steps)
    _, vx, vy = compute_macroscopic(np.asarray(s.f))
    return float(np.sum(vx * vx + vy * vy)) / float(np.sum(ux * ux + uy * uy))


def test_local_form_decays_a_taylor_green_vortex_as_the_finite_difference_form():
    print("[gpu] local LES Taylor-Green", flush=True)
    off = _taylor_green(False, "fd")
    fd = _taylor_green(True, "fd")
    loc = _taylor_green(True, "local")
    rate_off, rate_fd, rate_loc = -np.log(off), -np.log(fd), -np.log(loc)
    assert rate_fd > 1.2 * rate_off, (rate_off, rate_fd)  # the eddy viscosity matters here
    assert abs(rate_loc - rate_fd) < 0.02 * rate_fd, (rate_fd, rate_loc)
