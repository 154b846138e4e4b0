"""Pins the CPU oracle: golden vectors from the reference + closed forms.

These tests run on the CPU only.  They check the checker: a C restatement
that drifts from the reference (a dropped term, a swapped direction, a wrong
boundary order) fails one of the composite-step cases below, each of which
exercises a different collision/boundary combination.
"""

import ctypes

import numpy as np
import pytest

from oracle import oracle as orc
from tests import _golden as G


@pytest.fixture(scope="module", autouse=True)
def _built():
    orc.build_library()


# ----------------------------------------------------------------- lattices
@pytest.mark.parametrize("E,W,opp", [(orc.E9, orc.W9, orc.OPP9), (orc.E19, orc.W19, orc.OPP19)])
def test_lattice_moments(E, W, opp):
    """Isotropy: sum w = 1, sum w e = 0, sum w e_a e_b = delta/3, opp reverses e."""
    assert abs(W.sum() - 1.0) < 1e-15
    np.testing.assert_allclose(W @ E, 0.0, atol=1e-15)
    second = np.einsum("i,ia,ib->ab", W, E, E)
    np.testing.assert_allclose(second, np.eye(E.shape[1]) / 3.0, atol=1e-15)
    np.testing.assert_array_equal(E[opp], -E)


def test_feq_moments():
    rng = np.random.default_rng(0)
    rho = 1.0 + 0.01 * rng.standard_normal((5, 6))
    ux, uy = 0.05 * rng.standard_normal((2, 5, 6))
    f = orc.feq(orc.E9, orc.W9, rho, ux, uy)
    np.testing.assert_allclose(f.sum(0), rho, rtol=1e-14)
    np.testing.assert_allclose(np.einsum("i,iyx->yx", orc.E9[:, 0], f), rho * ux, atol=1e-15)
    np.testing.assert_allclose(np.einsum("i,iyx->yx", orc.E9[:, 1], f), rho * uy, atol=1e-15)


def test_mrt_inverses_match_reference():
    with np.load(G.GOLDEN / "mrt_inverses.npz") as d:
        np.testing.assert_allclose(orc.M9_INV, d["M9_inv"], rtol=0, atol=1e-15)
        np.testing.assert_allclose(orc.M19_INV, d["M19_inv"], rtol=0, atol=1e-15)
    # the reference's D3Q19 basis is rank 16 (its ghost rows are dependent)
    assert np.linalg.matrix_rank(orc.M19) == 16
    assert np.linalg.matrix_rank(orc.M9) == 9


# -------------------------------------------------------------------- links
def test_links_match_reference_c1_c2():
    with np.load(G.GOLDEN / "links_c1_c2.npz") as d:
        solid = d["c1_solid"]
        yy, xx = np.mgrid[0:200, 0:400].astype(np.float64)
        phi = np.sqrt((xx + 0.5 - 100.0) ** 2 + (yy + 0.5 - 100.0) ** 2) - 20.0
        links, q = orc.build_links(solid, phi)
        np.testing.assert_array_equal(links, d["c1_links"])
        np.testing.assert_array_equal(q, d["c1_q"])
        zz, yy, xx = np.mgrid[0:64, 0:64, 0:128].astype(np.float64)
        solid3 = (xx - 32.0) ** 2 + (yy - 32.0) ** 2 + (zz - 32.0) ** 2 <= 64.0
        assert int(solid3.sum()) == int(d["c2_solid_count"])
        phi3 = np.sqrt((xx + 0.5 - 32.0) ** 2 + (yy + 0.5 - 32.0) ** 2 + (zz + 0.5 - 32.0) ** 2) - 8.0
        links3, q3 = orc.build_links(solid3, phi3)
        np.testing.assert_array_equal(links3, d["c2_links"])
        np.testing.assert_array_equal(q3, d["c2_q"])


def test_links_brute_force_small():
    rng = np.random.default_rng(3)
    solid = rng.random((5, 6, 7)) < 0.3
    links, _ = orc.build_links(solid)
    expect = []
    for i in range(1, 19):
        e = orc.E19[i]
        for z in range(5):
            for y in range(6):
                for x in range(7):
                    zn = min(max(z + e[2], 0), 4)
                    yn = min(max(y + e[1], 0), 5)
                    xn = min(max(x + e[0], 0), 6)
                    if not solid[z, y, x] and solid[zn, yn, xn]:
                        expect.append((i, z, y, x))
    np.testing.assert_array_equal(links, np.array(expect).reshape(-1, 4))


# --------------------------------------------------------------- composites
def make_oracle(dim, case):
    kw = dict(case["kw"])
    solid = case["solid"]
    use_phi = kw.get("bouzidi", False) or kw.get("van_driest", False) or kw.get("ibm_enabled", False)
    phi = case.get("phi") if use_phi else None
    args = dict(omega=float(case["omega"]), u0=float(case["u0"]), D=float(case["D"]), phi=phi,
                ibm=bool(kw.pop("ibm_enabled", False)))
    inlet_mode = "zouhe"
    if kw.pop("synthetic_inflow", False):
        inlet_mode = "field"
    kw.pop("synthetic_inflow_intensity", None)
    if kw.pop("sem_inlet", False):
        inlet_mode = "sem"
    for k in ("sem_Tu", "sem_L_int", "sem_N"):
        kw.pop(k, None)
    bf = (kw.pop("body_force_x", 0.0), kw.pop("body_force_y", 0.0), kw.pop("body_force_z", 0.0))
    s = orc.OracleSolver(solid, inlet_mode=inlet_mode, body_force=bf, **args, **kw)
    s.set_state(G.initial_f(case))
    return s


def run_and_compare(dim, name, f_tol=1e-10, c_tol=1e-9):
    case = G.load(dim, name)
    s = make_oracle(dim, case)
    assert abs(s.f.sum() - float(case["f0_sum"])) < 1e-10
    np.testing.assert_array_equal(s.links_ref, case["links"])
    np.testing.assert_array_equal(s.q_vals, case["q_vals"])
    syn = None
    if s.inlet_mode == "field":
        shape = (s.ny,) if dim == 2 else (s.nz, s.ny)
        syn = G.SyntheticInflow(shape, s.u0, float(case["kw"]["synthetic_inflow_intensity"]))
    snaps = G.SNAP2 if dim == 2 else G.SNAP3
    coeffs = []
    for step in range(1, max(snaps) + 1):
        plane = None
        if syn is not None:
            plane = syn.next()
        elif s.inlet_mode == "sem":
            plane = case["sem_planes"][step - 1]
        raw = s.step(plane)
        coeffs.append(s.coefficients(raw))
        if step in snaps:
            assert G.rel_err(s.f_ref(), case[f"f_{step}"]) < f_tol, (name, step)
            assert G.rel_err(s.prev_ref(), case[f"fprev_{step}"]) < f_tol, (name, step)
            if s.g is not None:
                assert G.rel_err(s.g_ref(), case[f"g_{step}"]) < f_tol, (name, step)
    c = np.array(coeffs)
    if dim == 2:
        cols = dict(Cd=0, Cl=1, Cm=5, Cd_p=6, Cd_v=7)
    else:
        cols = dict(Cd=0, Cly=1, Clz=2, Cmx=3, Cmy=4, Cmz=5, Cd_p=6, Cd_v=7)
    for key, col in cols.items():
        ref = case[f"{key}_hist"]
        scale = max(np.max(np.abs(ref)), 1e-3)
        assert np.max(np.abs(c[:, col] - ref)) / scale < c_tol, (name, key)
    prof = s.last["profile"]
    for k, ax in (("fx", 0), ("fy", 1)) + ((("fz", 2),) if dim == 3 else ()):
        np.testing.assert_allclose(prof[:, ax], case[f"profile_{k}"], rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("name", G.CASES_2D)
def test_oracle_2d_matches_reference(name):
    run_and_compare(2, name)


@pytest.mark.parametrize("name", G.CASES_3D)
def test_oracle_3d_matches_reference(name):
    run_and_compare(3, name)


def test_oracle_detects_planted_mistake():
    """A plausible slip (slip-wall pairs swapped for no-slip) must fail parity."""
    case = G.load(2, "bgk_bb_conv_slip")
    s = make_oracle(2, case)
    s.wall_bc = "noslip"
    s._wall_tables()
    for _ in range(G.SNAP2[-1]):
        s.step()
    assert G.rel_err(s.f_ref(), case[f"f_{G.SNAP2[-1]}"]) > 1e-8


# ------------------------------------------ labelled non-reference operators
def _noisy_state(solid, u0=0.05, seed=5):
    rng = np.random.default_rng(seed)
    sh = solid.shape
    if solid.ndim == 3:
        f = orc.feq(orc.E19, orc.W19, 1.0 + 0.01 * rng.standard_normal(sh), np.full(sh, u0),
                    0.01 * rng.standard_normal(sh), 0.01 * rng.standard_normal(sh))
    else:
        f = orc.feq(orc.E9, orc.W9, 1.0 + 0.01 * rng.standard_normal(sh), np.full(sh, u0),
                    0.01 * rng.standard_normal(sh))[:, None]
    return f * (1.0 + 0.02 * rng.standard_normal(f.shape))


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("omega", [1.9, 1.2])
def test_lambda_trt_reduces_to_bgk(dim, omega):
    """Standard TRT with Lambda = (tau - 1/2)^2 has tau- = tau, so both parts
    relax at omega: the operator must equal BGK exactly.  The reference's TRT
    rate formula with the same lambda does not, so a mix-up of the two fails."""
    solid = np.zeros((4, 5, 6) if dim == 3 else (5, 6), dtype=bool)
    solid[..., 2, 3] = True
    f = _noisy_state(solid)
    tau = 1.0 / omega
    lam = (tau - 0.5) ** 2
    kw = dict(omega=omega, u0=0.05, D=4.0)
    bgk = orc.OracleSolver(solid, collision="bgk", **kw).collide(f)
    trt = orc.OracleSolver(solid, collision="trt", trt_lambda=lam, trt_variant="lambda", **kw).collide(f)
    ref = orc.OracleSolver(solid, collision="trt", trt_lambda=lam, trt_variant="reference", **kw).collide(f)
    assert np.max(np.abs(trt - bgk)) < 1e-14
    assert np.max(np.abs(ref - bgk)) > 1e-6


def _poiseuille(q, g=1e-6, tau=0.8, ny=12, steps=6000, qshift=None):
    """Body-force-driven channel, periodic in x and z, solid rows y = 0 and
    y = ny-1, every wall link given the fraction q: the walls sit at
    y = 1 - q and y = ny - 2 + q."""
    nz, nx = 2, 3
    solid = np.zeros((nz, ny, nx), dtype=bool)
    solid[:, 0, :] = True
    solid[:, -1, :] = True
    links, _ = orc.build_links(solid)
    qv = np.full(len(links), q if qshift is None else qshift, dtype=np.float32)
    o = orc.OracleSolver(solid, omega=1.0 / tau, u0=0.0, D=1.0, wall_bc="none", streamwise_bc="periodic",
                         collision="trt", trt_lambda=3.0 / 16.0, trt_variant="lambda", bouzidi=True,
                         bouzidi_variant="linear", links=links, q_vals=qv, body_force=(g, 0.0, 0.0))
    for _ in range(steps):
        o.step()
    _, ux, _, _ = o.macroscopic(o.f_ref())
    y = np.arange(1, ny - 1, dtype=np.float64)
    yb, yt = 1.0 - q, ny - 2.0 + q
    nu = (tau - 0.5) / 3.0
    # the uniform body force enters as (1 - omega/2) times the Guo source with
    # no half-force velocity shift (ibm_guo.py:264-301), i.e. a net force
    # (1 - 1/(2 tau)) g per step
    g_eff = (1.0 - 0.5 / tau) * g
    exact = g_eff / (2.0 * nu) * (y - yb) * (yt - y)
    num = ux[0, 1:-1, 1]
    return np.max(np.abs(num - exact)) / np.max(exact)


@pytest.mark.parametrize("q", [0.2, 0.5, 0.8])
def test_linear_bouzidi_puts_the_wall_at_the_link_fraction(q):
    """Linear interpolated bounce-back places a plane wall at fraction q along
    each link: the steady Poiseuille profile between walls at y = 1 - q and
    y = ny - 2 + q is matched exactly at q = 1/2 (mid-link bounce-back with
    Lambda = 3/16) and within the scheme's second-order error (< 3% of the
    peak on this 10-cell channel) elsewhere, while walls put at the mirrored
    fraction 1 - q miss it by 20% or more."""
    err = _poiseuille(q)
    if q == 0.5:
        assert err < 1e-9
    else:
        assert err < 3e-2
        assert _poiseuille(q, qshift=1.0 - q) > 0.1


def test_ibm_forcing_closed_forms():
    """Guo forcing f_i += w_i e_i.F (ibm_guo.py:110-182) keeps each band cell's
    mass, adds F/3 to its momentum (sum_i w_i e_i e_i = I/3) with
    F = (1 - 1/(2 tau)) (0 - u), and leaves solid, far and phi <= 0 cells alone."""
    rng = np.random.default_rng(12)
    for Q, E, W in ((9, orc.E9, orc.W9), (19, orc.E19, orc.W19)):
        nz, ny, nx = (1, 5, 7) if Q == 9 else (3, 4, 5)
        shape = (Q, nz, ny, nx)
        rho = 1.0 + rng.uniform(-0.05, 0.05, shape[1:])
        vel = [rng.uniform(-0.1, 0.1, shape[1:]) for _ in range(E.shape[1])]
        f = orc.feq(E, W, rho, *vel)
        f += rng.uniform(-1e-3, 1e-3, shape)
        phi = rng.uniform(-1.0, 3.0, shape[1:])
        solid = (rng.random(shape[1:]) < 0.2).astype(np.uint8)
        tau = 0.8
        g = f.copy()
        ex, ey = np.ascontiguousarray(E[:, 0], np.int32), np.ascontiguousarray(E[:, 1], np.int32)
        ez = np.ascontiguousarray(E[:, 2], np.int32) if Q == 19 else np.zeros(9, np.int32)
        orc.lib().orc_ibm_guo(Q, nz, ny, nx, orc._p(g), orc._p(np.ascontiguousarray(phi)), orc._p(solid),
                              ctypes.c_double(tau), orc._p(ex), orc._p(ey), orc._p(ez),
                              orc._p(np.ascontiguousarray(W)))
        band = (solid == 0) & (phi > 0) & (phi <= 1.5)
        np.testing.assert_array_equal(g[:, ~band], f[:, ~band])
        r0 = f.sum(axis=0)
        np.testing.assert_allclose(g.sum(axis=0)[band], r0[band], rtol=0, atol=1e-15)
        for a in range(E.shape[1]):
            m0 = np.tensordot(E[:, a].astype(float), f, axes=(0, 0))
            m1 = np.tensordot(E[:, a].astype(float), g, axes=(0, 0))
            F = (1.0 - 0.5 / tau) * (0.0 - m0 / r0)
            np.testing.assert_allclose((m1 - m0)[band], F[band] / 3.0, rtol=1e-12, atol=1e-16)


def test_kernel_only_loop_matches_collide_copy_stream():
    """orc_kernel_only (the CPU baseline's timed kernel leg) performs exactly
    collide -> copy -> stream, as the composed step's first two stages do,
    and leaves the solver's own state unchanged."""
    from oracle import oracle as orc
    rng = np.random.default_rng(3)
    nz, ny, nx = 5, 6, 9
    solid = np.zeros((nz, ny, nx), dtype=bool)
    solid[2, 3, 4] = True
    o = orc.OracleSolver(solid, omega=1.3, u0=0.04, D=2.0, collision="trt", trt_lambda=0.1875)
    f0 = orc.feq(orc.E19, orc.W19, np.ones((nz, ny, nx)), np.full((nz, ny, nx), 0.04),
                 rng.uniform(-1e-3, 1e-3, (nz, ny, nx)), rng.uniform(-1e-3, 1e-3, (nz, ny, nx)))
    o.set_state(f0)
    expect = o.stream(o.collide(o.stream(o.collide(o.f))))
    Q, _, _, _ = o._dims()
    f = o.f.copy()
    a, b = np.empty_like(f), np.empty_like(f)
    import ctypes
    sec = ctypes.c_double(-1.0)
    orc.lib().orc_kernel_only(Q, nz, ny, nx, orc._p(f), orc._p(a), orc._p(b), orc._p(o.solid_u8), orc._p(o.ex),
                              orc._p(o.ey), orc._p(o.ez), orc._p(o.w), orc._p(o.opp), o.collision,
                              ctypes.c_double(o.omega), ctypes.c_double(o.trt_lambda), orc._p(o.M), orc._p(o.Minv),
                              orc._p(o.s), o.trt_variant, 2, ctypes.byref(sec))
    np.testing.assert_array_equal(f, expect)
    assert sec.value >= 0.0
    np.testing.assert_array_equal(o.f, f0)
    assert orc.kernel_only_seconds(o, 1) >= 0.0


# ---------------------------------------------- local Smagorinsky closure
def _ce_state(dim, rho, u, S, tau):
    """Populations feq(rho, u) + f^neq whose non-equilibrium stress is the
    Chapman-Enskog Pi^neq = -2 rho c_s^2 tau S: f^neq_i = w_i / (2 c_s^4)
    (e_i e_i - c_s^2 I) : Pi^neq (zero mass and momentum; its second moment is
    Pi^neq by the lattice's fourth-order isotropy)."""
    E, W = (orc.E19, orc.W19) if dim == 3 else (orc.E9, orc.W9)
    E = E.astype(np.float64)
    cs2 = 1.0 / 3.0
    pi = -2.0 * rho * cs2 * tau * S
    fe = orc.feq(E.astype(np.int32), W, np.array(rho), *u)
    fneq = np.array([W[i] / (2.0 * cs2 * cs2) * np.sum((np.outer(E[i], E[i]) - cs2 * np.eye(dim)) * pi)
                     for i in range(len(W))])
    return fe + fneq


@pytest.mark.parametrize("dim", [2, 3])
def test_local_smagorinsky_closure_returns_the_chapman_enskog_tau(dim):
    """For populations whose stress is that of strain S relaxed at tau*, the
    local closure must return tau* = tau0 + 3 C^2 |S| (|S| = sqrt(2 S:S)),
    the Smagorinsky viscosity nu0 + C^2 |S| written as a relaxation time.
    u != 0 and off-diagonal, non-traceless S make a dropped rho u u term, a
    lost factor 2 on the shear entries or a wrong constant fail; C = 0 must
    give tau0; solid cells keep omega."""
    rng = np.random.default_rng(11)
    omega, cs = 1.7, 0.17
    tau0 = 1.0 / omega
    shape = (1, 2, 3)  # z, y, x
    n = int(np.prod(shape))
    Q = 19 if dim == 3 else 9
    f = np.empty((Q, n))
    expect = np.empty(n)
    damp = rng.uniform(0.3, 1.0, n)
    for c in range(n):
        A = rng.uniform(-2e-3, 2e-3, (dim, dim))
        S = 0.5 * (A + A.T)
        rho = 1.0 + 0.05 * rng.standard_normal()
        u = tuple(rng.uniform(-0.08, 0.08, dim))
        smag = np.sqrt(2.0 * np.sum(S * S))
        C = cs * damp[c]
        tau_star = tau0 + 3.0 * C * C * smag
        f[:, c] = _ce_state(dim, rho, u, S, tau_star)
        expect[c] = tau_star
    solid = np.zeros(shape if dim == 3 else shape[1:], dtype=bool)
    solid.reshape(-1)[4] = True
    phi = -25.0 * np.log(1.0 - np.sqrt(damp))  # (1 - exp(-|phi| / 25))^2 == damp
    o = orc.OracleSolver(solid, omega=omega, u0=0.05, D=2.0, les=True, les_cs=cs, les_form="local",
                         van_driest=True, phi=phi.reshape(solid.shape))
    om = o.omega_field(np.ascontiguousarray(f.reshape((Q,) + shape))).reshape(-1)
    fluid = ~solid.reshape(-1)
    np.testing.assert_allclose(1.0 / om[fluid], expect[fluid], rtol=1e-12, atol=0)
    assert om[4] == omega
    assert np.max(np.abs(expect[fluid] - tau0)) > 1e-4  # the eddy viscosity is not negligible here
    o0 = orc.OracleSolver(solid, omega=omega, u0=0.05, D=2.0, les=True, les_cs=0.0, les_form="local")
    om0 = o0.omega_field(np.ascontiguousarray(f.reshape((Q,) + shape))).reshape(-1)
    np.testing.assert_allclose(om0, omega, rtol=1e-15)


def test_local_smagorinsky_closure_detects_a_planted_mistake():
    """The same identity fails for a closure that forgets rho u u in Pi^eq."""
    rng = np.random.default_rng(2)
    S = np.array([[1e-3, 4e-4], [4e-4, -2e-4]])
    rho, u, tau0, C = 1.02, (0.07, -0.03), 1.0 / 1.7, 0.17
    tau_star = tau0 + 3.0 * C * C * np.sqrt(2.0 * np.sum(S * S))
    f = _ce_state(2, rho, u, S, tau_star)
    E = orc.E9.astype(np.float64)
    P = np.einsum("ia,ib,i->ab", E, E, f)
    wrong = P - rho / 3.0 * np.eye(2)  # the planted slip
    qn = np.sqrt(2.0 * np.sum(wrong * wrong))
    tau_wrong = 0.5 * (tau0 + np.sqrt(tau0 ** 2 + 18.0 * C * C * qn / rho))
    assert abs(tau_wrong - tau_star) > 1e-3 * tau_star
    del rng


def test_local_and_finite_difference_smagorinsky_decay_a_vortex_alike():
    """Both closures are the Smagorinsky model with different strain
    estimates: on a smooth 2D Taylor-Green vortex (64^2, tau = 0.5025,
    C_s = 0.3, 600 steps) the kinetic-energy decay rates agree within 2%,
    while the eddy viscosity raises the rate by more than 40% over no LES."""
    from tests.test_gpu_les_local import taylor_green_velocity
    n, U = 64, 0.04
    ux, uy = taylor_green_velocity(n, U)

    def rate(les, form):
        o = orc.OracleSolver(np.zeros((n, n), bool), omega=1.99, u0=U, D=1.0, inlet_bc="none", outlet_bc="none",
                             wall_bc="none", les=les, les_cs=0.3, les_form=form)
        o.set_state(orc.feq(orc.E9, orc.W9, np.ones((n, n)), ux, uy))
        for _ in range(600):
            o.step()
        f = o.f_ref()
        rho = f.sum(0)
        vx, vy = (f * orc.E9[:, 0, None, None]).sum(0) / rho, (f * orc.E9[:, 1, None, None]).sum(0) / rho
        return -np.log(np.sum(vx * vx + vy * vy) / np.sum(ux * ux + uy * uy))

    off, fd, loc = rate(False, "fd"), rate(True, "fd"), rate(True, "local")
    assert fd > 1.4 * off and loc > 1.4 * off
    assert abs(loc - fd) < 0.02 * fd
