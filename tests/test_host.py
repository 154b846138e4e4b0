"""Host-side pieces of the product (CPU only): geometry, links, inflow
generators, diagnostics and the solver's argument checking.

Each is compared with the reference's recorded outputs (golden fixtures) or,
where the reference's own tests fix a behaviour, with the same known-answer
inputs those tests use (tests/test_diagnostics_stationarity.py,
tests/test_phase2.py:424-453, tests/test_observables.py of the reference).
"""

import math

import numpy as np
import pytest

from aero_amd import diagnostics as dg
from aero_amd.geometry import Box, Cylinder, Cylinder3D, Polygon, Rectangle, Sphere
from aero_amd.inflow import SEMInlet, SyntheticInflow2D, SyntheticInflow3D
from aero_amd.links import build_surface_links, build_surface_links_3d
from oracle import oracle as orc
from tests import _golden as G


# ---------------------------------------------------------------- geometry
def test_geometry_masks_and_sdf_match_fixtures():
    shapes2 = {"cyl": Cylinder(radius=5.0, cx_frac=0.3, cy_frac=0.5),
               "rect": Rectangle(width=5.0, height=7.0, cx_frac=0.3, cy_frac=0.45),
               "poly": Polygon([(0.25, 0.35), (0.33, 0.42), (0.30, 0.62), (0.22, 0.60), (0.19, 0.45)])}
    by_case = {"bgk_bb_conv_slip": "cyl", "trt316_bb_zg_noslip": "rect", "bgk_polygon": "poly"}
    for name, shape in by_case.items():
        case = G.load(2, name)
        g = shapes2[shape]
        np.testing.assert_array_equal(g.mark_solid(20, 40), case["solid"])
        if "phi" in case:
            np.testing.assert_array_equal(g.sdf_field(20, 40), case["phi"])
    shapes3 = {"bgk_bouzidi_conv_slip": Sphere(radius=2.3, cx_frac=0.3, cy_frac=0.5, cz_frac=0.5),
               "mrt_bb_conv_sponge": Box(width=3.0, height=4.0, depth=3.0, cx_frac=0.3, cy_frac=0.45),
               "bgk_periodic_bodyforce": Cylinder3D(radius=2.0, length=4.0, cx_frac=0.4)}
    for name, g in shapes3.items():
        case = G.load(3, name)
        np.testing.assert_array_equal(g.mark_solid(6, 8, 16), case["solid"])
        if "phi" in case:
            np.testing.assert_array_equal(g.sdf_field(6, 8, 16), case["phi"])


# ------------------------------------------------------------------- links
def test_links_match_reference_fixtures():
    with np.load(G.GOLDEN / "links_c1_c2.npz") as d:
        g = Cylinder(radius=20.0, cx_frac=0.25, cy_frac=0.5)
        links, q = build_surface_links(g.mark_solid(200, 400), g.sdf_field(200, 400))
        np.testing.assert_array_equal(links, d["c1_links"])
        np.testing.assert_array_equal(q, d["c1_q"])
        g3 = Sphere(radius=8.0, cx_frac=0.25)
        links3, q3 = build_surface_links_3d(g3.mark_solid(64, 64, 128), g3.sdf_field(64, 64, 128))
        np.testing.assert_array_equal(links3, d["c2_links"])
        np.testing.assert_array_equal(q3, d["c2_q"])
    for dim in (2, 3):
        for name in (G.CASES_2D if dim == 2 else G.CASES_3D):
            case = G.load(dim, name)
            fn = build_surface_links if dim == 2 else build_surface_links_3d
            links, _ = fn(case["solid"])
            np.testing.assert_array_equal(links, case["links"])


@pytest.mark.parametrize("seed", range(6))
def test_links_brute_force(seed):
    """Against the oracle's scan (itself checked by brute force); solids touching
    the domain edges exercise the clipped neighbour index."""
    rng = np.random.default_rng(seed)
    sol3 = rng.random((5, 6, 7)) < 0.25
    np.testing.assert_array_equal(build_surface_links_3d(sol3)[0], orc.build_links(sol3)[0])
    sol2 = rng.random((6, 9)) < 0.25
    np.testing.assert_array_equal(build_surface_links(sol2)[0], orc.build_links(sol2)[0])


# ----------------------------------------------------------------- inflow
def test_sem_matches_reference_planes():
    with np.load(G.GOLDEN / "sem_inlet.npz") as d:
        sem = SEMInlet(Ny=20, Nz=16, u0=0.05, Tu=0.05, L_int=4.0, N_eddies=30, rng_seed=0)
        planes = [np.stack(sem.fluctuation())]
        for _ in range(5):
            sem.step()
            planes.append(np.stack(sem.fluctuation()))
        np.testing.assert_allclose(np.array(planes), d["planes"], rtol=0, atol=1e-15)
        np.testing.assert_array_equal(sem.xk, d["xk"])
        np.testing.assert_array_equal(sem.eps, d["eps"])


def test_sem_solver_planes_match_reference_step_planes():
    """The per-step planes the 3D SEM case fed its steps (solver3d.py:434-437)."""
    case = G.load(3, "bgk_sem")
    kw = case["kw"]
    sem = SEMInlet(8, 6, float(case["u0"]), kw["sem_Tu"], kw["sem_L_int"], kw["sem_N"])
    for k in range(len(case["sem_planes"])):
        sem.step()
        np.testing.assert_allclose(np.stack(sem.fluctuation()), case["sem_planes"][k], rtol=0, atol=1e-15)


def test_sem_divergence_free_and_energy():
    """Known properties the reference tests (tests/test_sem_inlet.py:46-70)."""
    sem = SEMInlet(Ny=32, Nz=32, u0=0.05, Tu=0.1, L_int=8.0, N_eddies=300, rng_seed=99)
    _, v, w = sem.fluctuation()
    vh, wh = np.fft.fft2(v), np.fft.fft2(w)
    kz = (np.fft.fftfreq(32) * 2 * np.pi)[:, None]
    ky = (np.fft.fftfreq(32) * 2 * np.pi)[None, :]
    div = ky * vh + kz * wh
    assert np.linalg.norm(div) / (np.linalg.norm(vh) + np.linalg.norm(wh)) < 1e-10
    assert SEMInlet(16, 16, 0.05, 0.0, 5.0, 50).fluctuation()[0].max() == 0.0


def test_synthetic_inflow_statistics():
    g2 = SyntheticInflow2D(64, 0.05, 0.03)
    for _ in range(5):
        uy = g2.next()
    assert abs(float(np.mean(uy))) < 1e-15
    assert abs(float(np.sqrt(np.mean(uy ** 2))) - 0.03 * 0.05) < 1e-12
    g3 = SyntheticInflow3D(8, 12, 0.05, 0.04)
    uy, uz = g3.next()
    assert abs(float(np.sqrt(np.mean(uy ** 2 + uz ** 2))) - 0.04 * 0.05) < 1e-12


# ------------------------------------------------------------ diagnostics
def test_convergence_and_strouhal_known_answers():
    hist = [1.0 + 1e-4 * math.sin(i) for i in range(600)]
    r = dg.analyze_convergence(hist, window=500, tol=0.01)
    assert r.converged and r.ratio < r.threshold
    D, u0 = 20.0, 0.05
    t = np.arange(8192.0)
    freq = 4 / t.size
    rep = dg.analyze_strouhal(list(0.2 * np.sin(2 * math.pi * freq * t)), D=D, u0=u0, window=8192)
    assert abs(rep.strouhal - freq * D / u0) / (freq * D / u0) < 0.02
    st = 0.164
    cl = 0.3 * np.sin(2 * math.pi * (st * u0 / 40.0) * np.arange(5000.0))
    assert abs(dg.compute_strouhal(list(cl), D=40.0, u0=u0) - st) / st < 0.05
    assert dg.compute_strouhal([0.1] * 100, D=40.0, u0=0.05) is None
    assert dg.compute_strouhal([0.0] * 2000, D=40.0, u0=0.05) is None


def test_stationarity_requires_both_metrics():
    D, u0 = 20.0, 0.05
    freq = 0.17 * u0 / D
    t = np.arange(5000.0)
    cd = list(1.5 + 1e-4 * np.sin(0.01 * t))
    cl = list(0.15 * np.sin(2 * math.pi * freq * t))
    cdr, str_, stop = dg.detect_statistical_stationarity(cd, cl, D=D, u0=u0, convergence_window=2000,
                                                         strouhal_window=2048)
    assert cdr.converged and str_.strouhal is not None and stop
    # lift whose frequency jumps from bin 10 to bin 14 (of 1024) between the two half windows
    f1, f2 = 10 / 1024, 14 / 1024
    drifting = list(np.where(t < 5000 - 1024, 0.15 * np.sin(2 * math.pi * f1 * t),
                             0.15 * np.sin(2 * math.pi * f2 * t)))
    _, st2, stop2 = dg.detect_statistical_stationarity(cd, drifting, D=D, u0=u0, convergence_window=2000,
                                                       strouhal_window=2048)
    assert not stop2 and st2.relative_drift is not None and st2.relative_drift > 0.05


def test_coefficient_spectrum_and_validation():
    vals = [0.2 * math.sin(2 * math.pi * 0.025 * i) for i in range(512)]
    rep = dg.coefficient_spectrum(vals, discard_fraction=0.0, min_samples=128)
    assert abs(rep["dominant_frequency"] - 0.025) < 0.005
    assert dg.coefficient_spectrum([1.0] * 50) is None
    msgs = dg.validate_parameters(tau=0.5, u0=0.1, D=90.0, Ny=200)
    assert msgs[0].startswith("ERROR") and any("Ma=" in m for m in msgs) and any("blockage" in m for m in msgs)


def test_check_stability_semantics():
    with pytest.raises(RuntimeError, match="NaN"):
        dg.check_stability_stats(np.array([1.0, 0.0, 0.0, 1.0, 1.0]), 3)
    with pytest.raises(RuntimeError, match="Non-positive"):
        dg.check_stability_stats(np.array([0.0, 1.0, 0.0, 1.0, 1.0]), 3)
    with pytest.warns(RuntimeWarning, match="Max velocity"):
        dg.check_stability_stats(np.array([0.0, 0.0, 0.5, 10.0, 10.0]), 3)
    with pytest.warns(RuntimeWarning, match="Mean density"):
        dg.check_stability_stats(np.array([0.0, 0.0, 0.0, 12.0, 10.0]), 3)


# --------------------------------------------------------- argument checks
def test_solver_argument_errors_match_reference():
    from aero_amd import Solver, Solver3D
    solid = np.zeros((8, 16), bool)
    with pytest.raises(ValueError, match="Unknown collision"):
        Solver(8, 16, solid, 1.0, 0.05, 4.0, collision="xyz")
    with pytest.raises(ValueError, match="Unknown backend"):
        Solver(8, 16, solid, 1.0, 0.05, 4.0, backend="cuda")
    with pytest.raises(ValueError, match="les_model"):
        Solver(8, 16, solid, 1.0, 0.05, 4.0, les_model="dynamic")
    with pytest.raises(ValueError, match="streamwise_bc"):
        Solver3D(4, 8, 16, np.zeros((4, 8, 16), bool), 1.0, 0.05, 4.0, streamwise_bc="inflow")


def _header_table(lattice_struct, name):
    """A constexpr integer / bool table of a lattice struct in lbm_device.h."""
    import pathlib
    import re
    src = (pathlib.Path(__file__).resolve().parents[1] / "aero_amd" / "csrc" / "lbm_device.h").read_text()
    seg = src[src.index(f"struct {lattice_struct} {{"):]
    j = seg.index(f" {name}[")
    body = seg[seg.index("=", j):seg.index("};", j)]
    return np.array([int(v) for v in re.findall(r"-?\d+", body)])


def test_mrt_tables_the_kernels_rely_on():
    """The MRT collision uses (a) D3Q19::PINV_NZ, the entries of the
    pseudo-inverse of the reference's rank-16 M19 that are nonzero beyond
    rounding, (b) that every row of M9 / M19 is even or odd under
    i <-> opp(i), with the pseudo-inverse keeping that symmetry, and (c) that
    M9's rows are orthogonal, so inv(M9) = M9^T diag(M9 M9^T)^-1."""
    from aero_amd import lattice
    P = lattice.M19_inv
    nz = _header_table("D3Q19", "PINV_NZ").reshape(19, 19).astype(bool)
    np.testing.assert_array_equal(nz, np.abs(P) > 1e-12)
    assert np.abs(P[~nz]).max() < 1e-14 and np.abs(P[nz]).min() > 1e-3
    np.testing.assert_array_equal(_header_table("D3Q19", "M").reshape(19, 19), lattice.M19)
    np.testing.assert_array_equal(_header_table("D2Q9", "M").reshape(9, 9), lattice.M9)
    for M, opp, Pi in ((lattice.M9, lattice.OPP, lattice.M9_inv), (lattice.M19, lattice.OPP3, P)):
        even = np.array([np.array_equal(r, r[opp]) for r in M])
        odd = np.array([np.array_equal(r, -r[opp]) for r in M])
        assert (even ^ odd).all()
        sign = np.where(even, 1.0, -1.0)
        assert np.abs(Pi[opp] - Pi * sign[None, :]).max() < 1e-13
    G9 = lattice.M9 @ lattice.M9.T
    assert np.abs(G9 - np.diag(np.diag(G9))).max() == 0
    np.testing.assert_allclose(lattice.M9.T / np.diag(G9), lattice.M9_inv, atol=1e-15)


def test_sem_state_round_trip_continues_the_same_planes():
    """SEMInlet.state() through JSON (the checkpoint's sem_state_json) into a
    fresh generator: the following steps regenerate the same eddies and give
    the same planes as the uninterrupted generator, bit for bit."""
    import json
    a = SEMInlet(Ny=24, Nz=18, u0=0.05, Tu=0.05, L_int=3.0, N_eddies=40, rng_seed=5)
    for _ in range(40):  # past several regenerations (x runs across 6 sigma in 360 steps at u0 = 0.05)
        a.step()
    b = SEMInlet(Ny=24, Nz=18, u0=0.05, Tu=0.05, L_int=3.0, N_eddies=40, rng_seed=123)
    b.set_state(json.loads(json.dumps(a.state())))
    for _ in range(400):
        a.step()
        b.step()
    np.testing.assert_array_equal(a.xk, b.xk)
    np.testing.assert_array_equal(a.eps, b.eps)
    for pa, pb in zip(a.fluctuation(), b.fluctuation()):
        np.testing.assert_array_equal(pa, pb)
    assert not np.array_equal(a.yk, SEMInlet(24, 18, 0.05, 0.05, 3.0, 40, rng_seed=5).yk)  # eddies did regenerate


def test_synthetic_inflow_state_round_trip():
    """SyntheticInflow2D/3D.state() through JSON into a generator with another
    seed: the next planes are those of the uninterrupted generator."""
    import json
    for make in (lambda seed: SyntheticInflow2D(40, 0.05, 0.05, seed),
                 lambda seed: SyntheticInflow3D(12, 20, 0.05, 0.05, seed)):
        a = make(7)
        for _ in range(9):
            a.next()
        b = make(99)
        b.set_state(json.loads(json.dumps(a.state())))
        for _ in range(5):
            pa, pb = a.next(), b.next()
            for x, y in zip(pa if isinstance(pa, tuple) else (pa,), pb if isinstance(pb, tuple) else (pb,)):
                np.testing.assert_array_equal(x, y)


def test_state_arrays_write_through_while_current():
    """DeviceSolver.f / .g hand out LiveField copies: an element assignment
    into one, or into a view of it, writes the array back while the state is
    the one it was read from; after a step (a new state generation) it is a
    detached copy, and arrays computed from it are plain copies."""
    from aero_amd._base import _live

    class Owner:  # the DeviceSolver attribute protocol, without a device
        _state_gen = 0

        def __init__(self):
            self.data, self.writes = np.zeros((3, 4, 5)), 0

        @property
        def f(self):
            return _live(self, "f", self.data.copy())

        @f.setter
        def f(self, v):
            self.data, self.writes = np.array(v), self.writes + 1
            self._state_gen += 1

    o = Owner()
    o.f[:, 1, 2] += 0.5  # the reference scripts' in-place edit
    assert o.data[0, 1, 2] == 0.5 and o.writes == 1
    a = o.f
    a[0][3, 4] = 7.0  # through a view
    a[1, 0, 0] = 2.0  # the same copy stays current after its own write-back
    assert o.data[0, 3, 4] == 7.0 and o.data[1, 0, 0] == 2.0 and o.writes == 3
    o._state_gen += 1  # a step
    a[2, 0, 0] = 9.0
    assert o.data[2, 0, 0] == 0.0 and o.writes == 3
    b = o.f * 2
    b[0, 0, 0] = 5.0
    assert o.writes == 3 and type(np.asarray(o.f)) is np.ndarray
    assert type(b) is np.ndarray and type(o.f.sum(axis=0)) is np.ndarray and isinstance(o.f.sum(), float)


@pytest.mark.parametrize("dim", [2, 3])
def test_lattice_tables_and_equilibrium(dim):
    """tests/test_d2q9.py:21-110 (and the same properties for D3Q19): weights
    sum to 1, opposite directions are mutual and reverse e, c_s^2 = 1/3, the
    equilibrium carries the density and momentum it is built from, stays
    positive at low Mach number, and compute_macroscopic inverts it."""
    from aero_amd import lattice as lat
    E, W, OPP = (lat.E, lat.W, lat.OPP) if dim == 2 else (lat.E3, lat.W3, lat.OPP3)
    assert abs(W.sum() - 1.0) < 1e-14 and abs(lat.CS2 - 1.0 / 3.0) < 1e-15
    for i in range(len(W)):
        assert OPP[OPP[i]] == i and (E[OPP[i]] == -E[i]).all()
    rng = np.random.default_rng(4)
    shape = (10, 20) if dim == 2 else (6, 8, 10)
    rho = 1.0 + 0.01 * rng.standard_normal(shape)
    u = [0.05 + 0.01 * rng.standard_normal(shape)] + [0.01 * rng.standard_normal(shape) for _ in range(dim - 1)]
    feq = lat.compute_feq(rho, *u) if dim == 2 else lat.compute_feq_3d(rho, *u)
    np.testing.assert_allclose(feq.sum(axis=0), rho, rtol=0, atol=1e-13)
    for d in range(dim):
        np.testing.assert_allclose(np.tensordot(E[:, d].astype(float), feq, axes=1), rho * u[d], rtol=0, atol=1e-14)
    assert feq.min() >= 0.0
    mac = lat.compute_macroscopic(feq) if dim == 2 else lat.compute_macroscopic_3d(feq)
    np.testing.assert_allclose(mac[0], rho, rtol=0, atol=1e-13)
    for d in range(dim):
        np.testing.assert_allclose(mac[1 + d], u[d], rtol=0, atol=1e-13)


def test_synthetic_inflow_chunks_equal_single_steps():
    """next_n(n) draws a chunk's noise in one block; its planes must be
    bitwise those of n successive next() calls, also across chunk borders."""
    a, b = SyntheticInflow2D(37, 0.05, 0.03, 11), SyntheticInflow2D(37, 0.05, 0.03, 11)
    ref = np.array([a.next() for _ in range(30)])
    np.testing.assert_array_equal(np.concatenate([b.next_n(7), b.next_n(1), b.next_n(22)]), ref)
    a3, b3 = SyntheticInflow3D(9, 14, 0.05, 0.03, 12), SyntheticInflow3D(9, 14, 0.05, 0.03, 12)
    steps = [a3.next() for _ in range(20)]
    y1, z1 = b3.next_n(13)
    y2, z2 = b3.next_n(7)
    np.testing.assert_array_equal(np.concatenate([y1, y2]), np.array([s[0] for s in steps]))
    np.testing.assert_array_equal(np.concatenate([z1, z2]), np.array([s[1] for s in steps]))


def test_les_form_selects_the_device_mode():
    """les_form "fd" keeps the reference's models; "local" is Smagorinsky
    only (WALE needs the rotation part of the velocity gradient)."""
    from aero_amd import _abi
    from aero_amd._base import les_code
    assert les_code(False, "smagorinsky", "local") == _abi.LES_OFF
    assert les_code(True, "smagorinsky", "fd") == _abi.LES_SMAGORINSKY
    assert les_code(True, "wale", "fd") == _abi.LES_WALE
    assert les_code(True, "smagorinsky", "local") == _abi.LES_SMAGORINSKY_LOCAL == 3
    with pytest.raises(ValueError, match="Smagorinsky only"):
        les_code(True, "wale", "local")
    with pytest.raises(ValueError, match="les_form"):
        les_code(True, "smagorinsky", "lattice")
