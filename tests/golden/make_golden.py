"""Write the golden input/output vectors for the LBM step path.

Run once, in a container that holds the reference checkout at
/root/reference (it is never needed at test time):

    python tests/golden/make_golden.py          # every fixture
    python tests/golden/make_golden.py ibm      # only the step cases named *ibm*
    python tests/golden/make_golden.py ckpt     # reference-written checkpoints
    python tests/golden/make_golden.py signatures
    python tests/golden/make_golden.py c1_long_run   # 20000 reference steps (~20 min)

The reference is pure Python.  Numba is absent here, so every call takes the
reference's NumPy fallback path, which its own tests hold equal to the Numba
path to rtol 1e-10 (tests/test_phase3.py:215-248, tests/test_phase4.py:427-438).

Each case builds the reference solver, overwrites ``f`` with a seeded
equilibrium-plus-noise state (the reference itself starts from plain
equilibrium, aero/lbm/solver.py:218-223), runs ``_step()`` and records f, g,
f_outlet_prev and every per-step observable.  Only data is written: inputs,
parameters and expected outputs, as compressed ``.npz`` files beside this
script.
"""

from __future__ import annotations

import json
import pathlib
import sys

import numpy as np

REF = pathlib.Path("/root/reference")
OUT = pathlib.Path(__file__).resolve().parent


def _import_reference():
    sys.path.insert(0, str(REF))
    # aero/lbm/kernels3d_trt.py:8 imports _trt_s_minus_numba, which
    # aero/lbm/kernels_trt.py:222 defines only when numba is installed.  The
    # NumPy TRT path (kernels3d_trt.py:19-56) never calls it, so a placeholder
    # that raises is enough for the module to import.
    import aero.lbm.kernels_trt as _kt
    if not hasattr(_kt, "_trt_s_minus_numba"):
        def _absent(*_a, **_k):
            raise RuntimeError("numba-only helper called on the NumPy path")
        _kt._trt_s_minus_numba = _absent
    from aero.lbm.solver import Solver
    from aero.lbm.solver3d import Solver3D
    from aero.lbm.d2q9 import compute_feq
    from aero.lbm.d3q19 import compute_feq_3d
    from aero.lbm.sem_inlet import SEMInlet
    from aero.lbm.boundary import build_surface_links
    from aero.lbm.boundary3d import build_surface_links_3d
    from aero.geometry.cylinder import Cylinder
    from aero.geometry.rectangle import Rectangle
    from aero.geometry.polygon import Polygon
    from aero.geometry3d.sphere import Sphere
    from aero.geometry3d.box import Box
    from aero.geometry3d.cylinder3d import Cylinder3D
    from aero.lbm.mrt3d import M19_inv
    from aero.lbm.mrt2d import M9_inv
    return dict(**locals())


def noisy_feq_2d(R, Ny, Nx, u0, seed, amp=1e-4):
    rng = np.random.default_rng(seed)
    uy = rng.uniform(-amp, amp, size=(Ny, Nx))
    return R["compute_feq"](np.ones((Ny, Nx)), np.full((Ny, Nx), u0), uy), uy


def noisy_feq_3d(R, Nz, Ny, Nx, u0, seed, amp=1e-4):
    rng = np.random.default_rng(seed)
    uy = rng.uniform(-amp, amp, size=(Nz, Ny, Nx))
    uz = rng.uniform(-amp, amp, size=(Nz, Ny, Nx))
    f0 = R["compute_feq_3d"](np.ones((Nz, Ny, Nx)), np.full((Nz, Ny, Nx), u0), uy, uz)
    return f0, uy, uz


CASES_2D = {
    # name: (solver kwargs, geometry, snapshot steps)
    "bgk_bb_conv_slip": dict(kw=dict(collision="bgk"), geom="cyl"),
    "bgk_bouzidi_zg_sponge": dict(
        kw=dict(collision="bgk", bouzidi=True, outlet_bc="zerogradient",
                sponge_thickness=8, sponge_strength=0.2), geom="cyl"),
    "trt_bouzidi_conv_slip": dict(
        kw=dict(collision="trt", trt_lambda=0.25, bouzidi=True), geom="cyl", omega=1.2),
    "trt316_bb_zg_noslip": dict(
        kw=dict(collision="trt", trt_lambda=0.1875, outlet_bc="zerogradient",
                wall_bc="noslip"), geom="rect", omega=1.1),
    "mrt_bb_conv_noslip": dict(kw=dict(collision="mrt", wall_bc="noslip"), geom="cyl"),
    "mrt_bouzidi_zg_sponge": dict(
        kw=dict(collision="mrt", bouzidi=True, outlet_bc="zerogradient",
                sponge_thickness=6), geom="rect"),
    "bgk_perturb_moving": dict(
        kw=dict(collision="bgk", inlet_perturbation=0.3, wall_bc="moving",
                wall_velocity_top=0.02, wall_velocity_bottom=-0.01), geom="cyl"),
    "bgk_pressure_io": dict(
        kw=dict(collision="bgk", inlet_bc="pressure", outlet_bc="pressure",
                rho_in=1.002, rho_out=0.998), geom="cyl"),
    "bgk_les_smag": dict(kw=dict(collision="bgk", les=True, les_cs=0.2), geom="cyl", omega=1.95),
    "trt_les_wale": dict(
        kw=dict(collision="trt", les=True, les_model="wale", les_cs=0.3), geom="cyl", omega=1.3),
    "bgk_les_smag_vandriest": dict(
        kw=dict(collision="bgk", les=True, les_cs=0.2, van_driest=True, van_driest_A=3.0,
                bouzidi=True), geom="cyl", omega=1.9),
    "bgk_thermal_buoy": dict(
        kw=dict(collision="bgk", thermal=True, alpha_T=0.02, buoyancy=True, g_gravity=0.01,
                beta=0.5, T_hot=1.0, T_cold=0.0, T_ref=0.5), geom="cyl"),
    "bgk_synth_inflow": dict(
        kw=dict(collision="bgk", synthetic_inflow=True, synthetic_inflow_intensity=0.05),
        geom="cyl"),
    "bgk_polygon": dict(kw=dict(collision="bgk"), geom="poly"),
    "bgk_ibm": dict(kw=dict(collision="bgk", ibm_enabled=True), geom="cyl"),
    "mrt_ibm_zg_sponge": dict(
        kw=dict(collision="mrt", ibm_enabled=True, outlet_bc="zerogradient", sponge_thickness=6), geom="rect"),
}

CASES_3D = {
    "bgk_bouzidi_conv_slip": dict(kw=dict(collision="bgk", bouzidi=True), geom="sphere"),
    "trt_bouzidi_conv_slip": dict(
        kw=dict(collision="trt", trt_lambda=0.1875, bouzidi=True), geom="sphere", omega=1.1),
    "mrt_bouzidi_zg_noslip": dict(
        kw=dict(collision="mrt", bouzidi=True, outlet_bc="zerogradient", wall_bc="noslip"),
        geom="sphere"),
    "mrt_bb_conv_sponge": dict(
        kw=dict(collision="mrt", sponge_thickness=5, sponge_strength=0.2), geom="box"),
    "bgk_perturb_moving": dict(
        kw=dict(collision="bgk", inlet_perturbation=0.3, wall_bc="moving",
                wall_velocity_top=0.02, wall_velocity_bottom=-0.01), geom="box"),
    "bgk_periodic_bodyforce": dict(
        kw=dict(collision="bgk", streamwise_bc="periodic", body_force_x=1e-5,
                wall_bc="noslip"), geom="cyl3d"),
    "trt_recycling": dict(
        kw=dict(collision="trt", streamwise_bc="recycling"), geom="sphere", omega=1.2),
    "bgk_les_smag": dict(kw=dict(collision="bgk", les=True, les_cs=0.2), geom="sphere", omega=1.95),
    "trt_les_wale": dict(
        kw=dict(collision="trt", les=True, les_model="wale", les_cs=0.3), geom="sphere", omega=1.3),
    "bgk_thermal_buoy": dict(
        kw=dict(collision="bgk", thermal=True, alpha_T=0.02, buoyancy=True, g_gravity=0.01,
                beta=0.5, T_ref=0.5), geom="box"),
    "bgk_sem": dict(
        kw=dict(collision="bgk", sem_inlet=True, sem_Tu=0.05, sem_L_int=2.5, sem_N=20),
        geom="sphere"),
    "bgk_synth_inflow": dict(
        kw=dict(collision="bgk", synthetic_inflow=True, synthetic_inflow_intensity=0.05),
        geom="sphere"),
    "bgk_ibm": dict(kw=dict(collision="bgk", ibm_enabled=True), geom="sphere"),
    "trt_ibm_les": dict(kw=dict(collision="trt", ibm_enabled=True, les=True, les_cs=0.2), geom="box", omega=1.3),
    # the 512 x 256 x 256 MRT + WALE LES + scalar + Boussinesq + SEM inlet configuration, small
    "mrt_les_wale_thermal_sem": dict(
        kw=dict(collision="mrt", les=True, les_model="wale", les_cs=0.3, bouzidi=True, thermal=True,
                alpha_T=0.01, buoyancy=True, g_gravity=0.01, beta=0.5, T_ref=0.5, sem_inlet=True,
                sem_Tu=0.05, sem_L_int=2.5, sem_N=20), geom="box", omega=1.3),
}

NY2, NX2 = 20, 40
NZ3, NY3, NX3 = 6, 8, 16
SNAP2 = (1, 40)
SNAP3 = (1, 6)


def geom2d(R, name):
    if name == "cyl":
        g = R["Cylinder"](radius=5.0, cx_frac=0.3, cy_frac=0.5)
    elif name == "rect":
        g = R["Rectangle"](width=5.0, height=7.0, cx_frac=0.3, cy_frac=0.45)
    elif name == "poly":
        g = R["Polygon"]([(0.25, 0.35), (0.33, 0.42), (0.30, 0.62), (0.22, 0.60), (0.19, 0.45)])
    solid = g.mark_solid(NY2, NX2)
    phi = g.sdf_field(NY2, NX2)
    D = g.reference_length() if name != "poly" else 8.0
    return solid, phi, D


def geom3d(R, name):
    if name == "sphere":
        g = R["Sphere"](radius=2.3, cx_frac=0.3, cy_frac=0.5, cz_frac=0.5)
    elif name == "box":
        g = R["Box"](width=3.0, height=4.0, depth=3.0, cx_frac=0.3, cy_frac=0.45, cz_frac=0.5)
    elif name == "cyl3d":
        g = R["Cylinder3D"](radius=2.0, length=4.0, cx_frac=0.4, cy_frac=0.5, cz_frac=0.5)
    solid = g.mark_solid(NZ3, NY3, NX3)
    phi = g.sdf_field(NZ3, NY3, NX3)
    return solid, phi, g.reference_length()


def run_2d(R, name, spec, seed):
    solid, phi, D = geom2d(R, spec["geom"])
    u0 = 0.05
    omega = spec.get("omega", 1.0 / (3.0 * u0 * D / 60.0 + 0.5))
    kw = dict(spec["kw"])
    use_phi = kw.get("bouzidi", False) or kw.get("van_driest", False) or kw.get("ibm_enabled", False)
    s = R["Solver"](Ny=NY2, Nx=NX2, solid=solid, omega=omega, u0=u0, D=D,
                    backend="numpy", phi=phi if use_phi else None, **kw)
    f0, uy0 = noisy_feq_2d(R, NY2, NX2, u0, seed)
    s.f = f0.copy()
    s._f_outlet_prev = f0[:, :, -1].copy()
    out = dict(noise_uy=uy0, f0_sum=np.float64(f0.sum()), solid=solid, D=np.float64(D), omega=np.float64(omega),
               u0=np.float64(u0), links=s.surface_links, q_vals=s.q_vals,
               params_json=np.array(json.dumps(kw)))
    if phi is not None:
        out["phi"] = phi
    hist = {k: [] for k in ("Cd", "Cl", "Cm", "Cd_p", "Cd_v")}
    for step in range(1, max(SNAP2) + 1):
        cd, cl = s._step()
        hist["Cd"].append(cd)
        hist["Cl"].append(cl)
        hist["Cm"].append(s._last_cm)
        hist["Cd_p"].append(s._last_cd_p)
        hist["Cd_v"].append(s._last_cd_v)
        if step in SNAP2:
            out[f"f_{step}"] = np.array(s.f, dtype=np.float64)
            out[f"fprev_{step}"] = np.array(s._f_outlet_prev, dtype=np.float64)
            if s.g is not None:
                out[f"g_{step}"] = np.array(s.g, dtype=np.float64)
    for k, v in hist.items():
        out[f"{k}_hist"] = np.array(v, dtype=np.float64)
    prof = s._last_force_profile
    out["profile_fx"] = np.array(prof["fx"])
    out["profile_fy"] = np.array(prof["fy"])
    return out


def run_3d(R, name, spec, seed):
    solid, phi, D = geom3d(R, spec["geom"])
    u0 = 0.05
    omega = spec.get("omega", 1.0 / (3.0 * u0 * D / 40.0 + 0.5))
    kw = dict(spec["kw"])
    use_phi = kw.get("bouzidi", False) or kw.get("ibm_enabled", False)
    s = R["Solver3D"](Nz=NZ3, Ny=NY3, Nx=NX3, solid=solid, omega=omega, u0=u0, D=D,
                      backend="numpy", phi=phi if use_phi else None, **kw)
    f0, uy0, uz0 = noisy_feq_3d(R, NZ3, NY3, NX3, u0, seed)
    s.f = f0.copy()
    s._f_outlet_prev = f0[:, :, :, -1].copy()
    out = dict(noise_uy=uy0, noise_uz=uz0, f0_sum=np.float64(f0.sum()), solid=solid, D=np.float64(D), omega=np.float64(omega),
               u0=np.float64(u0), links=s.surface_links, q_vals=s.q_vals,
               params_json=np.array(json.dumps(kw)))
    if phi is not None:
        out["phi"] = phi
    keys = ("Cd", "Cly", "Clz", "Cmx", "Cmy", "Cmz", "Cd_p", "Cd_v")
    hist = {k: [] for k in keys}
    sem_planes = []
    for step in range(1, max(SNAP3) + 1):
        if s._sem is not None:
            # Record the SEM plane the step will use without disturbing the RNG:
            # snapshot the generator state, draw, and restore.
            import copy
            sem_copy = copy.deepcopy(s._sem)
            sem_copy.step()
            sem_planes.append(np.stack(sem_copy.fluctuation()))
        cd, cly, clz = s._step()
        hist["Cd"].append(cd)
        hist["Cly"].append(cly)
        hist["Clz"].append(clz)
        hist["Cmx"].append(s._last_cmx)
        hist["Cmy"].append(s._last_cmy)
        hist["Cmz"].append(s._last_cmz)
        hist["Cd_p"].append(s._last_cd_p)
        hist["Cd_v"].append(s._last_cd_v)
        if step in SNAP3:
            out[f"f_{step}"] = np.array(s.f, dtype=np.float64)
            out[f"fprev_{step}"] = np.array(s._f_outlet_prev, dtype=np.float64)
            if s.g is not None:
                out[f"g_{step}"] = np.array(s.g, dtype=np.float64)
    for k, v in hist.items():
        out[f"{k}_hist"] = np.array(v, dtype=np.float64)
    if sem_planes:
        out["sem_planes"] = np.array(sem_planes)
    prof = s._last_spanwise_profile
    for c in ("fx", "fy", "fz"):
        out[f"profile_{c}"] = np.array(prof[c])
    return out


def links_fixtures(R):
    """Link lists and q for the 400x200 cylinder and the 128x64x64 sphere."""
    out = {}
    g = R["Cylinder"](radius=20.0, cx_frac=0.25, cy_frac=0.5)
    solid = g.mark_solid(200, 400)
    phi = g.sdf_field(200, 400)
    links, q = R["build_surface_links"](solid, phi)
    out.update(c1_solid=solid, c1_links=links, c1_q=q)
    g3 = R["Sphere"](radius=8.0, cx_frac=0.25, cy_frac=0.5, cz_frac=0.5)
    solid3 = g3.mark_solid(64, 64, 128)
    phi3 = g3.sdf_field(64, 64, 128)
    links3, q3 = R["build_surface_links_3d"](solid3, phi3)
    out.update(c2_solid_count=np.int64(solid3.sum()), c2_links=links3, c2_q=q3)
    return out


def sem_fixture(R):
    sem = R["SEMInlet"](Ny=20, Nz=16, u0=0.05, Tu=0.05, L_int=4.0, N_eddies=30, rng_seed=0)
    planes = [np.stack(sem.fluctuation())]
    for _ in range(5):
        sem.step()
        planes.append(np.stack(sem.fluctuation()))
    return dict(planes=np.array(planes), xk=sem.xk, yk=sem.yk, zk=sem.zk, eps=sem.eps)


def signatures_fixture(R):
    """Parameter names of the reference constructors (solver.py:75-116,
    solver3d.py:71-119): a checkpoint's params_json must use only these."""
    import inspect
    out = {}
    for cls in ("Solver", "Solver3D"):
        out[cls] = [n for n in inspect.signature(R[cls].__init__).parameters if n != "self"]
    return out


# Checkpoints written by the reference itself: the reference runs CKPT_AT
# steps from the seeded state, saves with its own save_checkpoint, then runs
# CKPT_MORE steps more; the fixture pair holds the checkpoint and the state
# the continued run reaches.
CKPT_CASES_2D = ("bgk_ibm", "trt_bouzidi_conv_slip")
CKPT_CASES_3D = ("trt_ibm_les", "bgk_bouzidi_conv_slip")
CKPT_AT, CKPT_MORE = 12, 9


def ckpt_fixtures(R):
    for dim, names, table in ((2, CKPT_CASES_2D, CASES_2D), (3, CKPT_CASES_3D, CASES_3D)):
        for name in names:
            spec = table[name]
            seed = 300 + list(table).index(name)
            kw = dict(spec["kw"])
            if dim == 2:
                solid, phi, D = geom2d(R, spec["geom"])
                omega = spec.get("omega", 1.0 / (3.0 * 0.05 * D / 60.0 + 0.5))
                use_phi = kw.get("bouzidi", False) or kw.get("ibm_enabled", False)
                s = R["Solver"](Ny=NY2, Nx=NX2, solid=solid, omega=omega, u0=0.05, D=D, backend="numpy",
                                phi=phi if use_phi else None, **kw)
                f0, _ = noisy_feq_2d(R, NY2, NX2, 0.05, seed)
            else:
                solid, phi, D = geom3d(R, spec["geom"])
                omega = spec.get("omega", 1.0 / (3.0 * 0.05 * D / 40.0 + 0.5))
                use_phi = kw.get("bouzidi", False) or kw.get("ibm_enabled", False)
                s = R["Solver3D"](Nz=NZ3, Ny=NY3, Nx=NX3, solid=solid, omega=omega, u0=0.05, D=D,
                                  backend="numpy", phi=phi if use_phi else None, **kw)
                f0, _, _ = noisy_feq_3d(R, NZ3, NY3, NX3, 0.05, seed)
            s.f = f0.copy()
            s._f_outlet_prev = f0[..., -1].copy()
            for _ in range(CKPT_AT):
                s._step()
            ck = OUT / f"ckpt_ref{dim}d_{name}.npz"
            s.save_checkpoint(str(ck))
            cd = []
            for _ in range(CKPT_MORE):
                cd.append(s._step()[0])
            np.savez_compressed(OUT / f"ckpt_ref{dim}d_{name}_cont.npz", f=np.array(s.f, dtype=np.float64),
                                Cd_hist=np.array(cd), steps=np.int64(CKPT_MORE))
            print("ckpt", dim, name, flush=True)


def c1_long_run(R, steps=20000):
    """The reference's CPU-runnable configuration run to its end: D2Q9 BGK
    cylinder 400x200, r = 20 at (100, 100), u0 = 0.05, Re = 100, velocity
    inlet, zero-gradient outlet with a 40-column sponge, free-slip walls,
    equilibrium plus seeded 1e-4 transverse noise (seed 0), float64.
    Observables as cli.py:429-439 reports them."""
    from aero.diagnostics import compute_strouhal
    ny, nx, u0, D = 200, 400, 0.05, 40.0
    solid = R["Cylinder"](radius=20.0, cx_frac=0.25, cy_frac=0.5).mark_solid(ny, nx)
    s = R["Solver"](Ny=ny, Nx=nx, solid=solid, omega=1.0 / (3.0 * u0 * D / 100.0 + 0.5), u0=u0, D=D,
                    backend="numpy", collision="bgk", inlet_bc="velocity", outlet_bc="zerogradient",
                    sponge_thickness=40, wall_bc="slip")
    rng = np.random.default_rng(0)
    uy = rng.uniform(-1e-4, 1e-4, (ny, nx))
    f0 = R["compute_feq"](np.ones((ny, nx)), np.full((ny, nx), u0), uy)
    s.f = f0.copy()
    s._f_outlet_prev = f0[:, :, -1].copy()
    cd, cl = [], []
    for k in range(steps):
        a, b = s._step()
        cd.append(a)
        cl.append(b)
        if (k + 1) % 1000 == 0:
            print("c1 step", k + 1, a, b, flush=True)
    cd, cl = np.array(cd), np.array(cl)
    st = compute_strouhal(list(cl), D, u0)
    np.savez_compressed(OUT / "c1_long_run.npz", Cd_hist=cd, Cl_hist=cl, steps=np.int64(steps),
                        Cd_mean_last4000=np.float64(cd[-4000:].mean()), Cl_std_last4000=np.float64(cl[-4000:].std()),
                        St=np.float64(st if st is not None else np.nan), f_end=np.array(s.f, dtype=np.float64)[:, ::8, ::8])


def main(only=None):
    """Write every fixture, or with `only` (a substring) just the matching
    step cases; a case's seed is its position in its table, so adding cases at
    the end leaves the existing fixtures' inputs unchanged."""
    R = _import_reference()
    if only == "c1_long_run":
        c1_long_run(R)
        return
    if only in (None, "signatures"):
        (OUT / "ref_signatures.json").write_text(json.dumps(signatures_fixture(R), indent=1))
    if only in (None, "ckpt"):
        ckpt_fixtures(R)
    if only in ("signatures", "ckpt"):
        return
    if only is None:
        np.savez_compressed(OUT / "mrt_inverses.npz", M9_inv=R["M9_inv"], M19_inv=R["M19_inv"])
    for k, (name, spec) in enumerate(CASES_2D.items()):
        if only is not None and only not in name:
            continue
        out = run_2d(R, name, spec, seed=100 + k)
        np.savez_compressed(OUT / f"ref2d_{name}.npz", **out)
        print("2d", name, "links", len(out["links"]), "Cd", out["Cd_hist"][-1], flush=True)
    for k, (name, spec) in enumerate(CASES_3D.items()):
        if only is not None and only not in name:
            continue
        out = run_3d(R, name, spec, seed=200 + k)
        np.savez_compressed(OUT / f"ref3d_{name}.npz", **out)
        print("3d", name, "links", len(out["links"]), "Cd", out["Cd_hist"][-1], flush=True)
    if only is None:
        np.savez_compressed(OUT / "links_c1_c2.npz", **links_fixtures(R))
        np.savez_compressed(OUT / "sem_inlet.npz", **sem_fixture(R))
    print("done")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else None)
