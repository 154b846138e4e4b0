"""Throughput benchmark of the MI355X lattice-Boltzmann step.

Workload (one step = one lattice update of every cell, solid included):
3D D3Q19 TRT sphere, 512 x 256 x 256 (Nx x Ny x Nz), radius 32 centred at
(128, 128, 128), Bouzidi links, u0 = 0.05, Re = 100, slip walls, convective
outlet, equilibrium start plus seeded 1e-4 transverse velocity noise, float32
population storage: the configuration the FP32 roofline target is quoted on.
At this tau the reference's TRT rates are linearly unstable (s- < 0 for tau
below ~0.76) and its interpolated bounce-back diverges even in float64, so the
timed run uses the standard magic-parameter TRT (Lambda = 3/16) and the linear
Bouzidi rule with link fractions, labelled "trt_variant": "lambda",
"bouzidi_variant": "linear".

Prints ONE JSON line.  ``value`` is whole-job MLUPS over the timed steps
(all launches included, populations resident in HBM).  ``roofline`` is the
fused step kernel's algorithmic HBM rate (2 Q sizeof(float) bytes per cell)
over its average HIP-event duration.  ``cpu_baseline`` times the float64 C
oracle (OpenMP) on a bounded sample of full steps of the same workload, with
all host cores and with one.  ``latency`` gives the wall time of one
``_step()`` call (how the reference's own tests drive the solver) and the
host overhead of a multi-step call.
"""

from __future__ import annotations

import argparse
import json
import os
import pathlib
import sys
import time

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0


def make_case(nx, ny, nz, radius, cx, storage, steps_hint=0, seed=0, **extra):
    import aero_amd
    from aero_amd.geometry import Sphere
    from aero_amd.lattice import compute_feq_3d
    g = Sphere(radius=radius, cx_frac=cx / nx, cy_frac=0.5, cz_frac=0.5)
    solid, phi = g.mark_solid(nz, ny, nx), g.sdf_field(nz, ny, nx, offset=0.0)
    D = 2.0 * radius
    u0, re = 0.05, 100.0
    omega = 1.0 / (3.0 * u0 * D / re + 0.5)
    s = aero_amd.Solver3D(Nz=nz, Ny=ny, Nx=nx, solid=solid, omega=omega, u0=u0, D=D, phi=phi, bouzidi=True,
                          collision="trt", trt_lambda=0.1875, trt_variant="lambda", bouzidi_variant="linear",
                          storage=storage, **extra)
    rng = np.random.default_rng(seed)
    sh = (nz, ny, nx)
    f0 = compute_feq_3d(np.ones(sh), np.full(sh, u0), rng.uniform(-1e-4, 1e-4, sh), rng.uniform(-1e-4, 1e-4, sh))
    s.f = f0
    s._f_outlet_prev = f0[:, :, :, -1].copy()
    return s, dict(solid=solid, phi=phi, omega=omega, u0=u0, D=D, f0=f0)


def timed(solver, steps, warmup):
    """Wall time of exactly `steps` steps after `warmup` untimed ones, then the
    step kernel's average duration from HIP events recorded at the kernel's
    start and end on its own stream, over a further (untimed) run."""
    import ctypes
    solver._advance(warmup)
    solver._ctx.call("aero_synchronize")
    t0 = time.perf_counter()
    rows = solver._advance(steps)
    solver._ctx.call("aero_synchronize")
    t1 = time.perf_counter()
    solver._ctx.call("aero_set_timing", ctypes.c_int32(1))
    solver._advance(min(steps, 512))
    ms = ctypes.c_double(0.0)
    n = ctypes.c_int64(0)
    solver._ctx.call("aero_get_timing", ctypes.byref(ms), ctypes.byref(n))
    solver._ctx.call("aero_set_timing", ctypes.c_int32(0))
    return t1 - t0, ms.value / max(n.value, 1), rows


def cell_counts(solver):
    """(cells, link cells) of the solver's domain."""
    import ctypes
    tot, links = ctypes.c_int64(), ctypes.c_int64()
    solver._ctx.call("aero_cell_counts", ctypes.byref(tot), ctypes.byref(links))
    return int(tot.value), int(links.value)


def cpu_baseline(case, nx, ny, nz, budget_s=8.0):
    """Float64 C oracle (OpenMP) on the same workload: a bounded sample of
    full reference steps (collide, copy, stream, Bouzidi, inlet, outlet,
    walls, forces) and of kernel-only steps (collide, copy, stream, timed
    inside the C library), with all host cores and with one.  The per-step
    rate of the sample is the figure; the run is not carried to the end."""
    from oracle import oracle as orc
    o = orc.OracleSolver(case["solid"], omega=case["omega"], u0=case["u0"], D=case["D"], phi=case["phi"],
                         bouzidi=True, collision="trt", trt_lambda=0.1875, trt_variant="lambda",
                         bouzidi_variant="linear")
    o.set_state(case["f0"])
    cells = nx * ny * nz
    cores = len(os.sched_getaffinity(0))
    orc.set_threads(cores)

    def full(budget, min_steps=1, warm=True):
        if warm:
            o.step()  # first touch of the step's arrays
        t0 = time.perf_counter()
        n = 0
        while n < min_steps or time.perf_counter() - t0 < budget:
            o.step()
            n += 1
        return n, time.perf_counter() - t0

    n, dt = full(budget_s)
    kn = max(2, n // 2)
    kt = orc.kernel_only_seconds(o, kn)
    orc.set_threads(1)
    n1, dt1 = full(0.0, 1, warm=False)
    kt1 = orc.kernel_only_seconds(o, 1)
    orc.set_threads(cores)
    return {"value": round(cells * n / dt / 1e6, 3), "unit": "MLUPS", "cores": cores, "kind": "port",
            "kernel_only_MLUPS": round(cells * kn / kt / 1e6, 3),
            "one_core_MLUPS": round(cells * n1 / dt1 / 1e6, 3),
            "one_core_kernel_only_MLUPS": round(cells / kt1 / 1e6, 3),
            "sample": f"{n} full float64 steps of the {nx}x{ny}x{nz} D3Q19 TRT sphere on {cores} threads "
                      f"(collide, copy, stream, Bouzidi, inlet, outlet, walls, forces; kernel-only = collide, "
                      f"copy, stream over {kn} steps); one thread: {n1} full step, 1 kernel-only step",
            "cpu_model": _cpu_model()}


def latency(solver, calls=100):
    """Wall time of single ``_step()`` calls (the reference's drive pattern:
    one call per step, coefficients returned to the host each time), and of
    20-step calls beside the sum of their 20 kernel durations (HIP events):
    the difference is the per-call host and launch overhead."""
    import ctypes
    lat = []
    for _ in range(calls):
        t0 = time.perf_counter()
        solver._step()
        lat.append((time.perf_counter() - t0) * 1e3)
    walls20 = []
    for _ in range(5):
        solver._ctx.call("aero_synchronize")
        t0 = time.perf_counter()
        solver._advance(20)
        walls20.append((time.perf_counter() - t0) * 1e3)
    solver._ctx.call("aero_set_timing", ctypes.c_int32(1))
    solver._advance(20)
    ms = ctypes.c_double(0.0)
    n = ctypes.c_int64(0)
    solver._ctx.call("aero_get_timing", ctypes.byref(ms), ctypes.byref(n))
    solver._ctx.call("aero_set_timing", ctypes.c_int32(0))
    w20 = float(np.median(walls20))
    return {"step_call_ms_median": round(float(np.median(lat)), 4),
            "step_call_ms_p99": round(float(np.percentile(lat, 99)), 4),
            "calls": calls, "steps20_call_ms_median": round(w20, 4),
            "steps20_kernel_ms": round(ms.value, 4), "steps20_overhead_ms": round(w20 - ms.value, 4)}


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def traffic_from_profiles(name):
    p = ROOT / "profiles" / f"{name}.json"
    if p.exists():
        try:
            return json.loads(p.read_text()).get("hbm_bytes_per_launch")
        except (ValueError, OSError):
            return None
    return None


FORMS = {  # name: (nx, ny, nz, sphere radius, centre x, storage)
    "f32_128x64x64": (128, 64, 64, 8.0, 32.0, "f32"),
    "f16_128x64x64": (128, 64, 64, 8.0, 32.0, "f16"),
    "bf16_128x64x64": (128, 64, 64, 8.0, 32.0, "bf16"),
    "f64_128x64x64": (128, 64, 64, 8.0, 32.0, "f64"),
    "f32_512x256x256": (512, 256, 256, 32.0, 128.0, "f32"),
    "f16_512x256x256": (512, 256, 256, 32.0, 128.0, "f16"),
    "bf16_512x256x256": (512, 256, 256, 32.0, 128.0, "bf16"),
}


def run_form(name, steps, warmup):
    """One storage form of the TRT sphere alone: MLUPS over `steps` steps,
    the step kernel's average duration and its fraction of the HBM roofline
    (also the workload the per-form rocprof summaries are taken on)."""
    nx, ny, nz, r, cx, st = FORMS[name]
    s, _ = make_case(nx, ny, nz, r, cx, st)
    w, k, rows = timed(s, steps, warmup)
    cells = nx * ny * nz
    bpl = 2 * 19 * {"f32": 4, "f16": 2, "bf16": 2, "f64": 8}[st] * cells
    s.close()
    out = {"form": name, "MLUPS": round(cells * steps / w / 1e6, 2), "kernel_ms": round(k, 5),
           "roofline_frac": round(bpl / (k * 1e-3) / 1e9 / HBM_PEAK_GBS, 4), "cd_last": float(rows[-1, 0])}
    t = traffic_from_profiles(f"batch_step_kernel_d3q19_trt_{name}")
    if t:
        out["traffic"] = t
    return out


LES_FORMS = {"mrt_les_scalar_512_f32": "f32", "mrt_les_scalar_512_f16": "f16"}
# the TRT sphere with Smagorinsky LES (C_s = 0.16): the reference's
# finite-difference form (split step) and the fused local-stress form
TRT_LES_FORMS = {"trt_les_fd_512_f32": ("fd", "f32"), "trt_les_local_512_f32": ("local", "f32"),
                 "trt_les_local_512_f16": ("local", "f16")}
CYL2D_FORMS = {"cylinder_2d_400x200_f64": "f64", "cylinder_2d_400x200_f32": "f32",
               "cylinder_2d_400x200_f16": "f16"}


def run_cylinder_2d_form(name, steps, warmup):
    """The reference's CPU-runnable case on the device: D2Q9 BGK cylinder,
    400x200, radius 20 at (100, 100), u0 = 0.05, Re = 100, velocity inlet,
    zero-gradient outlet with a 40-column sponge, free-slip walls, seeded
    1e-4 transverse noise.  One case alone: 80,000 cells fill about 1,250
    wavefronts, so the step is launch-latency-bound."""
    import aero_amd
    from aero_amd.geometry import Cylinder
    from aero_amd.lattice import compute_feq
    ny, nx = 200, 400
    st = CYL2D_FORMS[name]
    solid = Cylinder(radius=20.0, cx_frac=0.25, cy_frac=0.5).mark_solid(ny, nx)
    u0, D = 0.05, 40.0
    s = aero_amd.Solver(Ny=ny, Nx=nx, solid=solid, omega=1.0 / (3.0 * u0 * D / 100.0 + 0.5), u0=u0, D=D,
                        collision="bgk", inlet_bc="velocity", outlet_bc="zerogradient", sponge_thickness=40,
                        wall_bc="slip", storage=st)
    rng = np.random.default_rng(0)
    s.f = compute_feq(np.ones((ny, nx)), np.full((ny, nx), u0), rng.uniform(-1e-4, 1e-4, (ny, nx)))
    w, k, rows = timed(s, steps, warmup)
    cells = nx * ny
    bpl = 2 * 9 * {"f64": 8, "f32": 4, "f16": 2}[st] * cells
    s.close()
    return {"form": name, "MLUPS": round(cells * steps / w / 1e6, 2), "kernel_ms": round(k, 5),
            "roofline_frac": round(bpl / (k * 1e-3) / 1e9 / HBM_PEAK_GBS, 4), "cd_last": float(rows[-1, 0])}


def run_les_form(name, steps, warmup):
    """The MRT + WALE LES 512x256x256 configuration: box obstacle 32x48x48 at
    x = 128 with Bouzidi links, u0 = 0.08, Re = 20000 (on the box height),
    SEM inlet turbulence (Tu = 5%, 200 eddies, length scale 16 cells),
    passive scalar (diffusivity 0.01) with Boussinesq buoyancy g*beta = 1e-5,
    in FP32 or FP16-shifted storage.  The reference's MRT relaxes without the
    eddy viscosity, so the step is the fused kernel with the scalar; the
    algorithmic bytes are both population sets, 2 * 2 * 19 * s per cell.
    The SEM planes are generated on the host every step, so the wall rate
    includes that work; the kernel figure does not."""
    import aero_amd
    from aero_amd.geometry import Box
    nx, ny, nz = 512, 256, 256
    st = LES_FORMS[name]
    g = Box(width=32.0, height=48.0, depth=48.0, cx_frac=128.0 / nx, cy_frac=0.5, cz_frac=0.5)
    solid, phi = g.mark_solid(nz, ny, nx), g.sdf_field(nz, ny, nx)
    u0, D, re = 0.08, 48.0, 20000.0
    omega = 1.0 / (3.0 * u0 * D / re + 0.5)
    s = aero_amd.Solver3D(Nz=nz, Ny=ny, Nx=nx, solid=solid, omega=omega, u0=u0, D=D, phi=phi, bouzidi=True,
                          collision="mrt", les=True, les_model="wale", thermal=True, alpha_T=0.01,
                          buoyancy=True, g_gravity=1e-5, beta=1.0, sem_inlet=True, sem_Tu=0.05, sem_N=200,
                          sem_L_int=16.0, storage=st)
    w, k, rows = timed(s, steps, warmup)
    cells = nx * ny * nz
    bpl = 2 * 2 * 19 * {"f32": 4, "f16": 2}[st] * cells
    s.close()
    return {"form": name, "MLUPS_wall": round(cells * steps / w / 1e6, 2),
            "MLUPS_kernel": round(cells / (k * 1e-3) / 1e6, 2), "kernel_ms": round(k, 5),
            "roofline_frac": round(bpl / (k * 1e-3) / 1e9 / HBM_PEAK_GBS, 4), "cd_last": float(rows[-1, 0])}


def run_trt_les_form(name, steps, warmup):
    """The 512x256x256 TRT sphere with Smagorinsky LES.  "fd" is the
    reference's finite-difference model: each step writes the post-boundary
    populations and velocity in compute precision, then a second kernel
    collides with the eddy viscosity from central differences.  "local" is one
    fused launch with the eddy viscosity from each cell's non-equilibrium
    stress, so its algorithmic bytes are the plain step's, 2 * 19 * s per
    cell; the figure for "fd" is the wall rate only."""
    nx, ny, nz = 512, 256, 256
    form, st = TRT_LES_FORMS[name]
    s, _ = make_case(nx, ny, nz, 32.0, 128.0, st, les=True, les_cs=0.16, les_form=form)
    w, k, rows = timed(s, steps, warmup)
    cells = nx * ny * nz
    s.close()
    out = {"form": name, "MLUPS": round(cells * steps / w / 1e6, 2), "cd_last": float(rows[-1, 0])}
    if form == "local":
        bpl = 2 * 19 * {"f32": 4, "f16": 2}[st] * cells
        out.update(kernel_ms=round(k, 5), roofline_frac=round(bpl / (k * 1e-3) / 1e9 / HBM_PEAK_GBS, 4))
    return out


SLAB_FORMS = {"slab_512x256x32_f32": "f32", "slab_512x256x32_f16": "f16"}
BATCH_FORMS = {"c4_batch_256_f32": "f32", "c4_batch_256_f16": "f16"}


def run_batch_form(name, steps, warmup):
    """The batched 2D workload's stepping alone: the 256 cases of
    serve.c4_cases(256, seed=0) (nx 128-1024, ny = nx/2; BGK / TRT / MRT),
    all live, advanced together through Batch.advance -- per step one launch
    for each collision group.  Reports lattice updates over the wall time of
    the timed steps (launches, chunk-end reductions and coefficient-row
    downloads included) and the algorithmic bytes per step: 2 * 9 * s per
    cell."""
    from concurrent.futures import ThreadPoolExecutor
    from aero_amd import serve
    from aero_amd.batch import Batch
    st = BATCH_FORMS[name]
    specs = serve.c4_cases(256, seed=0, storage=st)
    with ThreadPoolExecutor(16) as ex:
        solvers = list(ex.map(lambda c: c.build(), specs))
    cells = {}
    for c in specs:
        cells[c.collision] = cells.get(c.collision, 0) + c.cells
    total = sum(cells.values())
    b = Batch()
    b.advance(solvers, warmup, record=False)
    t0 = time.perf_counter()
    b.advance(solvers, steps, record=False)
    wall = time.perf_counter() - t0
    b.close()
    for s in solvers:
        s.close()
    esz = {"f32": 4, "f16": 2}[st]
    return {"form": name, "cases": len(specs), "cells": total, "cells_by_collision": cells, "steps": steps,
            "MLUPS": round(total * steps / wall / 1e6, 1), "ms_per_step": round(wall / steps * 1e3, 4),
            "algorithmic_bytes_per_step": 2 * 9 * esz * total}


def run_slab_form(name, steps, warmup):
    """One GPU's share of the 512x256x256 TRT sphere split into 8 z slabs:
    the slabs [96, 128) and [128, 160) (the sphere spans both), each 32 own
    planes plus 2 halo planes, stepped with a halo exchange before every step
    (aero_halo_exchange: the 5 crossing populations of each face, packed and
    unpacked on the device).  Reports the step kernel of the upper slab (HIP
    events; its own 512x256x32 cells' algorithmic bytes, the halo planes'
    work counted as overhead) and the exchange's wall time per step for both
    faces of the pair, including the host synchronisation the one-device
    exchange does."""
    import ctypes
    from aero_amd import slabs as S
    from aero_amd.geometry import Sphere
    nx, ny, nz, r = 512, 256, 256, 32.0
    st = SLAB_FORMS[name]
    g = Sphere(radius=r, cx_frac=128.0 / nx, cy_frac=0.5, cz_frac=0.5)
    solid, phi = g.mark_solid(nz, ny, nx), g.sdf_field(nz, ny, nx, offset=0.0)
    u0, D = 0.05, 2 * r
    kw = dict(omega=1.0 / (3.0 * u0 * D / 100.0 + 0.5), u0=u0, D=D, bouzidi=True, collision="trt",
              trt_lambda=0.1875, trt_variant="lambda", bouzidi_variant="linear", storage=st)
    bounds = S.slab_bounds(nz, 8)[3:5]
    slabs = [S.make_slab(solid, z0, z1, phi=phi, **kw) for z0, z1 in bounds]
    del solid, phi
    for s in slabs:
        s.set_equilibrium()
    S.advance(slabs, warmup, periodic=False)
    for s in slabs:
        s._ctx.call("aero_synchronize")
    t0 = time.perf_counter()
    S.advance(slabs, steps, periodic=False)
    for s in slabs:
        s._ctx.call("aero_synchronize")
    wall = time.perf_counter() - t0
    up = slabs[1]
    up._ctx.call("aero_set_timing", ctypes.c_int32(1))
    for _ in range(min(steps, 100)):
        up._advance(1)
    ms = ctypes.c_double(0.0)
    n = ctypes.c_int64(0)
    up._ctx.call("aero_get_timing", ctypes.byref(ms), ctypes.byref(n))
    up._ctx.call("aero_set_timing", ctypes.c_int32(0))
    k = ms.value / max(n.value, 1)
    t1 = time.perf_counter()
    nx_ex = 50
    for _ in range(nx_ex):
        S.exchange(slabs, periodic=False)
    ex = (time.perf_counter() - t1) / nx_ex
    hb = ctypes.c_int64(0)
    up._ctx.call("aero_halo_bytes", ctypes.byref(hb))
    own = nx * ny * (bounds[1][1] - bounds[1][0])
    bpc = 2 * 19 * {"f32": 4, "f16": 2}[st]
    for s in slabs:
        s.close()
    return {"form": name, "slab_cells": own, "halo_planes": 2, "kernel_ms": round(k, 5),
            "MLUPS_kernel": round(own / (k * 1e-3) / 1e6, 1),
            "roofline_frac": round(own * bpc / (k * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "pair_step_ms_wall": round(wall / steps * 1e3, 4), "exchange_ms_wall": round(ex * 1e3, 4),
            "face_bytes": int(hb.value), "faces_per_exchange": 2}


def stl_geometry(nz=128, ny=128, nx=256, sample=4000):
    """STL preprocessing on the device: the sample airplane (3,458 triangles)
    voxelised and its narrow-band SDF on a 256x128x128 grid, beside the
    numpy oracle's parity test timed on a bounded sample of cells."""
    from aero_amd import mesh
    from oracle import mesh_oracle
    raw = mesh.load_stl_triangles(str(ROOT / "tests" / "golden" / "stl" / "simple_plane.stl"))
    tris, _ = mesh.prepare_mesh_triangles(raw, nz, ny, nx, fit_frac=0.35)
    mesh.grid_inside_mesh(tris, 8, 8, 8)
    t0 = time.perf_counter()
    ins = mesh.grid_inside_mesh(tris, nz, ny, nx)
    t_vox = time.perf_counter() - t0
    t0 = time.perf_counter()
    phi = mesh.compute_phi_field(raw, nz, ny, nx, fit_frac=0.35)
    t_phi = time.perf_counter() - t0
    cells = nz * ny * nx
    idx = np.random.default_rng(0).choice(cells, sample, replace=False)
    t0 = time.perf_counter()
    ref = mesh_oracle.inside(mesh_oracle.grid_centres(nz, ny, nx)[idx], tris)
    t_cpu = time.perf_counter() - t0
    return {"mesh": "simple_plane.stl", "triangles": int(len(tris)), "grid": [nz, ny, nx],
            "voxelise_s": round(t_vox, 4), "ray_triangle_tests_per_s": round(cells * len(tris) / t_vox, 1),
            "phi_field_s": round(t_phi, 4), "band_cells": int((np.abs(phi) < 16.0).sum()),
            "solid_cells": int(ins.sum()), "sample_agrees_with_oracle": bool((ref == ins.reshape(-1)[idx]).all()),
            "cpu_oracle_voxelise_s_extrapolated": round(t_cpu * cells / sample, 1),
            "cpu_oracle_sample": f"{sample} cells, numpy, 1 thread"}


def copy_calibration(mib=1024, reps=4):
    """A plain float32 device-to-device copy of `mib` MiB, `reps` times: the
    kernel whose FETCH_SIZE / WRITE_SIZE per dispatch (known to be `mib` MiB
    each way) scales the counters read for the step kernel."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.rand(n, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_each_way": n * 4, "reps": reps}), flush=True)


def _sequential(cases, chunk=500):
    """Each case alone through the plain single-case API, one after another:
    Solver.run with the reference's stability check and auto-stop every
    `chunk` steps.  Returns (stop reason, steps, Cd_mean, warned) per case."""
    import warnings
    out = []
    t0 = time.perf_counter()
    updates = 0
    for k, c in enumerate(cases):
        if k and k % 32 == 0:
            print(f"[serving] sequential_2d: {k} cases, {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
        s = c.build()
        with warnings.catch_warnings(record=True) as caught:
            warnings.simplefilter("always")
            try:
                r = s.run(steps=c.max_steps, check_every=chunk, verbose=False, auto_stop=True)
                reason, cd = r["stop_reason"], np.asarray(r["Cd_history"])
            except RuntimeError:
                reason, cd = "diverged", np.asarray(s.Cd_history)
        warned = any(issubclass(w.category, RuntimeWarning) for w in caught)
        updates += s.step_count * c.cells
        w = max(1, len(cd) // 5)
        out.append((reason, len(cd), float(np.mean(cd[-w:])) if len(cd) else float("nan"), warned))
        s.close()
    return out, time.perf_counter() - t0, updates


class _Progress:
    """Prints a line to stderr every `every` decided requests of a serving leg."""

    def __init__(self, leg, every=32):
        self.leg, self.every, self.n, self.t0 = leg, every, 0, time.perf_counter()

    def __call__(self, res):
        self.n += 1
        if self.n % self.every == 0:
            print(f"[serving] {self.leg}: {self.n} requests decided, {time.perf_counter() - self.t0:.1f} s",
                  file=sys.stderr, flush=True)


def serving(full=False):
    """Batched 2D solves (the 256-case workload at its stated size, FP32 and
    FP16-shifted), the same cases run one at a time through the plain
    single-case API, and Poisson arrival streams of mixed 2D/3D requests (two
    bounded rates unless `full`, then the 1-50 requests/s sweep at full
    size); returns solves/hour, p50/p99 time-to-result and aggregate MLUPS
    for each."""
    from aero_amd import serve
    out = {}
    n_cases, max_steps = 256, 30000
    cases = serve.c4_cases(n_cases, seed=0, max_steps=max_steps)
    srv = serve.Server(chunk=500, on_result=_Progress("batched_2d"))
    t0 = time.perf_counter()
    res = srv.serve([(0.0, c) for c in cases])
    wall = time.perf_counter() - t0
    out["batched_2d"] = dict(serve.summarize(res, wall, srv.steps_done, srv.busy_seconds), cases=n_cases,
                             max_steps=max_steps)
    # the same cases in the FP16-shifted storage form (FP32 arithmetic)
    cases16 = serve.c4_cases(n_cases, seed=0, max_steps=max_steps, storage="f16")
    srv16 = serve.Server(chunk=500, on_result=_Progress("batched_2d_f16"))
    t0 = time.perf_counter()
    res16 = srv16.serve([(0.0, c) for c in cases16])
    wall16 = time.perf_counter() - t0
    by32 = {r.rid: r for r in res}
    rel = [abs(r.Cd_mean - by32[r.rid].Cd_mean) / abs(by32[r.rid].Cd_mean) for r in res16
           if r.stop_reason != "diverged" and by32[r.rid].stop_reason != "diverged"]
    out["batched_2d_f16"] = dict(serve.summarize(res16, wall16, srv16.steps_done, srv16.busy_seconds),
                                 cases=n_cases, max_steps=max_steps,
                                 cd_mean_rel_diff_vs_f32_max=float(max(rel, default=0.0)),
                                 cd_mean_rel_diff_vs_f32_median=float(np.median(rel)) if rel else 0.0)
    seq, wall_s, upd_s = _sequential(cases)
    reasons = {}
    for r in seq:
        reasons[r[0]] = reasons.get(r[0], 0) + 1
    solved = sum(1 for r in seq if r[0] != "diverged")
    out["sequential_2d"] = {"requests": len(seq), "solves": solved, "diverged": len(seq) - solved,
                            "warned": sum(1 for r in seq if r[3]), "wall_s": round(wall_s, 3),
                            "solves_per_hour": round(solved / wall_s * 3600.0, 1),
                            "MLUPS": round(upd_s / wall_s / 1e6, 1), "stop_reasons": reasons}
    # batched and sequential runs stop at the same step with the same drag
    by = {r.rid: r for r in res}
    same = [k for k, r in enumerate(seq) if r[0] != "diverged" and by[k].stop_reason != "diverged"]
    out["sequential_2d"]["same_stop_step"] = all(seq[k][1] == by[k].steps for k in same)
    out["sequential_2d"]["max_abs_cd_mean_diff"] = float(max((abs(seq[k][2] - by[k].Cd_mean) for k in same),
                                                              default=0.0))
    out["batched_over_sequential_solves_per_hour"] = round(
        out["batched_2d"]["solves_per_hour"] / out["sequential_2d"]["solves_per_hour"], 2)
    out["batched_f16_over_sequential_f32_solves_per_hour"] = round(
        out["batched_2d_f16"]["solves_per_hour"] / out["sequential_2d"]["solves_per_hour"], 2)
    legs = [(r, "f32") for r in ((2.0, 10.0) if not full else (1.0, 5.0, 10.0, 20.0, 50.0))]
    if full:  # the FP16-shifted storage form of the same streams
        legs += [(r, "f16") for r in (5.0, 20.0, 50.0)]
    for rate, st in legs:
        n = 48 if not full else 200
        reqs = serve.c5_requests(n, rate=rate, seed=int(rate), max_steps=4000 if not full else 30000, storage=st)
        name = f"poisson_{rate:g}" + ("" if st == "f32" else f"_{st}")
        srv = serve.Server(chunk=500, on_result=_Progress(name, every=16))
        t0 = time.perf_counter()
        r5 = srv.serve(reqs)
        out[f"{name}_per_s" if st == "f32" else f"poisson_{rate:g}_per_s_{st}"] = dict(
            serve.summarize(r5, time.perf_counter() - t0, srv.steps_done, srv.busy_seconds),
            requests_rate_per_s=rate, storage=st)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-extra", action="store_true", help="skip the extra storage forms")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving / batched-solves leg")
    ap.add_argument("--serve-full", action="store_true",
                    help="serving leg with the full arrival-rate sweep of Poisson streams")
    ap.add_argument("--form", choices=sorted(FORMS) + sorted(LES_FORMS) + sorted(CYL2D_FORMS) + sorted(TRT_LES_FORMS)
                    + sorted(SLAB_FORMS) + sorted(BATCH_FORMS), default=None,
                    help="only time one storage form of the TRT sphere (prints one JSON line)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain device copy of known size (calibrates the HBM byte counters)")
    a = ap.parse_args()
    if a.copy_calibration:
        copy_calibration()
        return
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the solver keeps one resident domain per device and "
              "has no multi-GPU path (a large grid would split into x-slabs with one face exchange "
              "per step).", file=sys.stderr)
        sys.exit(2)
    if a.form in CYL2D_FORMS:
        print(json.dumps(run_cylinder_2d_form(a.form, a.steps, a.warmup)), flush=True)
        return
    if a.form in LES_FORMS:
        print(json.dumps(run_les_form(a.form, a.steps, a.warmup)), flush=True)
        return
    if a.form in TRT_LES_FORMS:
        print(json.dumps(run_trt_les_form(a.form, a.steps, a.warmup)), flush=True)
        return
    if a.form in SLAB_FORMS:
        print(json.dumps(run_slab_form(a.form, a.steps, a.warmup)), flush=True)
        return
    if a.form in BATCH_FORMS:
        print(json.dumps(run_batch_form(a.form, a.steps, a.warmup)), flush=True)
        return
    if a.form:
        print(json.dumps(run_form(a.form, a.steps, a.warmup)), flush=True)
        return

    nx, ny, nz = 512, 256, 256
    cells = nx * ny * nz
    s, case = make_case(nx, ny, nz, 32.0, 128.0, "f32")
    wall, kern_ms, rows = timed(s, a.steps, a.warmup)
    if not np.all(np.isfinite(rows[:, 0])):
        raise RuntimeError("non-finite drag in the timed run")
    mlups = cells * a.steps / wall / 1e6
    ncell, nlink = cell_counts(s)
    bytes_per_launch = 2 * 19 * 4 * ncell
    achieved = bytes_per_launch / (kern_ms * 1e-3) / 1e9
    out = {
        "metric": "MLUPS + %HBM roofline, D3Q19 FP32 vs mixed",
        "value": round(mlups, 2),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(wall / a.steps * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: analytic sphere, equilibrium + seeded 1e-4 noise",
        "config": {"workload": "D3Q19 TRT sphere 512x256x256 r=32 Bouzidi, u0=0.05 Re=100, fp32 storage",
                   "nx": nx, "ny": ny, "nz": nz, "collision": "trt", "trt_variant": "lambda",
                   "bouzidi_variant": "linear",
                   "trt_lambda": 0.1875, "storage": "f32", "links": int(len(s.surface_links))},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic_from_profiles("batch_step_kernel_d3q19_trt_f32_512x256x256"),
                     "kernel": "aero::batch_step_kernel<D3Q19, TRT, F32>", "kernel_ms": round(kern_ms, 5),
                     "algorithmic_bytes_per_launch": bytes_per_launch, "cells": ncell, "link_cells": nlink,
                     "bytes_per_cell": 2 * 19 * 4},
        "cd_last": float(rows[-1, 0]),
    }
    if not a.no_extra:
        out["latency"] = latency(s)
    s.close()
    if not a.no_extra:
        extra = {}
        # the mixed-precision forms of the same workload, then the 128 x 64 x 64 sphere (radius 8)
        for name in ("f16_512x256x256", "bf16_512x256x256", "f32_128x64x64", "f16_128x64x64", "bf16_128x64x64",
                     "f64_128x64x64"):
            n2 = max(200, a.steps) if "128x" in name else max(100, a.steps)
            extra[name] = run_form(name, n2, max(a.warmup, 20))
        extra["f16_over_f32_MLUPS"] = round(extra["f16_512x256x256"]["MLUPS"] / out["value"], 3)
        extra["f16_over_f32_kernel_time"] = round(kern_ms / extra["f16_512x256x256"]["kernel_ms"], 3)
        out["forms"] = extra
        out["stl_geometry"] = stl_geometry()
    if not a.no_serve:
        out["serving"] = serving(full=a.serve_full)
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(case, nx, ny, nz)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
