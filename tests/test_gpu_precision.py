"""Reduced-precision storage forms against float64.

1. Every golden case (31 reference-written fixtures: BGK/TRT/MRT, links,
   every inlet and outlet, sponge, walls, LES, IBM, SEM, body force, scalar +
   buoyancy, and the small MRT + WALE + scalar + SEM combination) replayed
   with FP32, FP16-shifted and BF16-shifted population storage (FP32
   arithmetic), against the reference's float64 outputs.  Stated tolerances
   per storage, for the populations f and the scalar g at the last snapshot:
     max  = max |a - ref| / max |ref|
     rms  = sqrt(mean((a - ref)^2)) / sqrt(mean(ref^2))
   and for every coefficient history (drag, lifts, moments, pressure and
   viscous drag) the largest deviation over the run, relative to the case's
   largest coefficient magnitude (lifts and moments of these near-symmetric
   cases stay close to 0, so their own peaks are no scale).  Measured maxima
   over all cases (MI355X): FP32 f 5.1e-6 / 6.7e-7, FP16 2.6e-4 / 8.6e-5,
   BF16 2.4e-3 / 7.2e-4; Cd 1.7e-6, 3.2e-4, 1.8e-3.  The tolerances leave a
   factor of about 3.
2. The 512 x 256 x 256 MRT + WALE LES + SEM inlet + scalar + Boussinesq
   configuration, FP32 and FP16-shifted storage against the FP64 storage of
   the same device path from the same state, over a bounded run: velocity,
   density and temperature fields (max, RMS) and the drag history.
MI355X only.
"""

import numpy as np
import pytest

from tests import _golden as G
from tests.test_gpu_parity import replay, _cols

pytestmark = pytest.mark.gpu

FIELD_TOL = {  # (max, rms) relative, f and g at the last snapshot
    "f32": (1.5e-5, 2e-6),
    "f16": (8e-4, 2.5e-4),
    "bf16": (8e-3, 2.5e-3),
}
CD_TOL = {"f32": 5e-6, "f16": 1e-3, "bf16": 6e-3}      # max |Cd - ref| / max |ref Cd|
COEF_TOL = {"f32": 1e-4, "f16": 2e-2, "bf16": 1.5e-1}  # every coefficient, relative to the largest one


def _field_errs(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    d = a - b
    return (float(np.max(np.abs(d)) / np.max(np.abs(b))),
            float(np.sqrt(np.mean(d * d)) / np.sqrt(np.mean(b * b))))


CASES = [(2, n) for n in G.CASES_2D] + [(3, n) for n in G.CASES_3D]


@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
@pytest.mark.parametrize("dim,name", CASES)
def test_golden_case_in_reduced_precision(dim, name, storage):
    print(f"[gpu] {storage} replay {dim}d {name}", flush=True)
    case, s, rows, fields = replay(dim, name, storage=storage)
    last = max(fields)
    f, prev, g = fields[last]
    tmax, trms = FIELD_TOL[storage]
    em, er = _field_errs(f, case[f"f_{last}"])
    assert em < tmax and er < trms, (name, storage, "f", em, er)
    if g is not None:
        em, er = _field_errs(g, case[f"g_{last}"])
        assert em < tmax and er < trms, (name, storage, "g", em, er)
    cols = _cols(dim)
    refs = {k: case[f"{k}_hist"] for k in cols}
    scale = max(float(np.max(np.abs(v))) for v in refs.values())
    cd_err = float(np.max(np.abs(rows[:, 0] - refs["Cd"])) / np.max(np.abs(refs["Cd"])))
    assert cd_err < CD_TOL[storage], (name, storage, cd_err)
    for k, col in cols.items():
        err = float(np.max(np.abs(rows[:, col] - refs[k])) / scale)
        assert err < COEF_TOL[storage], (name, storage, k, err)
    s.close()


C3_STEPS = 120
# vs the FP64 storage form: (field max, field rms, Cd max), relative (velocities
# to the largest velocity component).  Measured (MI355X, 120 steps): FP32
# 7.3e-6 / 1.4e-6 / 4.2e-7, FP16 9.0e-4 / 4.7e-5 / 1.7e-4.
C3_TOL = {
    "f32": (3e-5, 5e-6, 2e-6),
    "f16": (3e-3, 2e-4, 6e-4),
}


def _c3_solver(storage):
    import aero_amd
    from aero_amd.geometry import Box
    nx, ny, nz = 512, 256, 256
    g = Box(width=32.0, height=48.0, depth=48.0, cx_frac=128.0 / nx, cy_frac=0.5, cz_frac=0.5)
    solid, phi = g.mark_solid(nz, ny, nx), g.sdf_field(nz, ny, nx)
    u0, D, re = 0.08, 48.0, 20000.0
    omega = 1.0 / (3.0 * u0 * D / re + 0.5)
    return aero_amd.Solver3D(Nz=nz, Ny=ny, Nx=nx, solid=solid, omega=omega, u0=u0, D=D, phi=phi, bouzidi=True,
                             collision="mrt", les=True, les_model="wale", thermal=True, alpha_T=0.01,
                             buoyancy=True, g_gravity=1e-5, beta=1.0, sem_inlet=True, sem_Tu=0.05, sem_N=200,
                             sem_L_int=16.0, storage=storage)


def _c3_fields(s):
    rho, ux, uy, uz = s._macroscopic()
    T = s.g.sum(axis=0)
    return {"rho": rho, "ux": ux, "uy": uy, "uz": uz, "T": T}


def test_c3_512_fp32_and_fp16_against_fp64():
    """The MRT + WALE LES + SEM + scalar configuration at 512 x 256 x 256
    (the reference's MRT relaxes without the eddy viscosity), stepped
    C3_STEPS steps from the reference's equilibrium start in FP64, FP32 and
    FP16-shifted storage on the device.  Each form's SEM inlet draws the same
    eddies (seeded generator), so the three runs differ only by storage."""
    ref = _c3_solver("f64")
    cd64 = ref._advance(C3_STEPS)[:, 0]
    f64 = _c3_fields(ref)
    ref.close()
    del ref
    for storage in ("f32", "f16"):
        s = _c3_solver(storage)
        cd = s._advance(C3_STEPS)[:, 0]
        fields = _c3_fields(s)
        s.close()
        tmax, trms, tcd = C3_TOL[storage]
        u_scale = max(float(np.max(np.abs(f64[k]))) for k in ("ux", "uy", "uz"))
        for k, a in fields.items():
            b = f64[k]
            d = a - b
            scale = u_scale if k.startswith("u") else float(np.max(np.abs(b)))
            em = float(np.max(np.abs(d)) / scale)
            er = float(np.sqrt(np.mean(d * d)) / scale)
            print(f"[gpu] c3 512 {storage} {k}: max {em:.2e} rms {er:.2e}", flush=True)
            assert em < tmax and er < trms, (storage, k, em, er)
        cd_err = float(np.max(np.abs(cd - cd64)) / np.max(np.abs(cd64)))
        print(f"[gpu] c3 512 {storage} Cd: max {cd_err:.2e} (Cd64 last {cd64[-1]:.4f})", flush=True)
        assert np.all(np.isfinite(cd)) and cd_err < tcd, (storage, cd_err)
