"""SEM inlet planes formed on the device (aero_set_sem / aero_sem_planes)
against the reference's own planes (tests/golden/sem_inlet.npz, written from
aero/lbm/sem_inlet.py) and against the host SEMInlet.fluctuation(), from the
same eddy states.  The eddy sums are bitwise the host's; the FFT projection
(hipFFT vs numpy's pocketfft) and the reductions differ by rounding only,
hence the 1e-12 relative bound."""

import ctypes

import numpy as np
import pytest

from tests import _golden as G

pytestmark = pytest.mark.gpu


def _device_planes(sem, snaps):
    import aero_amd
    from aero_amd import _abi
    nz, ny = sem.Nz, sem.Ny
    s = aero_amd.Solver3D(Nz=nz, Ny=ny, Nx=8, solid=np.zeros((nz, ny, 8), dtype=bool), omega=1.0, u0=sem.u0, D=4.0,
                          sem_inlet=True, sem_Tu=sem.Tu, sem_L_int=sem.sigma, sem_N=sem.N)
    assert s._sem_device
    out = np.empty((len(snaps), 3, nz, ny))
    s._ctx.call("aero_sem_planes", ctypes.c_int32(len(snaps)), _abi.ptr(np.ascontiguousarray(snaps)), _abi.ptr(out))
    s.close()
    return out


def _snapshot(sem):
    return np.concatenate([sem.xk[None], sem.yk[None], sem.zk[None], sem.eps])


def test_device_planes_match_the_reference_planes():
    from aero_amd.inflow import SEMInlet
    with np.load(G.GOLDEN / "sem_inlet.npz") as d:
        ref = d["planes"]
    sem = SEMInlet(Ny=20, Nz=16, u0=0.05, Tu=0.05, L_int=4.0, N_eddies=30, rng_seed=0)
    snaps = [_snapshot(sem)]
    for _ in range(len(ref) - 1):
        sem.step()
        snaps.append(_snapshot(sem))
    print("[gpu] SEM planes vs reference", flush=True)
    got = _device_planes(sem, np.array(snaps))
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))


@pytest.mark.parametrize("ny,nz,n,L", [(63, 45, 200, 8.0), (64, 48, 120, 6.0), (33, 17, 5, 3.0)])
def test_device_planes_match_the_host_generator(ny, nz, n, L):
    """Odd and even plane sizes (the Nyquist row / column is zeroed only for
    even ones), many and few eddies; 70 steps, more than one chunk."""
    from aero_amd.inflow import SEMInlet
    sem = SEMInlet(Ny=ny, Nz=nz, u0=0.08, Tu=0.05, L_int=L, N_eddies=n, rng_seed=3)
    snaps, host = [], []
    for _ in range(70):
        sem.step()
        snaps.append(_snapshot(sem))
        host.append(np.stack(sem.fluctuation()))
    host = np.array(host)
    print(f"[gpu] SEM planes {ny}x{nz} N={n}", flush=True)
    got = _device_planes(sem, np.array(snaps))
    assert np.max(np.abs(got - host)) <= 1e-12 * np.max(np.abs(host))
    assert np.abs(host).max() > 0
