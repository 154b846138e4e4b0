"""Device convergence statistics against the host restatement (MI355X only).

aero_stationarity (FP64 on the device: window sums, lift choice, direct DFT
peak search) must take the same decisions as
diagnostics.detect_statistical_stationarity, which restates
aero/diagnostics.py:87-228: stop, drag convergence, lift series, the
Strouhal FFT bin of the window and of its halves, and the drift flag; the
statistics agree to rounding (1e-11 relative).  Cases: synthetic drag and
lift series of every regime (short histories, below and above the 512 and
1024 sample thresholds, converged and drifting drag, shedding at a fixed
and at a changing frequency, 3D with the stronger lift on z), and the rings
that real 2D and 3D runs append.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _host(cd, ly, lz, D, u0, **kw):
    from aero_amd.diagnostics import detect_statistical_stationarity
    lift = ly
    if lz is not None:
        w = kw.get("strouhal_window", 4096)
        if np.std(lz[-min(len(lz), w):]) > np.std(ly[-min(len(ly), w):]):
            lift = lz
    conv, st, stop = detect_statistical_stationarity(list(cd), list(lift), D, u0, **kw)
    return conv, st, stop, (2 if lift is lz and lz is not None else 1)


def _check(row, cd, ly, lz, D, u0, **kw):
    conv, st, stop, which = _host(cd, ly, lz, D, u0, **kw)
    assert bool(row[0]) == stop
    assert bool(row[1]) == conv.converged
    if np.isfinite(conv.ratio):
        assert row[2] == pytest.approx(conv.ratio, rel=1e-11, abs=1e-15)
        assert row[3] == pytest.approx(conv.mean_abs, rel=1e-11)
        assert row[4] == pytest.approx(conv.std, rel=1e-11, abs=1e-15)
    else:
        assert np.isinf(row[2])
    assert int(row[11]) == which
    if st.strouhal is None:
        assert np.isnan(row[5])
    else:
        assert row[5] == pytest.approx(st.strouhal, rel=1e-12)
    if st.relative_drift is None:
        assert np.isnan(row[9])
    else:
        assert row[9] == pytest.approx(st.relative_drift, rel=1e-9, abs=1e-14)
        assert bool(row[10]) == st.stationary


def _synthetic(rng, n, kind):
    t = np.arange(n, dtype=np.float64)
    cd = 1.3 + 0.5 * np.exp(-t / 800.0) + 1e-3 * rng.standard_normal(n)
    if kind == "drift":
        cd = cd + 2e-5 * t
    f0 = rng.uniform(0.005, 0.05)
    if kind == "chirp":
        ly = 0.3 * np.sin(2 * np.pi * (f0 + 4e-6 * t) * t)
    else:
        ly = 0.3 * np.sin(2 * np.pi * f0 * t + rng.uniform(0, 6.28))
    ly = ly + 1e-3 * rng.standard_normal(n)
    if kind == "flat":
        ly = np.full(n, 0.01) + 1e-9 * rng.standard_normal(n)
    return cd, ly


def test_synthetic_histories_match_host_decisions():
    from aero_amd.stationarity import stationarity_arrays
    rng = np.random.default_rng(0)
    cases = []
    for n in (100, 511, 600, 1023, 3000, 4095, 4096, 5000, 9000, 20000):
        for kind in ("steady", "drift", "chirp", "flat"):
            cd, ly = _synthetic(rng, n, kind)
            lz = None
            if kind != "flat" and rng.random() < 0.5:
                lz = 0.5 * ly[::-1].copy() + 0.2 * np.cos(2 * np.pi * 0.02 * np.arange(n))
            cases.append((cd, ly, lz, 20.0, 0.05))
    for kw in (dict(), dict(convergence_window=2000, convergence_tol=0.02, strouhal_window=2048, strouhal_tol=0.08)):
        out = stationarity_arrays(cases, **kw)
        for row, (cd, ly, lz, D, u0) in zip(out, cases):
            _check(row, cd, ly, lz, D, u0, **kw)
    # some of the converged, stationary synthetic cases stop
    assert out[:, 0].sum() > 0


def test_rings_of_real_runs_match_host_decisions():
    from aero_amd import serve
    from aero_amd.stationarity import enable_ring, stationarity
    kw = dict(convergence_window=1000, strouhal_window=512)
    specs = [serve.c4_cases(1, seed=3, max_steps=8000)[0],
             serve.CaseSpec(dim=3, nx=96, ny=48, nz=48, shape="sphere", size=(6.0,), Re=60.0, collision="trt",
                            seed=4)]
    solvers = [s.build() for s in specs]
    for s in solvers:
        enable_ring(s, 1500)
    for n in (300, 700, 2500, 3000):
        for s in solvers:
            s._record(s._advance(n))
        out = stationarity(solvers, **kw)
        for row, s in zip(out, solvers):
            lz = np.asarray(s.Clz_history) if s._is3d else None
            ly = np.asarray(s.Cly_history if s._is3d else s.Cl_history)
            _check(row, np.asarray(s.Cd_history), ly, lz, s.D, s.u0, **kw)
    for s in solvers:
        s.close()
