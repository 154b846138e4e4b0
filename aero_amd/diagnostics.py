"""Run-loop diagnostics: stability, convergence, Strouhal, spectra.

Restates aero/diagnostics.py:11-267 and aero/observables.py:10-43 with the
same thresholds, messages and report fields.  The stability test takes the
FP64 statistics the device reduces over the current state (NaN flag,
non-positive density flag, max |(ux, uy)|, sum and count of density) instead
of a host copy of the field.
"""

from __future__ import annotations

import warnings
from dataclasses import dataclass
from typing import Dict, List, Optional

import numpy as np


@dataclass
class ConvergenceReport:
    converged: bool
    window: int
    ratio: float
    mean_abs: float
    std: float
    threshold: float


@dataclass
class StrouhalReport:
    strouhal: Optional[float]
    peak_frequency: Optional[float]
    peak_amplitude: float
    sample_count: int
    window: int
    stationary: bool = False
    relative_drift: Optional[float] = None


def stability_findings(stats: np.ndarray, step: int):
    """The outcome of check_stability (aero/diagnostics.py:32-71) for device
    statistics, without acting on it: (error message or None, warning
    messages).  The error is the one the reference raises first."""
    nan_any, nonpos_any, umax, rho_sum, count = (float(v) for v in stats)
    if nan_any > 0.0:
        return (f"[step {step}] NaN detected in f — simulation has diverged. "
                "Check tau (must be > 0.5), inlet velocity, and grid resolution."), []
    if nonpos_any > 0.0:
        return f"[step {step}] Non-positive density detected — simulation has diverged.", []
    warns = []
    if umax > 0.3:
        warns.append(f"[step {step}] Max velocity = {umax:.4f} lattice units (Ma ≈ {umax * (3 ** 0.5):.3f}). "
                     "Compressibility artifacts likely. Reduce u0 or increase grid size.")
    rho_mean = rho_sum / max(count, 1.0)
    if abs(rho_mean - 1.0) > 0.05:
        warns.append(f"[step {step}] Mean density drifted to {rho_mean:.4f} (>5% from 1.0). "
                     "Check outlet BC or total mass conservation.")
    return None, warns


def check_stability_stats(stats: np.ndarray, step: int) -> None:
    """check_stability (aero/diagnostics.py:32-71) from device statistics:
    RuntimeError on divergence, RuntimeWarning for the soft limits."""
    err, warns = stability_findings(stats, step)
    if err is not None:
        raise RuntimeError(err)
    for w in warns:
        warnings.warn(w, RuntimeWarning, stacklevel=2)


def check_stability(f, rho, ux, uy, step: int) -> None:
    """Host-array form of the same test (aero/diagnostics.py:32-71)."""
    stats = np.array([float(np.any(np.isnan(f))), float(np.any(rho <= 0.0)),
                      float(np.sqrt(ux ** 2 + uy ** 2).max()), float(rho.sum()), float(rho.size)])
    check_stability_stats(stats, step)


def analyze_convergence(history: List[float], window: int = 5000, tol: float = 0.01) -> ConvergenceReport:
    """aero/diagnostics.py:87-120"""
    window = max(int(window), 1)
    if len(history) < window:
        return ConvergenceReport(False, window, float("inf"), 0.0, 0.0, float(tol))
    recent = np.asarray(history[-window:], dtype=np.float64)
    mean_abs = float(np.mean(np.abs(recent)))
    std = float(np.std(recent))
    if mean_abs < 1e-12:
        return ConvergenceReport(True, window, 0.0, mean_abs, std, float(tol))
    ratio = std / mean_abs
    return ConvergenceReport(ratio < tol, window, ratio, mean_abs, std, float(tol))


def check_convergence(Cd_history: List[float], window: int = 5000, tol: float = 0.01) -> bool:
    return analyze_convergence(Cd_history, window=window, tol=tol).converged


def analyze_strouhal(Cl_history: List[float], D: float, u0: float, *, window: Optional[int] = None,
                     min_samples: int = 512) -> StrouhalReport:
    """FFT-peak Strouhal number of the lift history (aero/diagnostics.py:153-194)."""
    if u0 <= 0.0 or D <= 0.0:
        return StrouhalReport(None, None, 0.0, len(Cl_history), int(window or 0))
    samples = np.asarray(Cl_history, dtype=np.float64)
    if window is not None and window > 0:
        samples = samples[-int(window):]
    used = int(samples.size)
    if used < max(int(min_samples), 2):
        return StrouhalReport(None, None, 0.0, len(Cl_history), used)
    signal = samples - float(np.mean(samples))
    peak_amplitude = float(np.max(np.abs(signal)))
    if peak_amplitude < 1e-6:
        return StrouhalReport(None, None, peak_amplitude, len(Cl_history), used)
    mag = np.abs(np.fft.rfft(signal))
    freqs = np.fft.rfftfreq(signal.size, d=1.0)
    if mag.size <= 1:
        return StrouhalReport(None, None, peak_amplitude, len(Cl_history), used)
    k = int(np.argmax(mag[1:])) + 1
    pf = float(freqs[k])
    return StrouhalReport(pf * D / u0, pf, peak_amplitude, len(Cl_history), used)


def compute_strouhal(Cl_history: List[float], D: float, u0: float) -> Optional[float]:
    """aero/diagnostics.py:123-150"""
    return analyze_strouhal(Cl_history, D=D, u0=u0).strouhal


def detect_statistical_stationarity(Cd_history, Cl_history, D, u0, *, convergence_window=5000,
                                    convergence_tol=0.01, strouhal_window=4096, strouhal_tol=0.05):
    """Cd window convergence + Strouhal drift between half windows (aero/diagnostics.py:197-228)."""
    cd = analyze_convergence(Cd_history, window=convergence_window, tol=convergence_tol)
    st = analyze_strouhal(Cl_history, D=D, u0=u0, window=strouhal_window)
    if st.strouhal is None or st.window < max(strouhal_window, 1024):
        return cd, st, cd.converged
    half = st.window // 2
    first = analyze_strouhal(Cl_history[-st.window:-half], D=D, u0=u0, min_samples=128)
    second = analyze_strouhal(Cl_history[-half:], D=D, u0=u0, min_samples=128)
    if first.strouhal is None or second.strouhal is None:
        return cd, st, cd.converged
    drift = abs(second.strouhal - first.strouhal) / max(abs(first.strouhal), 1e-12)
    st.stationary = drift < strouhal_tol
    st.relative_drift = drift
    return cd, st, cd.converged and st.stationary


def validate_parameters(tau: float, u0: float, D: float, Ny: int) -> List[str]:
    """Pre-flight checks (aero/diagnostics.py:231-267)."""
    msgs = []
    Ma = u0 * (3.0 ** 0.5)
    blockage = D / Ny
    if tau <= 0.5:
        msgs.append(f"ERROR: tau={tau:.4f} <= 0.5 — BGK is unconditionally unstable.")
    elif tau < 0.55:
        msgs.append(f"WARNING: tau={tau:.4f} is very close to 0.5. "
                    "Simulation may be marginally stable. Consider reducing Re or increasing u0/D.")
    if Ma > 0.15:
        msgs.append(f"WARNING: Ma={Ma:.4f} > 0.15. Compressibility errors will be significant. Reduce u0.")
    if blockage > 0.40:
        msgs.append(f"WARNING: blockage ratio D/Ny={blockage:.2f} > 40%. "
                    "Wall confinement will significantly inflate Cd. Increase Ny.")
    return msgs


def coefficient_spectrum(history: List[float], *, sample_spacing: float = 1.0, min_samples: int = 128,
                         discard_fraction: float = 0.2, top_n: int = 8) -> Optional[Dict[str, object]]:
    """Compact FFT summary (aero/observables.py:10-43)."""
    values = np.asarray(history, dtype=np.float64)
    if values.size < min_samples:
        return None
    discard = int(values.size * discard_fraction)
    if discard >= values.size - min_samples:
        discard = max(0, values.size - min_samples)
    values = values[discard:]
    signal = values - float(np.mean(values))
    if np.max(np.abs(signal)) < 1e-12:
        return None
    freqs = np.fft.rfftfreq(signal.size, d=sample_spacing)
    amps = np.abs(np.fft.rfft(signal))
    if amps.size <= 1:
        return None
    ranked = (np.argsort(amps[1:])[::-1] + 1)[:top_n]
    dom = int(ranked[0])
    return {"sample_count": int(signal.size), "dominant_frequency": float(freqs[dom]),
            "dominant_amplitude": float(amps[dom]), "frequencies": [float(freqs[k]) for k in ranked],
            "amplitudes": [float(amps[k]) for k in ranked]}
