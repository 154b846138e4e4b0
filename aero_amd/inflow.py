"""Host-side inflow plane generators (turbulent inlets).

Both generators are driven by numpy's PCG64 ``default_rng`` with the
reference's seeds and call sequences, so the planes they hand to the device
are the reference's planes:
* synthetic inflow: aero/lbm/solver.py:259-270 (2D), solver3d.py:297-321 (3D);
* SEM: aero/lbm/sem_inlet.py:15-136 (Jarrin et al. 2006 tent-kernel eddies
  with a 2D Fourier divergence-free projection of (v, w)).
"""

from __future__ import annotations

from typing import Tuple

import numpy as np


class SyntheticInflow2D:
    """Filtered-noise uy(y) at the inlet (aero/lbm/solver.py:259-270)."""

    def __init__(self, ny: int, u0: float, intensity: float, seed: int = 12345):
        self.ny, self.u0, self.intensity = ny, float(u0), float(intensity)
        self.rng = np.random.default_rng(int(seed))
        self.field = np.zeros(ny, dtype=np.float64)

    def state(self) -> dict:
        """The filtered field and the random generator's bit state (not in the
        reference's checkpoints; kept so a resumed run continues the same
        inflow)."""
        return dict(field=self.field.tolist(), rng=self.rng.bit_generator.state)

    def set_state(self, st: dict) -> None:
        field = np.array(st["field"], dtype=np.float64)
        if field.shape != self.field.shape:
            raise ValueError("inflow state of a different inlet")
        self.field = field
        self.rng.bit_generator.state = st["rng"]

    def next(self) -> np.ndarray:
        return self.next_n(1)[0]

    def next_n(self, n: int) -> np.ndarray:
        """(n, ny): the next n planes, equal to n calls of next().  The noise
        of all n steps is drawn in one block (numpy's generator gives the same
        values as n successive draws) and filtered at once; only the
        recursion runs per step."""
        noise = self.rng.standard_normal((n, self.ny))
        filtered = (noise + np.roll(noise, 1, axis=1) + np.roll(noise, -1, axis=1)) / 3.0
        out = np.empty((n, self.ny))
        target = self.intensity * self.u0
        for k in range(n):
            self.field = 0.75 * self.field + 0.25 * filtered[k]
            self.field -= float(np.mean(self.field))
            rms = float(np.sqrt(np.mean(self.field ** 2)))
            if rms > 1e-12 and target > 0.0:
                self.field *= target / rms
            out[k] = self.field
        return out


class SyntheticInflow3D:
    """Filtered-noise (uy, uz)(z, y) planes (aero/lbm/solver3d.py:297-321)."""

    def __init__(self, nz: int, ny: int, u0: float, intensity: float, seed: int = 12345):
        self.shape = (nz, ny)
        self.u0, self.intensity = float(u0), float(intensity)
        self.rng = np.random.default_rng(int(seed))
        self.fy = np.zeros(self.shape)
        self.fz = np.zeros(self.shape)

    def state(self) -> dict:
        """The filtered planes and the random generator's bit state (see
        SyntheticInflow2D.state)."""
        return dict(fy=self.fy.tolist(), fz=self.fz.tolist(), rng=self.rng.bit_generator.state)

    def set_state(self, st: dict) -> None:
        fy, fz = np.array(st["fy"], dtype=np.float64), np.array(st["fz"], dtype=np.float64)
        if fy.shape != self.shape or fz.shape != self.shape:
            raise ValueError("inflow state of a different inlet")
        self.fy, self.fz = fy, fz
        self.rng.bit_generator.state = st["rng"]

    def next(self) -> Tuple[np.ndarray, np.ndarray]:
        fy, fz = self.next_n(1)
        return fy[0], fz[0]

    def next_n(self, n: int) -> Tuple[np.ndarray, np.ndarray]:
        """(n, nz, ny) each: the next n (uy, uz) planes, equal to n calls of
        next().  The noise of all n steps is drawn in one block (the
        generator gives the same values in the same order); each step's two
        planes are filtered together."""
        noise = self.rng.standard_normal((n, 2) + self.shape)
        oy, oz = np.empty((n,) + self.shape), np.empty((n,) + self.shape)
        target = self.intensity * self.u0
        for k in range(n):
            a = noise[k]  # both planes of step k, filtered together
            filt = (a + np.roll(a, 1, axis=1) + np.roll(a, -1, axis=1) + np.roll(a, 1, axis=2)
                    + np.roll(a, -1, axis=2)) / 5.0
            self.fy = 0.8 * self.fy + 0.2 * filt[0]
            self.fz = 0.8 * self.fz + 0.2 * filt[1]
            self.fy -= float(np.mean(self.fy))
            self.fz -= float(np.mean(self.fz))
            rms = float(np.sqrt(np.mean(self.fy ** 2 + self.fz ** 2)))
            if rms > 1e-12 and target > 0.0:
                s = target / rms
                self.fy *= s
                self.fz *= s
            oy[k], oz[k] = self.fy, self.fz
        return oy, oz


class SEMInlet:
    """Synthetic Eddy Method inlet fluctuations (aero/lbm/sem_inlet.py:15-136)."""

    def __init__(self, Ny: int, Nz: int, u0: float, Tu: float, L_int: float, N_eddies: int = 200,
                 rng_seed: int = 0) -> None:
        self.Ny, self.Nz = int(Ny), int(Nz)
        self.u0, self.Tu = float(u0), float(Tu)
        self.sigma = float(L_int)
        self.N = int(N_eddies)
        self._r = 3.0 * self.sigma
        self._rng = np.random.default_rng(rng_seed)
        self.xk = self._rng.uniform(-self._r, self._r, self.N)
        self.yk = self._rng.uniform(0.0, float(self.Ny), self.N)
        self.zk = self._rng.uniform(0.0, float(self.Nz), self.N)
        self.eps = self._rng.choice(np.array([-1.0, 1.0]), size=(3, self.N)).astype(np.float64)

    def state(self) -> dict:
        """Everything the next planes depend on: the eddies and the random
        generator's bit state (not in the reference's checkpoints; kept so a
        resumed run continues the same turbulence)."""
        return dict(Ny=self.Ny, Nz=self.Nz, u0=self.u0, Tu=self.Tu, L_int=self.sigma, N=self.N,
                    xk=self.xk.tolist(), yk=self.yk.tolist(), zk=self.zk.tolist(), eps=self.eps.tolist(),
                    rng=self._rng.bit_generator.state)

    def set_state(self, st: dict) -> None:
        if (int(st["Ny"]), int(st["Nz"]), int(st["N"])) != (self.Ny, self.Nz, self.N):
            raise ValueError("SEM state of a different inlet")
        self.xk = np.array(st["xk"], dtype=np.float64)
        self.yk = np.array(st["yk"], dtype=np.float64)
        self.zk = np.array(st["zk"], dtype=np.float64)
        self.eps = np.array(st["eps"], dtype=np.float64).reshape(3, self.N)
        self._rng.bit_generator.state = st["rng"]

    def step(self, dt: float = 1.0) -> None:
        self.xk += self.u0 * dt
        wrap = self.xk > self._r
        n = int(np.sum(wrap))
        if n > 0:
            self.xk[wrap] = -self._r
            self.yk[wrap] = self._rng.uniform(0.0, float(self.Ny), n)
            self.zk[wrap] = self._rng.uniform(0.0, float(self.Nz), n)
            self.eps[:, wrap] = self._rng.choice(np.array([-1.0, 1.0]), size=(3, n)).astype(np.float64)

    def fluctuation(self) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
        zeros = np.zeros((self.Nz, self.Ny), dtype=np.float64)
        if self.Tu <= 0.0 or self.sigma <= 0.0 or self.N == 0:
            return zeros.copy(), zeros.copy(), zeros.copy()
        y = np.arange(self.Ny, dtype=np.float64) + 0.5
        z = np.arange(self.Nz, dtype=np.float64) + 0.5
        ZZ, YY = np.meshgrid(z, y, indexing="ij")
        dy = YY[None, :, :] - self.yk[:, None, None]
        dz = ZZ[None, :, :] - self.zk[:, None, None]
        r = np.sqrt(self.xk[:, None, None] ** 2 + dy ** 2 + dz ** 2)
        fk = np.maximum(0.0, 1.0 - r / self.sigma)
        inv_sqrt_n = 1.0 / np.sqrt(float(self.N))
        u_raw = inv_sqrt_n * (self.eps[0, :, None, None] * fk).sum(axis=0)
        v_raw = inv_sqrt_n * (self.eps[1, :, None, None] * fk).sum(axis=0)
        w_raw = inv_sqrt_n * (self.eps[2, :, None, None] * fk).sum(axis=0)
        v_hat = np.fft.fft2(v_raw)
        w_hat = np.fft.fft2(w_raw)
        if self.Nz % 2 == 0:
            v_hat[self.Nz // 2, :] = 0.0
            w_hat[self.Nz // 2, :] = 0.0
        if self.Ny % 2 == 0:
            v_hat[:, self.Ny // 2] = 0.0
            w_hat[:, self.Ny // 2] = 0.0
        kz = (np.fft.fftfreq(self.Nz) * 2.0 * np.pi)[:, None]
        ky = (np.fft.fftfreq(self.Ny) * 2.0 * np.pi)[None, :]
        k2 = kz ** 2 + ky ** 2
        k2[0, 0] = 1.0
        div = ky * v_hat + kz * w_hat
        v_hat -= ky / k2 * div
        w_hat -= kz / k2 * div
        v_df = np.real(np.fft.ifft2(v_hat))
        w_df = np.real(np.fft.ifft2(w_hat))
        target = self.Tu * self.u0

        def _scale(a, std):
            return a * (target / std) if std > 1e-14 else zeros.copy()

        u_std = float(np.std(u_raw))
        vw_rms = float(np.sqrt(np.mean(v_df ** 2 + w_df ** 2) / 2.0))
        return _scale(u_raw, u_std), _scale(v_df, vw_rms), _scale(w_df, vw_rms)
