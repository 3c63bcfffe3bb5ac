"""Seeded stationary random fields on the fault plane.

Spectral synthesis: white noise from numpy's PCG64 generator, filtered in
the Fourier domain by the square root of the target power spectrum, then
scaled to the requested standard deviation.  Computed in float64; callers
cast to the wavefield precision.

* Gaussian covariance  C(r) = s^2 exp(-r^2/a^2)
  -> P(k) ~ exp(-k^2 a^2 / 4)
* von Karman covariance with Hurst exponent H (2D)
  -> P(k) ~ (1 + k^2 a^2)^-(H+1)

The reference code has no random-field generator; its benchmark setups
(src/fd3d_init.f90:535-617) use rectangular patches.  These fields are the
seeded heterogeneity that the workload definitions ask for.
"""
from __future__ import annotations

import numpy as np


def _wavenumbers(n: int, d: float) -> np.ndarray:
    return 2.0 * np.pi * np.fft.fftfreq(n, d=d)


def random_field(nx: int, nz: int, dh: float, corr: float, std: float, seed: int,
                 kind: str = "gaussian", hurst: float = 0.75) -> np.ndarray:
    """Return an (nx, nz) float64 field with zero mean and standard deviation std.

    Index order is (x, z) with x fastest when raveled in Fortran order,
    matching the reference's 2D fault arrays (nxt, nzt).
    """
    if nx < 1 or nz < 1:
        raise ValueError("field dimensions must be positive")
    rng = np.random.Generator(np.random.PCG64(seed))
    noise = rng.standard_normal((nx, nz))
    kx = _wavenumbers(nx, dh)[:, None]
    kz = _wavenumbers(nz, dh)[None, :]
    k2 = kx * kx + kz * kz
    if kind == "gaussian":
        amp = np.exp(-k2 * corr * corr / 8.0)          # sqrt of exp(-k^2 a^2/4)
    elif kind == "vonkarman":
        amp = (1.0 + k2 * corr * corr) ** (-(hurst + 1.0) / 2.0)
    else:
        raise ValueError(f"unknown field kind {kind!r}")
    f = np.real(np.fft.ifft2(np.fft.fft2(noise) * amp))
    f -= f.mean()
    s = f.std()
    if s > 0:
        f *= std / s
    return f
