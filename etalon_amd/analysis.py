"""Numerical dispersion of the 2D Yee scheme, predicted and measured
(etalon/analysis/dispersion.py:20-94).

Both functions return a phase velocity normalised to the exact one, ṽp / (c/n):
1 means no dispersion error.
"""

from __future__ import annotations

import math
from typing import Literal

import numpy as np
from scipy.optimize import brentq
from scipy.special import hankel1

from etalon_amd.constants import C_0_SI, UM


def numerical_phase_velocity(wavelengths, grid, angle: float = 0.0, n: float = 1.0) -> np.ndarray:
    """ṽp / (c/n) along ``angle`` (degrees) for vacuum wavelengths (µm).

    The 2D Yee relation (Taflove & Hagness, 3rd ed., ch. 4)

        (sin(ωΔt/2) / (vΔt))² = (sin(k̃ cosφ Δx/2) / Δx)² + (sin(k̃ sinφ Δy/2) / Δy)²,

    v = c/n, solved for k̃ by Newton iteration from k = ω/v; returns ω / (k̃ v).
    """
    lam = np.asarray(wavelengths, dtype=np.float64) * UM
    v = C_0_SI / n
    omega = 2.0 * np.pi * C_0_SI / lam
    dt, dx, dy = grid.dt, grid.dx * UM, grid.dy * UM
    hx = math.cos(math.radians(angle)) * dx / 2.0
    hy = math.sin(math.radians(angle)) * dy / 2.0
    rhs = (np.sin(omega * dt / 2.0) / (v * dt)) ** 2
    k = omega / v
    for _ in range(100):
        residual = (np.sin(k * hx) / dx) ** 2 + (np.sin(k * hy) / dy) ** 2 - rhs
        slope = hx * np.sin(2.0 * k * hx) / dx**2 + hy * np.sin(2.0 * k * hy) / dy**2
        delta = residual / slope
        k = k - delta
        if np.all(np.abs(delta) <= 1e-14 * k):
            return np.asarray(omega / (k * v), dtype=np.float64)
    raise ValueError("Yee dispersion relation has no propagating solution for some wavelengths: "
                     "they are below the grid's cut-off (too few cells per wavelength)")


def _hankel_phase(z: np.ndarray) -> np.ndarray:
    """arg H0⁽¹⁾(z) − (z − π/4): the near-field phase of an outgoing cylindrical wave."""
    return np.angle(hankel1(0, z) * np.exp(-1j * (z - np.pi / 4.0)))


def measured_phase_velocity(near, far, r_near: float, r_far: float,
                            wave: Literal["cylindrical", "plane"], n: float = 1.0) -> np.ndarray:
    """ṽp / (c/n) from the DFT phase of the same component at two distances (µm).

    arg(far / near) is matched to k̃ (r_far − r_near), plus, for the cylindrical wave
    of a 2D point source, the difference of the Hankel near-field phases (time
    convention exp(−iωt)); the 2π ambiguity is resolved against the exact
    k = 2πn/λ.  Meaningful along a lattice symmetry direction (0° or 45°) only.
    ``near`` and ``far`` are monitor data with a ``wavelength`` coordinate.
    """
    lam = np.asarray(near.wavelength, dtype=np.float64)
    if not np.array_equal(lam, np.asarray(far.wavelength, dtype=np.float64)):
        raise ValueError("near and far must share the same wavelengths")
    if r_far <= r_near:
        raise ValueError("r_far must be larger than r_near")
    phase = np.angle(np.asarray(far.values) / np.asarray(near.values))
    k_exact = 2.0 * np.pi * n / lam

    def model(k: float) -> float:
        value = k * (r_far - r_near)
        if wave == "cylindrical":
            corr = _hankel_phase(np.array([k * r_far, k * r_near]))
            value += float(corr[0] - corr[1])
        return value

    k_meas = np.empty_like(k_exact)
    for m, (k0, phi) in enumerate(zip(k_exact, phase)):
        target = phi + 2.0 * np.pi * round((model(k0) - phi) / (2.0 * np.pi))
        k_meas[m] = brentq(lambda k, t=target: model(k) - t, 0.7 * k0, 1.3 * k0, xtol=1e-14)
    return k_exact / k_meas
