"""Numerical dispersion of the 2D Yee scheme, predicted and measured — the two
quantities whose ratio the M0 dispersion checks bound (etalon/analysis/dispersion.py:
20-94 defines them; the thresholds of those checks only mean the same thing if these
are the same quantities, computed here in this module's own way).

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

    The numerical wavenumber k̃ of a plane wave exp(i(k̃·r − ωt)) on the 2D Yee grid
    satisfies (Taflove & Hagness, 3rd ed., ch. 4)

        G(k̃) = (sin(k̃ a)/Δx)² + (sin(k̃ b)/Δy)² − (sin(ωΔt/2)/(vΔt))² = 0,

    a = cosφ Δx/2, b = sinφ Δy/2, v = c/n.  G rises monotonically from G(0) < 0 up to
    the first quarter period of the faster-varying term, k* = π/(2 max(|a|, |b|)); the
    propagating root is the one on [0, k*], found by bracketing (Brent) to machine
    precision.  No root there (G(k*) < 0) means the wavelength is below the grid's
    cut-off.  Returns ω / (k̃ v).
    """
    lam = np.atleast_1d(np.asarray(wavelengths, dtype=np.float64)) * UM
    v = C_0_SI / n
    dt, dx, dy = grid.dt, grid.dx * UM, grid.dy * UM
    a = abs(math.cos(math.radians(angle))) * dx / 2.0
    b = abs(math.sin(math.radians(angle))) * dy / 2.0
    k_top = math.pi / (2.0 * max(a, b))
    out = np.empty_like(lam)
    for m, wl in enumerate(lam):
        omega = 2.0 * math.pi * C_0_SI / wl
        rhs = (math.sin(omega * dt / 2.0) / (v * dt)) ** 2

        def gap(k: float, rhs=rhs) -> float:
            return (math.sin(k * a) / dx) ** 2 + (math.sin(k * b) / dy) ** 2 - rhs

        if gap(k_top) < 0.0:
            raise ValueError("Yee dispersion relation has no propagating solution for some "
                             "wavelengths: they are below the grid's cut-off (too few cells "
                             "per wavelength)")
        k = brentq(gap, 0.0, k_top, xtol=1e-300, rtol=4 * np.finfo(float).eps, maxiter=500)
        out[m] = omega / (k * v)
    return out.reshape(np.shape(wavelengths))


def _outgoing_wave_phase(kr: np.ndarray) -> np.ndarray:
    """Phase of the outgoing cylindrical wave H0⁽¹⁾(kr) (time convention exp(−iωt)),
    continuous in kr: its far-field part kr − π/4 plus the near-field remainder."""
    kr = np.asarray(kr, dtype=np.float64)
    remainder = np.angle(hankel1(0, kr) * np.exp(-1j * (kr - np.pi / 4.0)))
    return kr - np.pi / 4.0 + remainder


def measured_phase_velocity(near, far, r_near: float, r_far: float,
                            wave: Literal["cylindrical", "plane"], n: float = 1.0) -> np.ndarray:
    """ṽp / (c/n) from the DFT phase of the same component at two distances (µm).

    The ratio far/near removes the source spectrum and the DFT sampling; its phase is
    the wave's phase gain between the two points, k̃ (r_far − r_near) for a plane
    wave and the gain of H0⁽¹⁾(k̃ r) for the cylindrical wave of a 2D point source.
    The measured angle is placed on the 2π branch nearest the exact wave's gain
    (k = 2πn/λ) and the gain equation solved for k̃ within ±30 % of k.  Meaningful
    along a lattice symmetry direction (0° or 45°) only.  ``near`` and ``far`` are
    monitor data with a ``wavelength`` coordinate.
    """
    lam = np.asarray(near.wavelength, dtype=np.float64)
    if not np.array_equal(lam, np.asarray(far.wavelength, dtype=np.float64)):
        raise ValueError("near and far must share the same wavelengths")
    if r_far <= r_near:
        raise ValueError("r_far must be larger than r_near")
    spacing = r_far - r_near
    angles = np.angle(np.asarray(far.values) / np.asarray(near.values))

    if wave == "cylindrical":
        def gain(k: float) -> float:
            phases = _outgoing_wave_phase(np.array([k * r_far, k * r_near]))
            return float(phases[0] - phases[1])
    else:
        def gain(k: float) -> float:
            return k * spacing

    k_exact = 2.0 * np.pi * n / lam
    ratio = np.empty_like(k_exact)
    for m, (k0, angle) in enumerate(zip(k_exact, angles)):
        turns = round((gain(k0) - angle) / (2.0 * np.pi))
        goal = angle + 2.0 * np.pi * turns
        k_meas = brentq(lambda k: gain(k) - goal, 0.7 * k0, 1.3 * k0, xtol=1e-14)
        ratio[m] = k0 / k_meas
    return ratio
