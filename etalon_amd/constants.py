"""Physical constants (CODATA 2018) and unit conversions.

Lengths in the API are micrometres, time is seconds, frequency hertz; the solver
internals are SI.  Values follow etalon/constants.py:13-29 of the reference: c is
exact, μ0 = 1.25663706212e-6 H/m, ε0 = 1/(μ0 c²), η0 = μ0 c.  The expressions are
the same so that every derived coefficient has the same float64 bits.
"""

from __future__ import annotations

C_0_SI: float = 299_792_458.0
"""Speed of light in vacuum, m/s (exact)."""

C_0: float = C_0_SI * 1e6
"""Speed of light in vacuum in API units, µm/s."""

MU_0: float = 1.25663706212e-6
"""Vacuum permeability, H/m."""

EPSILON_0: float = 1.0 / (MU_0 * C_0_SI**2)
"""Vacuum permittivity, F/m."""

ETA_0: float = MU_0 * C_0_SI
"""Impedance of free space, Ω."""

UM: float = 1e-6
"""Metres per micrometre."""


def wvl_to_freq(wavelength):
    """Frequency in Hz of a vacuum wavelength in µm: f = c / λ."""
    return C_0 / wavelength


def freq_to_wvl(frequency):
    """Vacuum wavelength in µm of a frequency in Hz: λ = c / f."""
    return C_0 / frequency
