"""The dispersion analysis used by the M0 checks, pinned on the CPU by closed forms:
the 1D Yee relation along an axis, the limit of fine grids, the magic time step on
the diagonal, and phase-velocity recovery from synthetic cylindrical waves."""

import math
from pathlib import Path

import numpy as np
import pytest
from scipy.special import hankel1

from etalon_amd.analysis import measured_phase_velocity, numerical_phase_velocity
from etalon_amd.constants import C_0_SI, UM
from etalon_amd.grid import build_yee_grid
from etalon_amd.validation import point2d_simulation


class _G:
    def __init__(self, dl, courant):
        self.dx = self.dy = dl
        self.dt = courant * dl * UM / (C_0_SI * math.sqrt(2.0))


def test_axis_matches_closed_form():
    """Along 0° the relation is 1D: sin(k̃Δ/2)/Δ = sin(ωΔt/2)/(vΔt)."""
    g = _G(0.1, 0.99)
    lam = np.linspace(0.9, 1.8, 19)
    v = numerical_phase_velocity(lam, g, angle=0.0)
    omega = 2 * np.pi * C_0_SI / (lam * UM)
    dx = g.dx * UM
    k = 2.0 / dx * np.arcsin(dx / (C_0_SI * g.dt) * np.sin(omega * g.dt / 2.0))
    np.testing.assert_allclose(v, omega / (k * C_0_SI), rtol=1e-13)
    assert np.all(v < 1)  # the Yee grid is slow on axis


def test_fine_grid_limit_and_anisotropy():
    lam = np.array([1.55])
    coarse, fine = _G(0.1, 0.5), _G(0.001, 0.5)
    assert abs(numerical_phase_velocity(lam, fine, 0.0)[0] - 1) < 1e-5
    assert numerical_phase_velocity(lam, coarse, 45.0)[0] > numerical_phase_velocity(lam, coarse, 0.0)[0]


def test_diagonal_is_exact_at_the_courant_limit():
    """At courant 1 (Δt = Δ/(c√2)) a wave along 45° has no dispersion."""
    g = _G(0.1, 1.0)
    v = numerical_phase_velocity(np.linspace(1.0, 1.8, 9), g, angle=45.0)
    np.testing.assert_allclose(v, 1.0, rtol=1e-12)


class _Mon:
    def __init__(self, wavelength, values):
        self.wavelength = wavelength
        self.values = values


@pytest.mark.parametrize("wave", ["cylindrical", "plane"])
def test_measured_velocity_recovers_a_synthetic_wave(wave):
    lam = np.linspace(1.2, 1.8, 7)
    v_true = 0.987
    k = 2 * np.pi / lam / v_true
    r1, r2 = 1.5, 4.5
    if wave == "cylindrical":
        near, far = hankel1(0, k * r1), hankel1(0, k * r2)
    else:
        near, far = np.exp(1j * k * r1), np.exp(1j * k * r2)
    v = measured_phase_velocity(_Mon(lam, near), _Mon(lam, far), r1, r2, wave=wave)
    np.testing.assert_allclose(v, v_true, rtol=1e-10)


def test_point2d_grid_is_the_meep_grid():
    """500 x 300 cells (8 x 4 µm at resolution 50 plus 50 PML layers) and 8994 steps,
    as in the Meep fixture."""
    sim = point2d_simulation()
    grid = build_yee_grid(sim)
    ref = np.load(Path(__file__).resolve().parent / "golden" / "meep_point2d.npz")
    assert (grid.nx, grid.ny) == (500, 300)
    assert grid.n_steps(sim.run_time) == int(ref["steps"])
