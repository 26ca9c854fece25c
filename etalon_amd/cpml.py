"""CPML profiles b, c, 1/κ along one axis (Roden & Gedney 2000).

    ψ ← b ψ + c ∂F,   b = exp(−(σ/κ + α)Δt/ε0),   c = σ(b − 1)/(σκ + κ²α)
    σ = σ_max ρ^m,    κ = 1 + (κ_max − 1)ρ^m,      α = α_max(1 − ρ)
    σ_max = sigma_scale · 0.8(m + 1)/(η0 Δ sqrt(ε_r))

ρ is the normalised depth into the layer at integer (e) and half (h) nodes
(etalon/fdtd/cpml.py:43-92).  The float64 expressions keep the reference's
operation order, so the profiles — and the kernels' inputs cast from them — have
identical bits.  Outside the layer b = 1, c = 0, 1/κ = 1 exactly, which is what
lets the fused kernel skip ψ there.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from etalon_amd.constants import EPSILON_0, ETA_0, UM


@dataclass(frozen=True)
class CPMLAxis:
    """b, c, 1/κ: ``_e`` at integer nodes (n + 1), ``_h`` at half nodes (n)."""

    b_e: np.ndarray
    c_e: np.ndarray
    inv_kappa_e: np.ndarray
    b_h: np.ndarray
    c_h: np.ndarray
    inv_kappa_h: np.ndarray


def pml_depth(n_cells: int, layers: int) -> tuple[np.ndarray, np.ndarray]:
    """ρ at the n + 1 integer nodes and the n half nodes, mirrored on both faces."""
    rho_e = np.zeros(n_cells + 1)
    rho_h = np.zeros(n_cells)
    k = np.arange(layers + 1)
    rho_e[: layers + 1] = (layers - k) / layers
    k = np.arange(layers)
    rho_h[:layers] = (layers - k - 0.5) / layers
    return np.maximum(rho_e, rho_e[::-1]), np.maximum(rho_h, rho_h[::-1])


def _profile(rho, sigma_max, kappa_max, alpha_max, order, dt):
    sigma = sigma_max * rho**order
    kappa = 1.0 + (kappa_max - 1.0) * rho**order
    alpha = alpha_max * (1.0 - rho)
    b = np.exp(-(sigma / kappa + alpha) * dt / EPSILON_0)
    c = np.zeros_like(rho)
    layer = rho > 0.0
    s, k, a = sigma[layer], kappa[layer], alpha[layer]
    c[layer] = s * (b[layer] - 1.0) / (s * k + k**2 * a)
    b[~layer] = 1.0
    return b, c, 1.0 / kappa


def cpml_axis(n_cells: int, spacing: float, dt: float, pml, eps_r: float,
              alpha_max: float = 0.0) -> CPMLAxis:
    """Profiles for ``n_cells`` cells of ``spacing`` µm in a background of ε_r ``eps_r``."""
    order = pml.grading_order
    sigma_max = pml.sigma_scale * 0.8 * (order + 1.0) / (ETA_0 * spacing * UM * math.sqrt(eps_r))
    rho_e, rho_h = pml_depth(n_cells, pml.layers)
    b_e, c_e, ik_e = _profile(rho_e, sigma_max, pml.kappa_max, alpha_max, order, dt)
    b_h, c_h, ik_h = _profile(rho_h, sigma_max, pml.kappa_max, alpha_max, order, dt)
    return CPMLAxis(b_e, c_e, ik_e, b_h, c_h, ik_h)
