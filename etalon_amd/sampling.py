"""Monitor / source node selection and DFT phasors (etalon/fdtd/dft.py:1-119).

Per axis: size > 0 takes every node of the component inside [c − s/2, c + s/2]
(faces included within 1e-9 of a cell); size = 0 interpolates linearly between
the two bracketing nodes (one node, weight 1, when on a node).  Point sources use
the size = 0 rule on both axes.  Phasors are exp(+i2πft)Δt in float64.
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Optional

import numpy as np

from etalon_amd.constants import C_0
from etalon_amd.grid import TE_COMPONENTS, YeeGrid2D

_ON_NODE = 1e-9


@dataclass(frozen=True)
class AxisSampling:
    index: np.ndarray
    coords: np.ndarray
    weights: Optional[np.ndarray]
    point: Optional[float]


def axis_sampling(nodes: np.ndarray, center: float, size: float, what: str) -> AxisSampling:
    """Nodes of one uniform axis sampled by the rules above (etalon/fdtd/dft.py:39-66)."""
    spacing = nodes[1] - nodes[0]
    if size > 0.0:
        tol = _ON_NODE * spacing
        lo, hi = center - size / 2.0, center + size / 2.0
        index = np.nonzero((nodes >= lo - tol) & (nodes <= hi + tol))[0].astype(np.int64)
        if index.size == 0:
            raise ValueError(
                f"{what} contains no grid node along this axis (thinner than one cell, "
                f"{spacing} µm); give it size 0 to interpolate instead"
            )
        return AxisSampling(index, nodes[index], None, None)
    position = (center - nodes[0]) / spacing
    i0 = math.floor(position)
    frac = position - i0
    if frac < _ON_NODE:
        index, weights = np.array([i0]), np.array([1.0])
    elif frac > 1.0 - _ON_NODE:
        index, weights = np.array([i0 + 1]), np.array([1.0])
    else:
        index, weights = np.array([i0, i0 + 1]), np.array([1.0 - frac, frac])
    if index[0] < 0 or index[-1] >= nodes.size:
        raise ValueError(f"{what} at {center} µm is outside the grid")
    index = index.astype(np.int64)
    return AxisSampling(index, nodes[index], weights, center)


@dataclass(frozen=True)
class ComponentSampling:
    """The nodes one monitor records for one component (etalon/fdtd/dft.py:69-89)."""

    monitor: str
    component: str
    wavelengths: np.ndarray
    x: AxisSampling
    y: AxisSampling
    flat_index: np.ndarray

    @property
    def frequencies(self) -> np.ndarray:
        return C_0 / self.wavelengths

    @property
    def is_magnetic(self) -> bool:
        return self.component.startswith("H")


def monitor_samplings(monitor, grid: YeeGrid2D) -> list[ComponentSampling]:
    """One sampling per recorded component; None records Ex, Ey, Hz (dft.py:92-107)."""
    components = monitor.components or TE_COMPONENTS
    wavelengths = np.asarray(monitor.wavelengths, dtype=np.float64)
    out = []
    for component in components:
        xn, yn = grid.coords(component)
        what = f"monitor {monitor.name!r} ({component})"
        sx = axis_sampling(xn, monitor.center[0], monitor.size[0], what)
        sy = axis_sampling(yn, monitor.center[1], monitor.size[1], what)
        n_y = grid.shape(component)[1]
        flat = (sx.index[:, None] * n_y + sy.index[None, :]).ravel()
        out.append(ComponentSampling(monitor.name, component, wavelengths, sx, sy,
                                     flat.astype(np.int64)))
    return out


def dft_phasors(frequencies: np.ndarray, times: np.ndarray, dt: float) -> np.ndarray:
    """exp(+i 2π f t) Δt for every (time, frequency), shape (n_t, n_f), float64."""
    phase = 2.0 * np.pi * times[:, None] * frequencies[None, :]
    return np.asarray(np.exp(1j * phase) * dt, dtype=np.complex128)
