"""Uniform 2D Yee grid for TE fields and staircase permittivity on its nodes.

Node positions, i and j counting cells from the outer PML corner
(etalon/grid/yee.py:1-12):

    Hz[i, j] at ((i + ½)Δx, (j + ½)Δy)   shape (nx,     ny)
    Ex[i, j] at ((i + ½)Δx,  j      Δy)   shape (nx,     ny + 1)
    Ey[i, j] at ( i      Δx, (j + ½)Δy)   shape (nx + 1, ny)

Arrays are C-ordered with y fastest; API units are µm and s.
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from etalon_amd.constants import C_0

TE_COMPONENTS = ("Ex", "Ey", "Hz")
_HALF_NODES = {"Hz": (True, True), "Ex": (True, False), "Ey": (False, True)}
_TOL = 1e-9  # µm: a node on a structure boundary counts as inside


@dataclass(frozen=True)
class YeeGrid2D:
    """Cell counts, spacings, time step and node coordinates (etalon/grid/yee.py:29-79)."""

    nx: int
    ny: int
    dx: float
    dy: float
    dt: float
    pml_layers: int
    x_min: float
    y_min: float
    z: float
    courant: float

    @property
    def n_cells(self) -> int:
        return self.nx * self.ny

    @property
    def region_cells(self) -> tuple[int, int]:
        return self.nx - 2 * self.pml_layers, self.ny - 2 * self.pml_layers

    def shape(self, component: str) -> tuple[int, int]:
        half_x, half_y = _HALF_NODES[component]
        return (self.nx if half_x else self.nx + 1, self.ny if half_y else self.ny + 1)

    def coords(self, component: str) -> tuple[np.ndarray, np.ndarray]:
        half_x, half_y = _HALF_NODES[component]
        n_x, n_y = self.shape(component)
        x = self.x_min + (np.arange(n_x) + (0.5 if half_x else 0.0)) * self.dx
        y = self.y_min + (np.arange(n_y) + (0.5 if half_y else 0.0)) * self.dy
        return np.asarray(x, dtype=np.float64), np.asarray(y, dtype=np.float64)

    def n_steps(self, run_time: float) -> int:
        return math.ceil(round(run_time / self.dt, 9))


def cells_along(size: float, dl: float) -> int:
    """N = ceil(size / dl) ignoring float noise (etalon/grid/yee.py:82-84)."""
    return math.ceil(round(size / dl, 9))


def courant_time_step(courant: float, dx: float, dy: float) -> float:
    """Δt = courant / (c sqrt(1/Δx² + 1/Δy²)) in s, Δ in µm (etalon/grid/yee.py:87-93)."""
    return courant / (C_0 * math.sqrt(1.0 / dx**2 + 1.0 / dy**2))


def build_yee_grid(sim) -> YeeGrid2D:
    """Region cells plus PML cells outside it (etalon/grid/yee.py:96-115)."""
    if not sim.is_2d:
        raise NotImplementedError("3D grids arrive in M1; set size_z = 0 for a 2D simulation")
    layers = sim.boundary.layers
    n_x = cells_along(sim.size[0], sim.grid.dl)
    n_y = cells_along(sim.size[1], sim.grid.dl)
    dx = sim.size[0] / n_x
    dy = sim.size[1] / n_y
    return YeeGrid2D(
        nx=n_x + 2 * layers,
        ny=n_y + 2 * layers,
        dx=dx,
        dy=dy,
        dt=courant_time_step(sim.courant, dx, dy),
        pml_layers=layers,
        x_min=sim.center[0] - sim.size[0] / 2.0 - layers * dx,
        y_min=sim.center[1] - sim.size[1] / 2.0 - layers * dy,
        z=sim.center[2],
        courant=sim.courant,
    )


def _in_plane(z: float, cz: float, height: float) -> bool:
    return height == 0.0 or abs(z - cz) <= height / 2.0 + _TOL


def _paint(values: np.ndarray, x: np.ndarray, y: np.ndarray, s, z: float) -> None:
    """Write structure ``s`` onto the nodes (x, y) of ``values`` (later ones win)."""
    kind = s.type
    if kind == "Box":
        (cx, cy, cz), (sx, sy, sz) = s.center, s.size
        if abs(z - cz) > sz / 2.0 + _TOL:
            return
        inside = (np.abs(x - cx) <= sx / 2.0 + _TOL) & (np.abs(y - cy) <= sy / 2.0 + _TOL)
        values[inside] = s.medium.permittivity
    elif kind == "Cylinder":
        cx, cy, cz = s.center
        if not _in_plane(z, cz, s.height):
            return
        inside = (x - cx) ** 2 + (y - cy) ** 2 <= (s.radius + _TOL) ** 2
        values[inside] = s.medium.permittivity
    elif kind == "Ring":
        cx, cy, cz = s.center
        if not _in_plane(z, cz, s.height):
            return
        r2 = (x - cx) ** 2 + (y - cy) ** 2
        inside = (r2 <= (s.outer_radius + _TOL) ** 2) & (
            r2 >= max(s.inner_radius - _TOL, 0.0) ** 2
        )
        span = s.angle_stop - s.angle_start
        if span < 360.0:
            ang = np.degrees(np.arctan2(y - cy, x - cx))
            rel = np.mod(ang - s.angle_start, 360.0)
            inside &= rel <= span + 1e-9
        values[inside] = s.medium.permittivity
    elif kind == "PixelGrid":
        (cx, cy, cz), (sx, sy, sz) = s.center, s.size
        if abs(z - cz) > sz / 2.0 + _TOL:
            return
        eps = np.asarray(s.permittivity, dtype=np.float64)
        npx, npy = eps.shape
        x0, y0 = cx - sx / 2.0, cy - sy / 2.0
        inside = (np.abs(x - cx) <= sx / 2.0 + _TOL) & (np.abs(y - cy) <= sy / 2.0 + _TOL)
        p = np.clip(np.floor((x - x0) / (sx / npx)).astype(np.int64), 0, npx - 1)
        q = np.clip(np.floor((y - y0) / (sy / npy)).astype(np.int64), 0, npy - 1)
        values[inside] = eps[p[inside], q[inside]]
    else:  # pragma: no cover - the model union is closed
        raise TypeError(f"unknown structure type {kind!r}")


def _paint_cylinders(values: np.ndarray, xs: np.ndarray, ys: np.ndarray, cyls) -> None:
    """Paint cylinders of one permittivity at once: each one's test (the Cylinder rule
    of ``_paint``, the same arithmetic) runs on the block of nodes around it only, so a
    lattice of thousands of holes costs a few vector operations."""
    cx = np.array([c.center[0] for c in cyls], dtype=np.float64)
    cy = np.array([c.center[1] for c in cyls], dtype=np.float64)
    r = np.array([c.radius for c in cyls], dtype=np.float64)
    reach = float(r.max()) + _TOL
    dx = float(np.min(np.diff(xs))) if xs.size > 1 else 1.0
    dy = float(np.min(np.diff(ys))) if ys.size > 1 else 1.0
    hx, hy = int(np.ceil(reach / dx)) + 1, int(np.ceil(reach / dy)) + 1
    ii = np.clip(np.searchsorted(xs, cx)[:, None] + np.arange(-hx, hx + 1)[None, :], 0, xs.size - 1)
    jj = np.clip(np.searchsorted(ys, cy)[:, None] + np.arange(-hy, hy + 1)[None, :], 0, ys.size - 1)
    inside = ((xs[ii][:, :, None] - cx[:, None, None]) ** 2
              + (ys[jj][:, None, :] - cy[:, None, None]) ** 2) <= (r[:, None, None] + _TOL) ** 2
    n, a, b = np.nonzero(inside)
    values[ii[n, a], jj[n, b]] = cyls[0].medium.permittivity


def permittivity_on_nodes(sim, grid: YeeGrid2D) -> dict[str, np.ndarray]:
    """ε_r at the Ex and Ey nodes, staircased (etalon/grid/materials.py:12-35).
    Consecutive cylinders of one permittivity (a photonic-crystal lattice) are painted
    together; the order among them does not matter, their value being the same."""
    eps: dict[str, np.ndarray] = {}
    structures = list(sim.structures)
    for component in ("Ex", "Ey"):
        xs, ys = grid.coords(component)
        x, y = np.meshgrid(xs, ys, indexing="ij")
        values = np.full(x.shape, sim.medium.permittivity)
        k = 0
        while k < len(structures):
            s = structures[k]
            if s.type != "Cylinder":
                _paint(values, x, y, s, grid.z)
                k += 1
                continue
            m = k + 1
            while (m < len(structures) and structures[m].type == "Cylinder"
                   and structures[m].medium.permittivity == s.medium.permittivity):
                m += 1
            run = [c for c in structures[k:m] if _in_plane(grid.z, c.center[2], c.height)]
            if len(run) == 1:
                _paint(values, x, y, run[0], grid.z)
            elif run:
                _paint_cylinders(values, np.asarray(xs, np.float64), np.asarray(ys, np.float64),
                                 run)
            k = m
        eps[component] = values
    return eps


def pixel_node_map(grid: YeeGrid2D, center, size, shape) -> dict[str, tuple[np.ndarray, ...]]:
    """For a PixelGrid footprint: per E component, the (i, j) nodes inside it and the
    flat pixel index each one takes (the map the permittivity gradient sums over)."""
    (cx, cy, _), (sx, sy, _) = center, size
    npx, npy = shape
    x0, y0 = cx - sx / 2.0, cy - sy / 2.0
    out = {}
    for component in ("Ex", "Ey"):
        xs, ys = grid.coords(component)
        x, y = np.meshgrid(xs, ys, indexing="ij")
        inside = (np.abs(x - cx) <= sx / 2.0 + _TOL) & (np.abs(y - cy) <= sy / 2.0 + _TOL)
        ii, jj = np.nonzero(inside)
        p = np.clip(np.floor((x[ii, jj] - x0) / (sx / npx)).astype(np.int64), 0, npx - 1)
        q = np.clip(np.floor((y[ii, jj] - y0) / (sy / npy)).astype(np.int64), 0, npy - 1)
        out[component] = (ii.astype(np.int64), jj.astype(np.int64), p * npy + q)
    return out
