"""Source signals and point-source injection (etalon/fdtd/sources.py:1-83).

    E^(n+1) = E^n + (Δt/ε)(∇×H)^(n+½) − (Δt/ε) J^(n+½)
    H^(n+½) = H^(n−½) − (Δt/μ)(∇×E)^n − (Δt/μ) M^n

J is sampled at (n + ½)Δt and M at nΔt.  A point source of strength A·g(t)
injects A·g(t)/(ΔxΔy) shared bilinearly between the nearest component nodes.
The per-step amount is computed on the host in float64 exactly as the reference
does, then cast once to the run's precision.
"""

from __future__ import annotations

import functools
from dataclasses import dataclass

import numpy as np

from etalon_amd.constants import EPSILON_0, MU_0, UM
from etalon_amd.grid import YeeGrid2D
from etalon_amd.sampling import axis_sampling


def pulse_signal(pulse, t):
    """g(t) = A exp(−(t − t0)²/(2τ²)) sin(2π f0 (t − t0)); t in s."""
    s = np.asarray(t, dtype=np.float64) - pulse.delay
    envelope = np.exp(-(s**2) / (2.0 * pulse.width**2))
    return pulse.amplitude * envelope * np.sin(2.0 * np.pi * pulse.frequency * s)


@dataclass(frozen=True)
class PointInjection:
    component: str
    i: np.ndarray
    j: np.ndarray
    weights: np.ndarray
    coefficient: np.ndarray
    pulse: object

    @property
    def is_magnetic(self) -> bool:
        return self.component.startswith("H")

    def sample_time(self, n: int, dt: float) -> float:
        return n * dt if self.is_magnetic else (n + 0.5) * dt

    def delta(self, n: int, dt: float) -> np.ndarray:
        """Amount subtracted from the field at the nodes in step n (float64)."""
        return self.coefficient * pulse_signal(self.pulse, self.sample_time(n, dt))

    def delta_table(self, n0: int, rows: int, dt: float) -> np.ndarray:
        """delta(n) for n in [n0, n0 + rows), shape (rows, nodes), float64: the
        coefficients times the pulse column ``signal_column`` (the same product of
        the same float64 numbers as ``delta(n)``)."""
        g = signal_column(self.pulse, n0, rows, dt, self.is_magnetic)
        return self.coefficient[None, :] * g[:, None]


def _pulse_key(pulse) -> tuple:
    return (float(pulse.amplitude), float(pulse.delay), float(pulse.width),
            float(pulse.frequency))


@functools.lru_cache(maxsize=32)
def _signal_column_cached(key: tuple, n0: int, rows: int, dt: float, magnetic: bool) -> np.ndarray:
    amplitude, delay, width, frequency = key
    pulse = _Pulse(amplitude, delay, width, frequency)
    steps = np.arange(n0, n0 + rows, dtype=np.float64)
    t = steps * dt if magnetic else (steps + 0.5) * dt
    s = t - delay
    # Per step, ``s`` is a NumPy float64 scalar and ``s**2`` is the scalar power (C
    # ``pow``), which can differ by one ulp from the array square, so the squares are
    # taken per element the same way.
    s2 = np.fromiter((x**2 for x in s), dtype=np.float64, count=s.size)
    g = amplitude * np.exp(-s2 / (2.0 * width**2)) * np.sin(2.0 * np.pi * frequency * s)
    # Whether exp and sin give the same bits for an array as for one element depends
    # on the NumPy build and the CPU: sampled rows are checked against the per-step
    # evaluation and any difference switches to one evaluation per step.
    sample = sorted(set(range(min(rows, 16))) | set(range(0, rows, max(1, rows // 64)))
                    | ({rows - 1} if rows else set()))
    if not all(g[r].view(np.int64) == pulse_signal(pulse, t[r]).view(np.int64) for r in sample):
        g = np.array([pulse_signal(pulse, t[r]) for r in range(rows)], dtype=np.float64)
    g.setflags(write=False)
    return g


@dataclass(frozen=True)
class _Pulse:
    amplitude: float
    delay: float
    width: float
    frequency: float


def signal_column(pulse, n0: int, rows: int, dt: float, magnetic: bool) -> np.ndarray:
    """g(t_n) for n in [n0, n0 + rows) at the injection times of an H (n Δt) or E
    ((n + ½) Δt) source, float64, read-only; equal bit for bit to ``pulse_signal``
    at each single time.  Cached: the simulations of a batch that share a pulse and
    Δt share the column."""
    return _signal_column_cached(_pulse_key(pulse), int(n0), int(rows), float(dt), bool(magnetic))


def point_injection(source, grid: YeeGrid2D, eps_r: dict) -> PointInjection:
    """Resolve a point source onto its component's nodes (etalon/fdtd/sources.py:59-83)."""
    x_nodes, y_nodes = grid.coords(source.component)
    what = f"source at {source.center}"
    sx = axis_sampling(x_nodes, source.center[0], 0.0, what)
    sy = axis_sampling(y_nodes, source.center[1], 0.0, what)
    i = np.repeat(sx.index, sy.index.size)
    j = np.tile(sy.index, sx.index.size)
    weights = np.outer(sx.weights, sy.weights).ravel()
    area = (grid.dx * UM) * (grid.dy * UM)
    if source.component.startswith("H"):
        update = np.full(weights.shape, grid.dt / MU_0)
    else:
        update = grid.dt / (EPSILON_0 * eps_r[source.component][i, j])
    return PointInjection(source.component, i, j, weights, update * weights / area, source.pulse)
