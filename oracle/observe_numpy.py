"""ORACLE — test infrastructure only; never imported by the etalon_amd package.

CPU restatement of the run observers of etalon/studio/runner.py, the checker of
the HIP reductions (etl_field_energy, etl_downsample):

    energy_sums      the three float64 sums of field_energy, runner.py:258-280
    field_energy     W = ½ ΔxΔy (Δt (S_ex + S_ey) + Δt/db S_hz), runner.py:275-280
    block_means      the block means of downsample, runner.py:283-291, summed in a
                     stated order (below) instead of NumPy's
    downsample_steps the block sizes sx, sy of downsample, runner.py:285-287

What pins it (tests/test_observers.py): the reference test's known answers — the
energy of constant fields from the permittivity map (tests/studio/test_jobs.py:
277-304) and the block shape and first block mean of an arange field (:307-315) —,
exact sums (math.fsum) for the energy, and NumPy's own ``mean`` for the block
means (to float32 rounding).
"""

from __future__ import annotations

import math

import numpy as np

UM = 1e-6
SNAPSHOT_MAX_NODES = 256  # etalon/studio/runner.py:48


def energy_sums(state) -> tuple[float, float, float]:
    """(Σ ex²/cb_ex, Σ ey²/cb_ey, Σ hz²) in float64, each term (v·v)/cb with v widened
    first, summed by NumPy (pairwise) as runner.py:268-274 does."""

    def weighted(field, coefficient):
        values = np.asarray(field).astype(np.float64)
        return float(np.sum(values * values / np.asarray(coefficient)))

    hz = np.asarray(state.hz).astype(np.float64)
    return (weighted(state.ex, state.cb_ex), weighted(state.ey, state.cb_ey),
            float(np.sum(hz * hz)))


def exact_energy_sums(state) -> tuple[float, float, float]:
    """The same three sums, each correctly rounded (math.fsum of the float64 terms)."""

    def terms(field, coefficient=None):
        v = np.asarray(field).astype(np.float64).ravel()
        t = v * v
        if coefficient is not None:
            t = t / np.asarray(coefficient).astype(np.float64).ravel()
        return math.fsum(t)

    return terms(state.ex, state.cb_ex), terms(state.ey, state.cb_ey), terms(state.hz)


def field_energy(sums, dt: float, db: float, dx_um: float, dy_um: float) -> float:
    """runner.py:275-280 from the three sums."""
    electric = dt * (sums[0] + sums[1])
    magnetic = dt / float(db) * sums[2]
    cell_area = dx_um * UM * dy_um * UM
    return float(0.5 * cell_area * (electric + magnetic))


def downsample_steps(nx: int, ny: int, max_nodes: int = SNAPSHOT_MAX_NODES) -> tuple[int, int]:
    """runner.py:285-287."""
    return max(1, math.ceil(nx / max_nodes)), max(1, math.ceil(ny / max_nodes))


def block_means(field, sx: int, sy: int) -> np.ndarray:
    """Block means of ``field`` as float32: each block's values widened to float64 and
    added in (p, q) row-major order within the block, the sum divided by sx·sy in
    float64 and rounded to float32 once (the order the HIP kernel uses)."""
    f = np.asarray(field)
    nx, ny = f.shape
    mx, my = nx // sx, ny // sy
    acc = np.zeros((mx, my), np.float64)
    for p in range(sx):
        for q in range(sy):
            acc = acc + f[p: mx * sx: sx, q: my * sy: sy].astype(np.float64)
    return (acc / float(sx * sy)).astype(np.float32)
