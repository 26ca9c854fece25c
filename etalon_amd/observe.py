"""Run observers computed on the device (etalon/studio/runner.py:258-311).

A progress report of a long run needs the field energy and a small snapshot of Hz.
The reference computes both with its backend's array module (CuPy on the GPU); here
they are HIP reductions, so a report copies three doubles or one ≤ 256 × 256 float32
block to the host instead of whole fields:

* ``field_energy(solver)`` — W = ½ ΔxΔy Σ (ε0 εr E² + μ0 Hz²) per unit length, J/m,
  with ε0 εr from the solver's own cb = Δt/(ε0 εr) and μ0 = Δt/db, the sums in
  float64 (``etl_field_energy``: a fixed reduction order, equal to NumPy's sums to
  float64 rounding);
* ``downsample(field)`` — block means with at most 256 blocks per axis, float32
  (``etl_downsample``);
* ``snapshot(solver)`` — the downsampled Hz with the block-centre coordinates and
  its extent (the data of the reference's snapshot frame, without its wire format).
"""

from __future__ import annotations

import ctypes as C
import math
from typing import Any

import numpy as np

from etalon_amd.constants import UM
from etalon_amd.hip import _lib
from etalon_amd.hip.backend import DeviceArray, _precision_of

SNAPSHOT_MAX_NODES = 256  # blocks per axis of a snapshot (etalon/studio/runner.py:48)


def energy_sums(state) -> np.ndarray:
    """(Σ ex²/cb_ex, Σ ey²/cb_ey, Σ hz²) over the state's nodes, float64, on the device."""
    out = DeviceArray((3,), np.float64)
    _lib.call("etl_field_energy", _precision_of(state.hz.dtype), C.c_void_p(state.ex.ptr),
              C.c_void_p(state.ey.ptr), C.c_void_p(state.hz.ptr), C.c_void_p(state.cb_ex.ptr),
              C.c_void_p(state.cb_ey.ptr), int(state.nx), int(state.ny), C.c_void_p(out.ptr))
    return out.get()


def field_energy(solver: Any) -> float:
    """Discrete field energy per unit length, J/m (etalon/studio/runner.py:258-280):
    W = ½ ΔxΔy (Δt (S_ex + S_ey) + Δt/db S_hz), the same host arithmetic on the three
    device sums."""
    s = energy_sums(solver.state)
    dt = solver.grid.dt
    electric = dt * (float(s[0]) + float(s[1]))
    magnetic = dt / float(solver.state.db) * float(s[2])
    cell_area = solver.grid.dx * UM * solver.grid.dy * UM
    return float(0.5 * cell_area * (electric + magnetic))


def downsample(field: DeviceArray, max_nodes: int = SNAPSHOT_MAX_NODES) -> tuple[DeviceArray, int, int]:
    """Block means of a 2D device field with at most ``max_nodes`` blocks per axis:
    (float32 blocks on the device, sx, sy), as etalon/studio/runner.py:283-291."""
    if field.ndim != 2:
        raise ValueError(f"downsample takes a 2D field, not shape {field.shape}")
    n_x, n_y = field.shape
    sx = max(1, math.ceil(n_x / max_nodes))
    sy = max(1, math.ceil(n_y / max_nodes))
    out = DeviceArray((n_x // sx, n_y // sy), np.float32)
    _lib.call("etl_downsample", _precision_of(field.dtype), C.c_void_p(field.ptr), n_x, n_y,
              sx, sy, C.c_void_p(out.ptr))
    return out, sx, sy


def snapshot(solver: Any, max_nodes: int = SNAPSHOT_MAX_NODES) -> dict:
    """Hz downsampled on the device, then copied: ``data`` (mx, my) float32, block
    centres ``x``, ``y`` (µm), ``extent`` (x0, x1, y0, y1), ``step`` (the content of
    the reference's snapshot frame, etalon/studio/runner.py:294-311)."""
    blocks, sx, sy = downsample(solver.state.hz, max_nodes)
    data = blocks.get()
    mx, my = data.shape
    x, y = solver.grid.coords("Hz")
    xb = x[: mx * sx].reshape(mx, sx).mean(axis=1)
    yb = y[: my * sy].reshape(my, sy).mean(axis=1)
    return {"data": data, "x": xb, "y": yb, "sx": sx, "sy": sy,
            "extent": (float(xb[0]), float(xb[-1]), float(yb[0]), float(yb[-1])),
            "component": "Hz", "step": solver.n}
