"""SimulationResult and monitor data without xarray (etalon/io/result.py:15-91).

``result["probe"]["Hz"]`` is a ``MonitorData``: a labelled complex array with the
dims, coords and attrs that the reference's ``xarray.DataArray`` carries —
``wavelength`` last, an ``x``/``y`` dimension per monitor axis with size > 0
holding that component's node coordinates, interpolated axes as scalar coords.
Interpolation of size-0 axes happens once here, in float64, and the result is cast
back to the accumulator dtype, as in ``monitor_dataarray`` (result.py:54-91).
"""

from __future__ import annotations

from typing import Any

import numpy as np

from etalon_amd.constants import C_0


class MonitorData:
    """Complex monitor values with named dims and coordinates (DataArray subset)."""

    def __init__(self, values: np.ndarray, dims: tuple[str, ...], coords: dict[str, Any],
                 name: str, attrs: dict[str, Any]):
        self.values = values
        self.dims = tuple(dims)
        self.coords = coords
        self.name = name
        self.attrs = attrs

    @property
    def dtype(self):
        return self.values.dtype

    @property
    def shape(self):
        return self.values.shape

    def __getattr__(self, item: str):
        coords = self.__dict__.get("coords", {})
        if item in coords:
            c = coords[item]
            return np.asarray(c[1] if isinstance(c, tuple) else c)
        raise AttributeError(item)

    def isel(self, **indexers: int) -> "MonitorData":
        values = self.values
        dims = list(self.dims)
        coords = dict(self.coords)
        for dim, k in indexers.items():
            axis = dims.index(dim)
            values = np.take(values, k, axis=axis)
            for name, c in list(coords.items()):
                if isinstance(c, tuple) and c[0] == dim:
                    coords[name] = np.asarray(c[1])[k]
            dims.pop(axis)
        return MonitorData(values, tuple(dims), coords, self.name, self.attrs)

    def __array__(self, dtype=None, copy=None):
        return np.asarray(self.values, dtype=dtype)

    def __repr__(self) -> str:
        return f"MonitorData({self.name}, dims={self.dims}, shape={self.values.shape})"


class SimulationResult:
    """What ``Simulation.run()`` returns (etalon/io/result.py:15-51)."""

    def __init__(self, simulation, grid, data: dict, metadata: dict):
        self.simulation = simulation
        self.grid = grid
        self.data = data
        self.metadata = metadata

    @property
    def monitors(self) -> list[str]:
        return list(self.data)

    def __getitem__(self, name: str) -> dict:
        try:
            return self.data[name]
        except KeyError:
            raise KeyError(f"no monitor named {name!r}; monitors: {self.monitors}") from None

    def __repr__(self) -> str:
        m = self.metadata
        return (f"SimulationResult(monitors={self.monitors}, backend={m['backend']!r}, "
                f"steps={m['n_steps']}, wall_time={m['wall_time_s']:.3g} s)")


def monitor_data(sampling, acc: np.ndarray, z: float) -> MonitorData:
    """Accumulator (n_f, n_nodes) → labelled data, size-0 axes interpolated in float64."""
    n_f = sampling.wavelengths.size
    data = acc.reshape(n_f, sampling.x.index.size, sampling.y.index.size).astype(np.complex128)
    dims: list[str] = []
    coords: dict[str, Any] = {}
    if sampling.x.weights is None:
        dims.append("x")
        coords["x"] = ("x", sampling.x.coords)
    else:
        data = np.einsum("fab,a->fb", data, sampling.x.weights)
        coords["x"] = sampling.x.point
    if sampling.y.weights is None:
        dims.append("y")
        coords["y"] = ("y", sampling.y.coords)
    else:
        data = data @ sampling.y.weights
        coords["y"] = sampling.y.point
    coords["z"] = z
    coords["wavelength"] = ("wavelength", sampling.wavelengths)
    coords["frequency"] = ("wavelength", C_0 / sampling.wavelengths)
    data = np.moveaxis(data, 0, -1).astype(acc.dtype)
    return MonitorData(data, (*dims, "wavelength"), coords, sampling.component, {
        "monitor": sampling.monitor,
        "units": "A·s/m" if sampling.is_magnetic else "V·s/m",
        "length_units": "um",
        "time_convention": "exp(-i omega t)",
    })


def monitor_geometry(size) -> str:
    """How many axes of a monitor have size > 0: point, line, plane or box
    (etalon/studio/runner.py:359-362)."""
    return ("point", "line", "plane", "box")[sum(1 for s in size if s > 0.0)]


def monitor_layout(sim, result: SimulationResult) -> tuple[dict[str, np.ndarray], list[dict]]:
    """The arrays of ``result.npz`` and the monitor list of ``meta.json``
    (etalon/studio/runner.py:314-356).

    Keys: ``m<i>.<C>`` complex data shaped (n_x, n_y, n_wavelengths), with n = 1 along
    an interpolated axis; ``m<i>.<C>.x`` and ``.y`` its node coordinates (µm, float64);
    ``m<i>.wavelengths`` (µm).  ``i`` is the monitor's index in ``sim.monitors``.  Each
    list entry: name, index, geometry, components, wavelengths_um, and for lines and
    planes the first component's (n_x, n_y) shape and [x0, x1, y0, y1] extent.
    """
    arrays: dict[str, np.ndarray] = {}
    monitors = []
    for i, monitor in enumerate(sim.monitors):
        components = result[monitor.name]
        first = None
        for component, data in components.items():
            values = np.asarray(data.values)
            if "x" not in data.dims:
                values = values[np.newaxis, ...]
            if "y" not in data.dims:
                values = values[:, np.newaxis, ...]
            x = np.atleast_1d(np.asarray(data.x, dtype=np.float64))
            y = np.atleast_1d(np.asarray(data.y, dtype=np.float64))
            arrays[f"m{i}.{component}"] = values
            arrays[f"m{i}.{component}.x"] = x
            arrays[f"m{i}.{component}.y"] = y
            if first is None:
                first = ([int(values.shape[0]), int(values.shape[1])],
                         [float(x[0]), float(x[-1]), float(y[0]), float(y[-1])])
        arrays[f"m{i}.wavelengths"] = np.asarray(monitor.wavelengths, dtype=np.float64)
        geometry = monitor_geometry(monitor.size)
        point = geometry == "point"
        monitors.append({
            "name": monitor.name,
            "index": i,
            "geometry": geometry,
            "components": list(components),
            "wavelengths_um": [float(w) for w in monitor.wavelengths],
            "shape": None if point or first is None else first[0],
            "extent": None if point or first is None else first[1],
        })
    return arrays, monitors


def save_result(folder, sim, result: SimulationResult) -> dict:
    """Write ``result.npz`` (the arrays of ``monitor_layout``, replaced atomically) and
    ``meta.json`` (steps, wall time, cell-updates/s, the monitor list) into ``folder``,
    as a finished run's folder holds them (etalon/studio/runner.py:555-570); returns
    the metadata written."""
    import json
    import os
    from pathlib import Path

    folder = Path(folder)
    folder.mkdir(parents=True, exist_ok=True)
    arrays, monitors = monitor_layout(sim, result)
    tmp = folder / "result.tmp.npz"
    np.savez(tmp, **arrays)
    os.replace(tmp, folder / "result.npz")
    m = result.metadata
    wall = m["wall_time_s"]
    meta = {"steps": int(m["n_steps"]),
            "wall_time_s": float(wall) if wall is not None and np.isfinite(wall) else None,
            "cells_per_s": m["cell_updates_per_s"], "backend": m["backend"],
            "device": m["device"], "precision": m["precision"], "monitors": monitors,
            "has_result": True}
    (folder / "meta.json").write_text(json.dumps(meta, indent=1), encoding="utf-8")
    return meta
