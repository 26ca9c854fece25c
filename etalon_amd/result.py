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
