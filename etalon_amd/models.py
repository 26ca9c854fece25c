"""User-facing models: the reference's API (etalon/core/*, etalon/geometry/box.py).

Every model is frozen, rejects unknown keywords and NaN/inf, and round-trips
through JSON, as in etalon/core/base.py:31-39.  The fields, defaults and
validation messages follow the reference so that a simulation written for it
builds here unchanged; ``Simulation.backend`` takes ``"auto"`` or ``"hip"``.

Geometry beyond ``Box`` (needed by the waveguide, ring, photonic-crystal and
inverse-design workloads) is added as ``Cylinder``, ``Ring`` and ``PixelGrid``;
all use the reference's staircase rule (a node takes the medium of the last
structure containing it, boundaries included within 1e-9 µm;
etalon/grid/materials.py:10-35).
"""

from __future__ import annotations

import math
from typing import Annotated, Any, Literal, Optional, Union

import numpy as np
from pydantic import (
    BaseModel,
    BeforeValidator,
    ConfigDict,
    Field,
    NonNegativeFloat,
    PositiveFloat,
    field_validator,
    model_validator,
)

from etalon_amd.constants import C_0


def _array_to_list(value: Any) -> Any:
    if hasattr(value, "tolist"):
        return value.tolist()
    return value


class EtalonModel(BaseModel):
    """Immutable, typo-proof, JSON-safe base (etalon/core/base.py:31-39)."""

    model_config = ConfigDict(frozen=True, extra="forbid", allow_inf_nan=False)


Coordinate = Annotated[tuple[float, float, float], BeforeValidator(_array_to_list)]
Size = Annotated[
    tuple[NonNegativeFloat, NonNegativeFloat, NonNegativeFloat], BeforeValidator(_array_to_list)
]
Wavelengths = Annotated[
    tuple[PositiveFloat, ...], BeforeValidator(_array_to_list), Field(min_length=1)
]
FieldComponent = Literal["Ex", "Ey", "Ez", "Hx", "Hy", "Hz"]


class Medium(EtalonModel):
    """Lossless, non-dispersive, isotropic medium: ε_r = n² (etalon/core/medium.py:10-37)."""

    type: Literal["Medium"] = Field("Medium", repr=False)
    n: float = 1.0
    name: Optional[str] = None

    @field_validator("n")
    @classmethod
    def _index_at_least_one(cls, n: float) -> float:
        if n < 1.0:
            raise ValueError(
                f"refractive index n = {n} is below 1; a lossless, non-dispersive medium "
                "with n < 1 is superluminal and makes the FDTD time step unstable"
            )
        return n

    @property
    def permittivity(self) -> float:
        return self.n**2


class PML(EtalonModel):
    """CPML on every face (etalon/core/boundary.py:10-42); layers are outside the region."""

    type: Literal["PML"] = Field("PML", repr=False)
    layers: int = Field(12, ge=1)
    grading_order: float = Field(3.0, ge=1.0, le=6.0)
    sigma_scale: float = Field(1.0, gt=0.0)
    kappa_max: float = Field(1.0, ge=1.0)


class UniformGrid(EtalonModel):
    """N = ceil(size / dl) cells per axis of Δ = size / N (etalon/core/grid_spec.py:12-29)."""

    type: Literal["UniformGrid"] = Field("UniformGrid", repr=False)
    dl: PositiveFloat


class GaussianPulse(EtalonModel):
    """g(t) = A exp(−(t − t0)²/(2τ²)) sin(2π f0 (t − t0)) (etalon/core/source.py:12-68)."""

    type: Literal["GaussianPulse"] = Field("GaussianPulse", repr=False)
    wavelength: PositiveFloat
    bandwidth: PositiveFloat
    amplitude: float = 1.0

    @model_validator(mode="after")
    def _band_stays_positive(self) -> "GaussianPulse":
        if self.bandwidth >= 2.0 * self.wavelength:
            raise ValueError(
                f"bandwidth {self.bandwidth} µm must be smaller than 2 × wavelength "
                f"({2.0 * self.wavelength} µm), otherwise the band reaches λ ≤ 0"
            )
        return self

    @property
    def frequency(self) -> float:
        return C_0 / self.wavelength

    @property
    def frequency_width(self) -> float:
        f_hi = C_0 / (self.wavelength - self.bandwidth / 2.0)
        f_lo = C_0 / (self.wavelength + self.bandwidth / 2.0)
        return (f_hi - f_lo) / 2.0

    @property
    def width(self) -> float:
        return 1.0 / (2.0 * math.pi * self.frequency_width)

    @property
    def delay(self) -> float:
        return 5.0 * self.width


class PointSource(EtalonModel):
    """Point dipole current, bilinear over the nearest nodes (etalon/core/source.py:71-97)."""

    type: Literal["PointSource"] = Field("PointSource", repr=False)
    center: Coordinate
    component: FieldComponent
    pulse: GaussianPulse
    name: Optional[str] = None


class DFTMonitor(EtalonModel):
    """Running DFT F(f) = Σ F(t_n) exp(+i2πf t_n) Δt (etalon/core/monitor.py:10-52)."""

    type: Literal["DFTMonitor"] = Field("DFTMonitor", repr=False)
    name: str = Field(min_length=1)
    center: Coordinate = (0.0, 0.0, 0.0)
    size: Size = (0.0, 0.0, 0.0)
    wavelengths: Wavelengths
    components: Optional[tuple[FieldComponent, ...]] = None

    @field_validator("components")
    @classmethod
    def _components_unique(cls, components):
        if components is not None:
            if len(components) == 0:
                raise ValueError("components must name at least one component, or be None")
            if len(set(components)) != len(components):
                raise ValueError(f"components contains repeats: {components}")
        return components


class Box(EtalonModel):
    """Axis-aligned block (etalon/geometry/box.py:11-32)."""

    type: Literal["Box"] = Field("Box", repr=False)
    center: Coordinate = (0.0, 0.0, 0.0)
    size: Size
    medium: Medium
    name: Optional[str] = None


class Cylinder(EtalonModel):
    """Disk of ``radius`` µm around ``center`` (z-extent ``height``, 0 = infinite)."""

    type: Literal["Cylinder"] = Field("Cylinder", repr=False)
    center: Coordinate = (0.0, 0.0, 0.0)
    radius: PositiveFloat
    height: NonNegativeFloat = 0.0
    medium: Medium
    name: Optional[str] = None


class Ring(EtalonModel):
    """Annulus inner_radius ≤ r ≤ outer_radius, optionally restricted to the angular
    sector [angle_start, angle_stop] (degrees, counter-clockwise from +x): a ring
    resonator, or a circular waveguide bend."""

    type: Literal["Ring"] = Field("Ring", repr=False)
    center: Coordinate = (0.0, 0.0, 0.0)
    inner_radius: NonNegativeFloat
    outer_radius: PositiveFloat
    angle_start: float = 0.0
    angle_stop: float = 360.0
    height: NonNegativeFloat = 0.0
    medium: Medium
    name: Optional[str] = None

    @model_validator(mode="after")
    def _radii(self) -> "Ring":
        if self.inner_radius >= self.outer_radius:
            raise ValueError("inner_radius must be smaller than outer_radius")
        return self


class PixelGrid(EtalonModel):
    """Pixelated permittivity over the box center ± size/2: ``permittivity[p][q]``
    is ε_r of pixel (p, q), p along x.  A node inside the box (faces included)
    takes the ε of the pixel containing it (the last pixel on the upper faces).
    This is the design region of inverse design."""

    type: Literal["PixelGrid"] = Field("PixelGrid", repr=False)
    center: Coordinate = (0.0, 0.0, 0.0)
    size: Size
    permittivity: Annotated[tuple[tuple[PositiveFloat, ...], ...], BeforeValidator(_array_to_list)]
    name: Optional[str] = None

    @model_validator(mode="after")
    def _shape(self) -> "PixelGrid":
        rows = {len(r) for r in self.permittivity}
        if not self.permittivity or len(rows) != 1 or 0 in rows:
            raise ValueError("permittivity must be a non-empty 2D array")
        if self.size[0] <= 0.0 or self.size[1] <= 0.0:
            raise ValueError("a PixelGrid needs size_x > 0 and size_y > 0")
        return self

    @property
    def shape(self) -> tuple[int, int]:
        return len(self.permittivity), len(self.permittivity[0])


Structure = Annotated[Union[Box, Cylinder, Ring, PixelGrid], Field(discriminator="type")]


class Simulation(EtalonModel):
    """One 2D TE FDTD run (etalon/core/simulation.py:19-125).

    Δt = courant / (c sqrt(1/Δx² + 1/Δy²)); PML layers are added outside ``size``.
    """

    center: Coordinate = (0.0, 0.0, 0.0)
    size: Size
    medium: Medium = Medium()
    structures: tuple[Structure, ...] = ()
    grid: UniformGrid
    boundary: PML = PML()
    sources: tuple[PointSource, ...] = Field(min_length=1)
    monitors: tuple[DFTMonitor, ...] = ()
    run_time: PositiveFloat
    courant: float = Field(0.99, gt=0.0, le=1.0)
    precision: Literal["single", "double"] = "single"
    backend: Literal["auto", "hip"] = "auto"

    @property
    def is_2d(self) -> bool:
        return self.size[2] == 0.0

    @model_validator(mode="after")
    def _check_layout(self) -> "Simulation":
        if self.size[0] == 0.0 or self.size[1] == 0.0:
            raise ValueError(
                f"size = {self.size}: size_x and size_y must be > 0 "
                "(set size_z = 0 for a 2D simulation)"
            )
        names = [m.name for m in self.monitors]
        repeated = sorted({n for n in names if names.count(n) > 1})
        if repeated:
            raise ValueError(f"monitor names must be unique; repeated: {', '.join(repeated)}")
        for i, source in enumerate(self.sources):
            label = source.name or f"sources[{i}]"
            self._check_inside(f"source {label!r}", source.center, (0.0, 0.0, 0.0))
        for monitor in self.monitors:
            self._check_inside(f"monitor {monitor.name!r}", monitor.center, monitor.size)
        return self

    def _check_inside(self, what, center, size) -> None:
        for axis, c, s, rc, rs in zip("xyz", center, size, self.center, self.size):
            lo, hi = c - s / 2.0, c + s / 2.0
            rlo, rhi = rc - rs / 2.0, rc + rs / 2.0
            tol = 1e-9 * max(1.0, abs(rlo), abs(rhi))
            if lo >= rlo - tol and hi <= rhi + tol:
                continue
            if axis == "z" and rs == 0.0:
                raise ValueError(
                    f"{what} must lie in the simulation plane z = {rc} µm "
                    "with size_z = 0 (this is a 2D simulation)"
                )
            raise ValueError(
                f"{what} spans {axis} = [{lo}, {hi}] µm, outside the simulation region "
                f"{axis} = [{rlo}, {rhi}] µm; sources and monitors must be inside "
                "the region (the PML is outside it)"
            )

    def summary(self) -> str:
        from etalon_amd.summary import summary

        return summary(self)

    def run(self):
        from etalon_amd.solver import run

        return run(self)


def eps_array(values: Any) -> np.ndarray:
    """A PixelGrid permittivity as a float64 array."""
    return np.asarray(values, dtype=np.float64)
