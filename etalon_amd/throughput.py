"""Cell-update throughput of the HIP backend: the calibration a run-time estimate
reads (etalon/validation/throughput.py:27-69, etalon/studio/runner.py:147-162).

``measure_throughput`` times the reference's calibration workload — a square vacuum
grid of ``region_cells``² cells at dl = 0.02 µm plus a 12-layer CPML, one Hz point
source (Gaussian pulse 1.55 µm, bandwidth 0.1 µm) at the centre — after ``warmup``
steps, bracketed by device synchronisation: cells (PML included) × steps / wall.
``fused=True`` (the default) advances with the fused plan, as ``run()`` does;
``fused=False`` with one launch per half-step (``step()``, the reference's loop).
``calibrate(seconds)`` sizes a second measurement to about ``seconds`` of work from a
short trial, as Studio does.
"""

from __future__ import annotations

import time
from dataclasses import dataclass

from etalon_amd import models as et


@dataclass
class Throughput:
    backend: str
    device: str
    precision: str
    cells: tuple[int, int]
    """(nx, ny) including PML."""
    steps: int
    wall_time_s: float
    cell_updates_per_s: float


def calibration_sim(region_cells: int = 2000, precision: str = "single"):
    """The reference's calibration workload (etalon/validation/throughput.py:36-52)."""
    dl = 0.02
    return et.Simulation(
        size=(region_cells * dl, region_cells * dl, 0.0),
        grid=et.UniformGrid(dl=dl),
        boundary=et.PML(layers=12),
        sources=[et.PointSource(center=(0.0, 0.0, 0.0), component="Hz",
                                pulse=et.GaussianPulse(wavelength=1.55, bandwidth=0.1))],
        run_time=1.0,  # not used: a fixed number of steps is timed
        precision=precision,
        backend="hip",
    )


def measure_throughput(region_cells: int = 2000, steps: int = 500, backend: str = "hip",
                       precision: str = "single", warmup: int = 20,
                       fused: bool = True) -> Throughput:
    """Time ``steps`` steps of the calibration grid after ``warmup`` untimed ones."""
    from etalon_amd.solver import FDTDSolver

    if backend not in ("hip", "auto"):
        raise ValueError(f"measure_throughput runs the HIP backend, not {backend!r}")
    solver = FDTDSolver(calibration_sim(region_cells, precision))

    def advance(count: int) -> None:
        if fused:
            solver.advance(count)
        else:
            for _ in range(count):
                solver.step()

    advance(warmup)
    solver.backend.synchronize()
    start = time.perf_counter()
    advance(steps)
    solver.backend.synchronize()
    wall = time.perf_counter() - start
    grid = solver.grid
    return Throughput(backend=solver.backend.name, device=solver.backend.device,
                      precision=precision, cells=(grid.nx, grid.ny), steps=steps,
                      wall_time_s=wall, cell_updates_per_s=grid.n_cells * steps / wall)


def calibrate(seconds: float, region_cells: int = 2000, precision: str = "single") -> dict:
    """Cells per second over about ``seconds`` of stepping (etalon/studio/runner.py:
    147-162): a 10-step trial sizes the measured run, clamped to 20-200 000 steps."""
    trial = measure_throughput(region_cells=region_cells, steps=10, precision=precision,
                               warmup=5)
    n_cells = trial.cells[0] * trial.cells[1]
    steps = int(seconds * trial.cell_updates_per_s / n_cells)
    steps = min(max(steps, 20), 200_000)
    result = measure_throughput(region_cells=region_cells, steps=steps, precision=precision,
                                warmup=5)
    label, _, name = result.device.partition(" ")
    return {"cells_per_s": result.cell_updates_per_s, "backend": label, "device": name or label,
            "steps": steps}
