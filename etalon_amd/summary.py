"""Simulation.summary(): cost of a run before it starts (etalon/fdtd/summary.py:14-39)."""

from __future__ import annotations

from etalon_amd import __version__
from etalon_amd.capabilities import check_supported
from etalon_amd.grid import build_yee_grid
from etalon_amd.sampling import monitor_samplings


def summary(sim) -> str:
    check_supported(sim)
    grid = build_yee_grid(sim)
    real_bytes = 4 if sim.precision == "single" else 8
    nodes = sum(a * b for a, b in map(grid.shape, ("Ex", "Ey", "Hz")))
    # fields, psi and dt/eps (three arrays per component), and the ping-pong copies
    # of the fields and psi used by the fused path
    field_bytes = 3 * nodes * real_bytes
    pingpong_bytes = (2 * nodes + grid.n_cells) * real_bytes
    monitor_bytes = 0
    for monitor in sim.monitors:
        for sampling in monitor_samplings(monitor, grid):
            monitor_bytes += 2 * real_bytes * sampling.flat_index.size * sampling.wavelengths.size
    rx, ry = grid.region_cells
    steps = grid.n_steps(sim.run_time)
    return "\n".join([
        f"etalon_amd {__version__} | 2D TE (Ex, Ey, Hz) | backend {sim.backend}, "
        f"{sim.precision} precision",
        f"grid    {grid.nx} x {grid.ny} cells = {rx} x {ry} region "
        f"+ {grid.pml_layers}-cell CPML per face; dx = {grid.dx:.4g} um, dy = {grid.dy:.4g} um",
        f"time    dt = {grid.dt:.5g} s (courant {grid.courant}); {steps} steps "
        f"for {sim.run_time * 1e15:.4g} fs",
        f"memory  {field_bytes / 1e6:.3g} MB fields + {pingpong_bytes / 1e6:.3g} MB ping-pong "
        f"+ {monitor_bytes / 1e6:.3g} MB monitors",
    ])
