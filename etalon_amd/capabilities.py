"""What the solver runs: 2D TE (etalon/fdtd/capabilities.py:7-24)."""

from __future__ import annotations

from etalon_amd.grid import TE_COMPONENTS


def check_supported(sim) -> None:
    """Raise NotImplementedError for a valid description the solver cannot run."""
    if not sim.is_2d:
        raise NotImplementedError("3D simulations arrive in M1; set size_z = 0 for a 2D simulation")
    for i, source in enumerate(sim.sources):
        if source.component not in TE_COMPONENTS:
            label = source.name or f"sources[{i}]"
            raise NotImplementedError(
                f"source {label!r} drives {source.component}, a 2D TM component; this build "
                "simulates 2D TE only (Ex, Ey, Hz)"
            )
    for monitor in sim.monitors:
        for component in monitor.components or ():
            if component not in TE_COMPONENTS:
                raise NotImplementedError(
                    f"monitor {monitor.name!r} asks for {component}, a 2D TM component; this "
                    "build simulates 2D TE only (Ex, Ey, Hz)"
                )
