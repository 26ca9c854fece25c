"""etalon_amd — the MI355X (gfx950) backend of etalon's 2D TE FDTD solver.

A drop-in for the reference's solver path: the same user models
(``Simulation``, ``Medium``, ``Box``, ``PointSource``, ``GaussianPulse``,
``DFTMonitor``, ``PML``, ``UniformGrid``; etalon/__init__.py:29-72), the same
backend protocol (etalon/fdtd/backends/__init__.py:1-5) implemented over the
hand-written HIP kernels of ``libetalon_hip.so``, and an ``FDTDSolver`` whose
``run()`` replays the time loop on the GPU.  Public names load lazily.
"""

from __future__ import annotations

import importlib
from typing import Any

__version__ = "0.1.0"

_LAZY = {
    "C_0": "etalon_amd.constants",
    "freq_to_wvl": "etalon_amd.constants",
    "wvl_to_freq": "etalon_amd.constants",
    "PML": "etalon_amd.models",
    "UniformGrid": "etalon_amd.models",
    "Medium": "etalon_amd.models",
    "DFTMonitor": "etalon_amd.models",
    "Simulation": "etalon_amd.models",
    "GaussianPulse": "etalon_amd.models",
    "PointSource": "etalon_amd.models",
    "Box": "etalon_amd.models",
    "Cylinder": "etalon_amd.models",
    "Ring": "etalon_amd.models",
    "PixelGrid": "etalon_amd.models",
    "SimulationResult": "etalon_amd.result",
    "FDTDSolver": "etalon_amd.solver",
    "get_backend": "etalon_amd.backends",
}

__all__ = sorted([*_LAZY, "__version__"])


def __getattr__(name: str) -> Any:
    if name in _LAZY:
        value = getattr(importlib.import_module(_LAZY[name]), name)
        globals()[name] = value
        return value
    raise AttributeError(f"module 'etalon_amd' has no attribute {name!r}")


def __dir__() -> list[str]:
    return sorted(set(globals()) | set(__all__))
