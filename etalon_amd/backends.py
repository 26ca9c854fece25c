"""Backend selection (etalon/fdtd/backends/__init__.py:16-53).

``"hip"`` raises RuntimeError when the library or a device is missing — never a
silent fallback.  ``"auto"`` is ``"hip"``: this package has no CPU compute path
(the NumPy reference is the reference package's own backend), so on a machine
without a GPU it raises the same RuntimeError instead of degrading.
"""

from __future__ import annotations

import logging

log = logging.getLogger("etalon_amd")


def get_backend(name: str, precision: str):
    if name in ("hip", "auto"):
        from etalon_amd.hip.backend import HipBackend

        return HipBackend(precision)
    raise ValueError(f"unknown backend {name!r}; use 'auto' or 'hip'")
