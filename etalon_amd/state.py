"""Arrays a 2D TE step reads and writes (etalon/fdtd/state.py:15-99).

This dataclass is the kernel ABI: field names, shapes and the coefficient
definitions (cb = Δt/(ε0 ε_r), db = Δt/μ0, inv_d = 1/(Δ·1e-6)) are those of the
reference.  Coefficients are computed in float64 and cast once, so any two
backends start from bit-identical inputs.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Any, Callable, Optional

import numpy as np

from etalon_amd.constants import EPSILON_0, MU_0, UM
from etalon_amd.cpml import CPMLAxis
from etalon_amd.grid import YeeGrid2D

FIELDS = ("ex", "ey", "hz", "psi_ex_y", "psi_ey_x", "psi_hz_x", "psi_hz_y")
PROFILES = ("bx_e", "cx_e", "ikx_e", "bx_h", "cx_h", "ikx_h",
            "by_e", "cy_e", "iky_e", "by_h", "cy_h", "iky_h")


@dataclass
class TEState:
    ex: Any
    ey: Any
    hz: Any
    psi_ex_y: Any
    psi_ey_x: Any
    psi_hz_x: Any
    psi_hz_y: Any
    cb_ex: Any
    cb_ey: Any
    bx_e: Any
    cx_e: Any
    ikx_e: Any
    bx_h: Any
    cx_h: Any
    ikx_h: Any
    by_e: Any
    cy_e: Any
    iky_e: Any
    by_h: Any
    cy_h: Any
    iky_h: Any
    db: Any
    inv_dx: Any
    inv_dy: Any
    nx: int
    ny: int
    # host copies of the CPML profiles as cast (PROFILES names), when the builder kept
    # them: readers of the profiles need no device copy
    host_profiles: Optional[dict] = None


def cb_arrays(grid: YeeGrid2D, eps_r: dict) -> dict:
    """cb_ex, cb_ey = Δt / (ε0 εr) at the Ex and Ey nodes, float64
    (etalon/fdtd/state.py:80-81)."""
    return {"cb_ex": grid.dt / (EPSILON_0 * eps_r["Ex"]),
            "cb_ey": grid.dt / (EPSILON_0 * eps_r["Ey"])}


def te_coefficients(grid: YeeGrid2D, eps_r: dict, cpml_x: CPMLAxis, cpml_y: CPMLAxis) -> dict:
    """The float64 coefficient arrays and scalars of a TE state, before the cast."""
    return {
        **cb_arrays(grid, eps_r),
        "bx_e": cpml_x.b_e, "cx_e": cpml_x.c_e, "ikx_e": cpml_x.inv_kappa_e,
        "bx_h": cpml_x.b_h, "cx_h": cpml_x.c_h, "ikx_h": cpml_x.inv_kappa_h,
        "by_e": cpml_y.b_e, "cy_e": cpml_y.c_e, "iky_e": cpml_y.inv_kappa_e,
        "by_h": cpml_y.b_h, "cy_h": cpml_y.c_h, "iky_h": cpml_y.inv_kappa_h,
        "db": grid.dt / MU_0,
        "inv_dx": 1.0 / (grid.dx * UM),
        "inv_dy": 1.0 / (grid.dy * UM),
    }


def build_te_state(grid: YeeGrid2D, eps_r: dict, cpml_x: CPMLAxis, cpml_y: CPMLAxis,
                   dtype, to_device: Callable[[np.ndarray], Any]) -> TEState:
    """Zero fields; coefficients cast once from float64 to ``dtype`` and moved by ``to_device``."""
    coef = te_coefficients(grid, eps_r, cpml_x, cpml_y)

    def put(a):
        return to_device(np.ascontiguousarray(a, dtype=dtype))

    shapes = {"ex": "Ex", "ey": "Ey", "hz": "Hz", "psi_ex_y": "Ex", "psi_ey_x": "Ey",
              "psi_hz_x": "Hz", "psi_hz_y": "Hz"}
    arrays = {name: put(np.zeros(grid.shape(comp))) for name, comp in shapes.items()}
    arrays.update({name: put(coef[name]) for name in ("cb_ex", "cb_ey", *PROFILES)})
    host = {name: np.ascontiguousarray(coef[name], dtype=dtype) for name in PROFILES}
    return TEState(
        **arrays,
        db=dtype(coef["db"]), inv_dx=dtype(coef["inv_dx"]), inv_dy=dtype(coef["inv_dy"]),
        nx=grid.nx, ny=grid.ny, host_profiles=host,
    )
