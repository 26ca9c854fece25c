"""Host side of the HIP backend on the CPU: models, grid, CPML, sampling, sources
and state must reproduce the oracle's float64 setup bit for bit (they feed the
kernels), plus the reference's model/grid unit tests (plan-m0 U1–U4, U11–U12)."""

import ctypes
import math
import re
import sys
from pathlib import Path

import numpy as np
import pytest
from pydantic import ValidationError

import etalon_amd as et
from etalon_amd.cpml import cpml_axis
from etalon_amd.grid import build_yee_grid, cells_along, courant_time_step, permittivity_on_nodes
from etalon_amd.sampling import axis_sampling, dft_phasors, monitor_samplings
from etalon_amd.sources import point_injection
from etalon_amd.state import te_coefficients
from oracle import te_numpy as ref
from tests.sims import bitwise_equal, full_sim, pulse, small_sim, vacuum_pulse

ROOT = Path(__file__).resolve().parents[1]


@pytest.mark.parametrize("make", [small_sim, full_sim, vacuum_pulse])
@pytest.mark.parametrize("precision", ["single", "double"])
def test_host_setup_equals_oracle_bitwise(make, precision):
    sim = make(precision=precision)
    grid = build_yee_grid(sim)
    g = ref.make_grid(sim)
    assert (grid.nx, grid.ny, grid.dt, grid.x_min, grid.y_min) == (g.nx, g.ny, g.dt, g.x_min, g.y_min)
    eps = permittivity_on_nodes(sim, grid)
    reps = ref.permittivity(sim, g)
    assert bitwise_equal(eps["Ex"], reps["Ex"]) and bitwise_equal(eps["Ey"], reps["Ey"])
    bg = sim.medium.permittivity
    cx = cpml_axis(grid.nx, grid.dx, grid.dt, sim.boundary, bg)
    cy = cpml_axis(grid.ny, grid.dy, grid.dt, sim.boundary, bg)
    coef = te_coefficients(grid, eps, cx, cy)
    real = np.float32 if precision == "single" else np.float64
    rs = ref.State(g, reps, ref.cpml(g.nx, g.dx, g.dt, sim.boundary, bg),
                   ref.cpml(g.ny, g.dy, g.dt, sim.boundary, bg), real)
    for name in ("cb_ex", "cb_ey", "bx_e", "cx_e", "ikx_e", "bx_h", "cx_h", "ikx_h",
                 "by_e", "cy_e", "iky_e", "by_h", "cy_h", "iky_h"):
        assert bitwise_equal(np.ascontiguousarray(coef[name], dtype=real), getattr(rs, name)), name
    for name in ("db", "inv_dx", "inv_dy"):
        assert real(coef[name]) == getattr(rs, name)
    for src in sim.sources:
        a, b = point_injection(src, grid, eps), ref.injection(src, g, reps)
        assert np.array_equal(a.i, b.i) and np.array_equal(a.j, b.j)
        assert bitwise_equal(a.coefficient, b.coef)
        # the whole-column evaluation against the oracle's one evaluation per step
        table = a.delta_table(3, 1200, grid.dt)
        assert bitwise_equal(table, np.array([b.delta(n, g.dt) for n in range(3, 1203)]))
    for mon in sim.monitors:
        for s, r in zip(monitor_samplings(mon, grid), ref.monitors_of(mon, g)):
            assert np.array_equal(s.flat_index, r.flat)
            t = (np.arange(5, 9, dtype=np.float64) + 0.5) * grid.dt
            assert bitwise_equal(dft_phasors(s.frequencies, t, grid.dt), ref.phasors(r.freqs, t, g.dt))


# ---------------------------------------------------------------- models (U1, U2)
def test_models_are_frozen_and_reject_typos():
    sim = small_sim()
    with pytest.raises(ValidationError, match="frozen"):
        sim.run_time = 1e-12
    with pytest.raises(ValidationError, match="wavelenght"):
        et.GaussianPulse(wavelength=1.55, bandwidth=0.1, wavelenght=1.5)
    with pytest.raises(ValidationError):
        et.GaussianPulse(wavelength=math.nan, bandwidth=0.1)


def test_json_round_trip_with_extra_geometry():
    sim = small_sim(structures=[
        et.Box(center=(0.1, 0.0, 0.0), size=(0.5, 0.2, 0.0), medium=et.Medium(n=3.476)),
        et.Ring(inner_radius=0.2, outer_radius=0.4, medium=et.Medium(n=2.0)),
        et.Cylinder(radius=0.1, medium=et.Medium()),
        et.PixelGrid(size=(0.2, 0.2, 0.0), permittivity=np.full((2, 2), 4.0)),
    ], precision="double")
    again = et.Simulation.model_validate_json(sim.model_dump_json())
    assert again == sim and hash(again) == hash(sim)


def test_validation_messages_follow_the_reference():
    with pytest.raises(ValidationError, match=r"n = 0\.9 is below 1"):
        et.Medium(n=0.9)
    with pytest.raises(ValidationError, match="smaller than 2 × wavelength"):
        pulse(bandwidth=3.1)
    with pytest.raises(ValidationError, match="size_x and size_y must be > 0"):
        small_sim(size=(0.0, 1.0, 0.0))
    m = et.DFTMonitor(name="probe", wavelengths=[1.55])
    with pytest.raises(ValidationError, match="repeated: probe"):
        small_sim(monitors=[m, m])
    with pytest.raises(ValidationError, match="repeats"):
        et.DFTMonitor(name="p", wavelengths=[1.55], components=("Hz", "Hz"))
    src = et.PointSource(center=(4.5, 0.0, 0.0), component="Hz", pulse=pulse())
    with pytest.raises(ValidationError, match=r"source 'sources\[0\]' spans x"):
        small_sim(sources=[src])
    src = et.PointSource(center=(0.0, 0.0, 0.1), component="Hz", pulse=pulse(), name="d")
    with pytest.raises(ValidationError, match="'d' must lie in the simulation plane z = 0"):
        small_sim(sources=[src])
    with pytest.raises(ValidationError):
        small_sim(backend="cuda")


def test_pulse_derived_quantities():
    p = pulse()
    assert p.frequency == pytest.approx(et.C_0 / 1.55)
    assert p.frequency_width == pytest.approx((et.C_0 / 1.30 - et.C_0 / 1.80) / 2)
    assert p.delay == pytest.approx(5 * p.width)
    assert 4e-15 < p.width < 6e-15


# ---------------------------------------------------------------- grid (U3, U4)
def test_cells_along_and_grid_counts():
    assert cells_along(0.56, 0.01) == 56 and cells_along(1.0, 0.3) == 4
    grid = build_yee_grid(small_sim(size=(8.0, 4.0, 0.0), boundary=et.PML(layers=12)))
    assert (grid.nx, grid.ny) == (184, 104) and grid.region_cells == (160, 80)
    x_ey, _ = grid.coords("Ey")
    assert x_ey[12] == pytest.approx(-4.0) and x_ey[-13] == pytest.approx(4.0)
    assert grid.shape("Hz") == (grid.nx, grid.ny)
    assert grid.shape("Ex") == (grid.nx, grid.ny + 1)
    assert grid.shape("Ey") == (grid.nx + 1, grid.ny)


def test_time_step_is_courant_fraction():
    grid = build_yee_grid(small_sim(size=(2.0, 1.2, 0.0), courant=0.7))
    limit = 1.0 / (et.C_0 * math.sqrt(1 / grid.dx**2 + 1 / grid.dy**2))
    assert grid.dt == pytest.approx(0.7 * limit, rel=1e-15)
    assert courant_time_step(1.0, 0.05, 0.05) == pytest.approx(0.05 / (et.C_0 * math.sqrt(2)))
    assert grid.n_steps(10 * grid.dt) == 10 and grid.n_steps(10.5 * grid.dt) == 11


def test_summary_and_capabilities():
    text = small_sim().summary()
    assert "56 x 46 cells = 40 x 30 region + 8-cell CPML" in text and "MB fields" in text
    text.encode("ascii")
    with pytest.raises(NotImplementedError, match="M1"):
        small_sim(size=(2.0, 1.5, 1.0)).summary()
    from etalon_amd.capabilities import check_supported

    src = et.PointSource(center=(0, 0, 0), component="Ez", pulse=pulse(), name="tm")
    with pytest.raises(NotImplementedError, match=r"'tm' drives Ez.*2D TE only"):
        check_supported(small_sim(sources=[src]))


def test_box_staircase_faces_inclusive_later_wins():
    si, sin = et.Medium(n=3.5), et.Medium(n=2.0)
    boxes = [et.Box(center=(0.0, 0.0, 0.0), size=(0.5, 0.4, 0.22), medium=si),
             et.Box(center=(0.2, 0.0, 0.0), size=(0.2, 0.2, 0.22), medium=sin)]
    sim = small_sim(structures=boxes)
    grid = build_yee_grid(sim)
    eps = permittivity_on_nodes(sim, grid)
    x, y = grid.coords("Ex")
    X, Y = np.meshgrid(x, y, indexing="ij")
    in_si = (np.abs(X) <= 0.25 + 1e-9) & (np.abs(Y) <= 0.2 + 1e-9)
    in_sin = (np.abs(X - 0.2) <= 0.1 + 1e-9) & (np.abs(Y) <= 0.1 + 1e-9)
    np.testing.assert_array_equal(eps["Ex"], np.where(in_sin, 4.0, np.where(in_si, 3.5**2, 1.0)))


def test_axis_sampling_rules_and_errors():
    nodes = np.arange(11) * 0.1
    np.testing.assert_array_equal(axis_sampling(nodes, 0.5, 0.2, "m").index, [4, 5, 6])
    s = axis_sampling(nodes, 0.53, 0.0, "m")
    np.testing.assert_array_equal(s.index, [5, 6])
    with pytest.raises(ValueError, match="thinner than one cell"):
        axis_sampling(nodes, 0.55, 0.02, "m")
    with pytest.raises(ValueError, match="outside the grid"):
        axis_sampling(nodes, 1.5, 0.0, "m")


# ---------------------------------------------------------------- backend selection
def test_unknown_backend_name():
    from etalon_amd.backends import get_backend

    with pytest.raises(ValueError, match="unknown backend"):
        get_backend("numpy", "single")


def test_hip_backend_fails_loudly_without_the_library(monkeypatch, tmp_path):
    """No silent CPU fallback: a missing library is a RuntimeError."""
    from etalon_amd.hip import _lib

    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(RuntimeError, match="not built"):
        _lib.load(tmp_path / "libetalon_hip.so")


# ---------------------------------------------------------------- the C ABI
def _declared_functions():
    text = (ROOT / "include" / "etalon_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^int\s+(etl_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    from etalon_amd.hip import _lib

    lib_path = _lib.LIB_PATH
    if not lib_path.exists():
        pytest.skip("libetalon_hip.so not built (python -m etalon_amd.build)")
    lib = ctypes.CDLL(str(lib_path))
    declared = _declared_functions()
    assert len(declared) >= 20
    for name in declared:
        assert hasattr(lib, name), name
    assert sorted(_lib.exported_symbols()) == declared


def test_error_path_without_init_reports_a_message():
    """Calls before etl_init return ETL_ESTATE with a message (no device touched)."""
    from etalon_amd.hip import _lib

    if not _lib.LIB_PATH.exists():
        pytest.skip("libetalon_hip.so not built")
    if _lib._device is not None:
        pytest.skip("library already initialised in this process")
    lib = _lib.load()
    rc = lib.etl_synchronize()
    assert rc == -3
    assert "etl_init" in _lib.last_error(lib)
