"""Monitor interpolation pinned by a known answer (the reference's
tests/unit/test_solver.py:69-79): a point monitor equals bilinear interpolation of
the same run's plane monitor at the point, to rel 1e-12 in float64.

The interpolation here is the textbook bilinear formula on the plane's node
coordinates (bracketing nodes found by search, weights from the distances), not the
sampling weights of etalon/fdtd/dft.py that both the oracle's ``monitor_values`` and
``etalon_amd.result.monitor_data`` apply (etalon/io/result.py:54-91).  So a wrong
weight, a swapped pair of nodes or a wrong axis order in either fails it.  The CPU
tests pin the oracle (and catch a planted weight swap); the GPU tests pin the HIP
path through ``FDTDSolver.run()`` and ``result.py``, at the reference's small case and
at the 184 x 184 vacuum-pulse case whose Hz probes sit mid-cell.
"""

import numpy as np
import pytest

import etalon_amd as et
from oracle import te_numpy as ref
from tests.sims import small_sim, vacuum_pulse

POINT = (0.31, 0.12)


def bilinear(xs, ys, values, x0, y0):
    """values[a, b] given at (xs[a], ys[b]) (both increasing), interpolated at (x0, y0)."""
    xs, ys = np.asarray(xs, np.float64), np.asarray(ys, np.float64)
    a = int(np.clip(np.searchsorted(xs, x0, side="right") - 1, 0, xs.size - 2))
    b = int(np.clip(np.searchsorted(ys, y0, side="right") - 1, 0, ys.size - 2))
    tx = (x0 - xs[a]) / (xs[a + 1] - xs[a])
    ty = (y0 - ys[b]) / (ys[b + 1] - ys[b])
    v = np.asarray(values, np.complex128)
    return ((1 - tx) * (1 - ty) * v[a, b] + tx * (1 - ty) * v[a + 1, b]
            + (1 - tx) * ty * v[a, b + 1] + tx * ty * v[a + 1, b + 1])


def _sim(precision, components=("Hz",), points=(POINT,)):
    wls = [1.55]
    monitors = [et.DFTMonitor(name=f"point{k}", center=(x, y, 0.0), wavelengths=wls,
                              components=components) for k, (x, y) in enumerate(points)]
    monitors.append(et.DFTMonitor(name="plane", size=(2.0, 1.5, 0.0), wavelengths=wls,
                                  components=components))
    return small_sim(monitors=monitors, precision=precision)


def _oracle_pairs(solver, names):
    """{(point name, comp): (point value, plane interpolated there)} from the oracle."""
    values = solver.monitor_values()
    planes = {m.comp: m for m in solver.mons if m.name == "plane"}
    out = {}
    for m in solver.mons:
        if m.name not in names:
            continue
        plane = planes[m.comp]
        v = values[("plane", m.comp)][..., 0]
        out[(m.name, m.comp)] = (complex(values[(m.name, m.comp)][0]),
                                 bilinear(plane.x.coords, plane.y.coords, v,
                                          m.x.point, m.y.point))
    return out


@pytest.mark.parametrize("components", [("Hz",), ("Ex", "Ey", "Hz")])
def test_oracle_point_equals_plane_interpolation(components):
    solver = ref.Solver(_sim("double", components)).run()
    pairs = _oracle_pairs(solver, {"point0"})
    assert len(pairs) == len(components)
    for key, (point, expected) in pairs.items():
        assert abs(expected) > 0, key
        assert point == pytest.approx(expected, rel=1e-12, abs=0), key


def test_oracle_single_precision_casts_back():
    """float32: the point value is the float64 interpolation of float32 accumulators,
    cast back to complex64 (agreement to float32 rounding)."""
    solver = ref.Solver(_sim("single")).run()
    values = solver.monitor_values()
    assert values[("point0", "Hz")].dtype == np.complex64
    (point, expected), = _oracle_pairs(solver, {"point0"}).values()
    assert point == pytest.approx(expected, rel=1e-6, abs=0)


def test_oracle_vacuum_pulse_probes_equal_plane_interpolation():
    """The 184 x 184 vacuum pulse in float64 on the oracle: every component of its four
    probes against the bilinear interpolation of a plane over the interior, to 1e-12
    of the plane's largest value (see the GPU twin below for why not of the value)."""
    base = vacuum_pulse("double")
    plane = et.DFTMonitor(name="plane", size=(8.0, 8.0, 0.0), wavelengths=[1.45])
    probes = [m.model_copy(update={"wavelengths": [1.45]}) for m in base.monitors]
    solver = ref.Solver(base.model_copy(update={"monitors": [*probes, plane]})).run()
    values = solver.monitor_values()
    pairs = _oracle_pairs(solver, {m.name for m in probes})
    assert len(pairs) == 12
    for (name, comp), (point, expected) in pairs.items():
        scale = np.abs(values[("plane", comp)]).max()
        assert abs(point - expected) <= 1e-12 * scale, (name, comp)


def test_planted_weight_swap_is_caught(monkeypatch):
    """A plausible mistake in the sampling weights — (fr, 1 - fr) for (1 - fr, fr) —
    planted in the oracle makes the check fail."""
    real_axis = ref.axis

    def swapped(nodes, center, size, what="m"):
        s = real_axis(nodes, center, size, what)
        if s.weights is not None and s.weights.size == 2:
            s = ref.Axis(s.index, s.coords, s.weights[::-1].copy(), s.point)
        return s

    monkeypatch.setattr(ref, "axis", swapped)
    solver = ref.Solver(_sim("double")).run()
    (point, expected), = _oracle_pairs(solver, {"point0"}).values()
    assert point != pytest.approx(expected, rel=1e-6, abs=0)


def _hip_pairs(result, names, components):
    """{(point, comp): (point value, plane interpolated there, max |plane value|)}."""
    out = {}
    for name in names:
        for comp in components:
            pt = result[name][comp]
            plane = result["plane"][comp].isel(wavelength=0)
            out[(name, comp)] = (complex(pt.values[0]),
                                 bilinear(plane.x, plane.y, plane.values, float(pt.x),
                                          float(pt.y)),
                                 float(np.abs(plane.values).max()))
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("components", [("Hz",), ("Ex", "Ey", "Hz")])
def test_hip_point_equals_plane_interpolation(hip, components):
    from etalon_amd.solver import FDTDSolver

    result = FDTDSolver(_sim("double", components)).run()
    assert result["point0"][components[0]].dims == ("wavelength",)
    for key, (point, expected, _) in _hip_pairs(result, ["point0"], components).items():
        assert abs(expected) > 0, key
        assert point == pytest.approx(expected, rel=1e-12, abs=0), key


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["double", "single"])
def test_hip_vacuum_pulse_probes_equal_plane_interpolation(hip, precision):
    """The 184 x 184 vacuum pulse (2000 steps), its four probes and a plane monitor over
    the interior at the first wavelength: every component of every probe (the Hz ones
    mid-cell) equals the bilinear interpolation of the plane, to ``rel`` of the plane's
    largest value.  (Ex vanishes on the source's symmetry axis y = 0, where p0 and p1
    sit: there only rounding is left, and the coordinate weights' offset of ~1e-15 from
    the node weighs in the much larger neighbouring row, so a relative error of the
    value itself means nothing.)"""
    from etalon_amd.solver import FDTDSolver

    base = vacuum_pulse(precision)
    plane = et.DFTMonitor(name="plane", size=(8.0, 8.0, 0.0), wavelengths=[1.45])
    probes = [m.model_copy(update={"wavelengths": [1.45]}) for m in base.monitors]
    sim = base.model_copy(update={"monitors": [*probes, plane]})
    result = FDTDSolver(sim).run()
    rel = 1e-12 if precision == "double" else 1e-6
    pairs = _hip_pairs(result, [m.name for m in probes], ("Ex", "Ey", "Hz"))
    assert len(pairs) == 12
    for key, (point, expected, scale) in pairs.items():
        assert scale > 0, key
        assert abs(point - expected) <= rel * scale, key
