"""M0 physics validation on the HIP path: the reference's V1–V6 checks with the
reference's thresholds (tests/validation/test_m0.py:17-72), and the point2d
convention check against the committed Meep result (bench/metrics.py:20 gate).

These run the product path (fused plan, or the plain kernels where a case steps)
end to end; the physics, not the bits, is what is checked here.
"""

from pathlib import Path

import numpy as np
import pytest

from etalon_amd import validation as m0

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module")
def vacuum_cases():
    """0° and 45° at Δ = 0.1 µm (9–18 cells/λ) and Δ = 0.05 µm (18–36 cells/λ)."""
    return m0.dispersion_cases(0.1) + m0.dispersion_cases(0.05)


def test_v1_dispersion_matches_yee_relation(vacuum_cases):
    for case in vacuum_cases:
        residual = case.v_measured / case.v_theory - 1
        print(f"dl={case.dl} angle={case.angle} max|residual|={np.abs(residual).max():.3e}")
        assert np.abs(residual).max() < 1e-3, (case.dl, case.angle, residual)


def test_v1_dispersion_anisotropy_is_resolved(vacuum_cases):
    axis, diagonal = (c for c in vacuum_cases if c.dl == 0.1)
    coarsest = np.argmin(axis.cells_per_wavelength)
    assert axis.v_measured[coarsest] < 0.99
    assert diagonal.v_measured[coarsest] - axis.v_measured[coarsest] > 0.01


def test_v1_run_time_is_converged():
    short = m0.dispersion_cases(0.05, run_time=200e-15)
    long = m0.dispersion_cases(0.05, run_time=400e-15)
    for a, b in zip(short, long):
        assert np.abs(a.v_measured - b.v_measured).max() < 1e-4


def test_v2_raw_phase_velocity_within_half_percent(vacuum_cases):
    for case in vacuum_cases:
        resolved = case.cells_per_wavelength >= 20
        assert np.all(np.abs(case.v_measured[resolved] - 1) < 5e-3), (case.dl, case.angle)


def test_v3_uniform_dielectric():
    for case in m0.dispersion_cases(0.05, n=1.5):
        residual = case.v_measured / case.v_theory - 1
        assert np.abs(residual).max() < 1e-3, (case.angle, residual)


def test_v4_cpml_reflection():
    """Plan threshold −60 dB (1e-3), guarded at −80 dB (1e-4); the reference measured
    2.68e-5 (side) and 1.95e-5 (corner) in f32."""
    case = m0.cpml_reflection_case()
    assert np.abs(case.hz_reference).max() > 0
    for probe, error in zip(case.probes, case.error):
        print(f"{probe}: {error:.3e}")
        assert error < 1e-4, (probe, error)


def test_v5_late_time_stability():
    case = m0.late_time_case()
    energy = case.energy / case.energy.max()
    peak = int(np.argmax(energy))
    print(f"final energy / peak = {energy[-1]:.3e}")
    assert energy[-1] < 1e-8
    assert np.all(np.diff(energy[peak:]) <= 1e-12)


def test_v6_field_matches_2d_greens_function():
    case = m0.green_function_case(dl=0.05)
    ratio = case.hz_measured / case.hz_analytic
    print(f"max |ratio - 1| = {np.abs(ratio - 1).max():.3e}")
    assert np.all(np.abs(ratio - 1) < 1e-2), ratio


def test_point2d_matches_meep():
    """max |ΔT21| against Meep 1.35 at resolution 50 below 1e-3 (the reference: 2.14e-6)."""
    ref = np.load(GOLDEN / "meep_point2d.npz")
    wavelengths, t21, steps = m0.point2d_t21()
    assert steps == int(ref["steps"])
    np.testing.assert_allclose(wavelengths, ref["wavelengths"], rtol=0, atol=1e-12)
    err = float(np.max(np.abs(t21 - ref["t21"])))
    print(f"max |dT21| = {err:.3e}")
    assert err < 1e-3


def test_m0_values_reproduce_the_reference_record():
    """The values bench/validate_m0.py records (:214, :245-252), against the ones the
    reference measured on its CUDA backend (bench/results/b2ff355.json:8-75, fixture
    tests/golden/m0_record.json) to 1e-6 relative: the kernels' arithmetic is the
    reference's, and host transcendentals differ by ulps across NumPy builds (measured
    on MI355X: 14 of the 16 values equal to every printed digit, the largest difference
    9e-8, in the minimum v/c at dl = 0.025)."""
    import json

    record = json.loads((GOLDEN / "m0_record.json").read_text())
    got = {}
    for dl in (0.1, 0.05, 0.025):
        for c in m0.dispersion_cases(dl):
            got[("dispersion", c.dl, c.angle)] = (
                float(np.max(np.abs(c.v_measured / c.v_theory - 1))), float(np.min(c.v_measured)))
    reflection = m0.cpml_reflection_case()
    late = m0.late_time_case()
    green = m0.green_function_case(dl=0.05)
    checks = []
    for d in record["dispersion"]:
        residual, v_min = got[("dispersion", d["dl_um"], d["angle_deg"])]
        checks.append((f"dispersion dl={d['dl_um']} angle={d['angle_deg']} residual", residual,
                       d["max_abs_residual_vs_yee"]))
        checks.append((f"dispersion dl={d['dl_um']} angle={d['angle_deg']} min v/c", v_min,
                       d["min_v_over_c"]))
    for probe, error in zip(reflection.probes, reflection.error):
        checks.append((f"cpml {probe}", float(error), record["cpml_reflection"][probe]))
    checks.append(("late-time energy", float(late.energy[-1] / np.max(late.energy)),
                   record["late_time_final_energy_rel"]))
    checks.append(("green", float(np.abs(green.hz_measured / green.hz_analytic - 1).max()),
                   record["green_function_max_rel_error"]))
    assert len(checks) == 2 * 6 + 2 + 2
    for label, value, want in checks:
        print(f"{label}: {value:.6e} (recorded {want:.6e}, rel {abs(value / want - 1):.1e})")
    for label, value, want in checks:
        assert value == pytest.approx(want, rel=1e-6), label
