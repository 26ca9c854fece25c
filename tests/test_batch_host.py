"""Batch set-up on the CPU: the workload generator and the checks a batch makes
before it touches a device."""

import numpy as np
import pytest

from etalon_amd.batch import BatchSolver
from etalon_amd.grid import build_yee_grid, permittivity_on_nodes
from etalon_amd.workloads import uneven_batch
from oracle import te_numpy as ref
from tests.sims import bitwise_equal


def test_uneven_batch_is_seeded_and_in_range():
    a, b = uneven_batch(64, seed=7), uneven_batch(64, seed=7)
    assert [s.model_dump_json() for s in a] == [s.model_dump_json() for s in b]
    assert uneven_batch(3, seed=8)[0].model_dump_json() != a[0].model_dump_json()
    for sim in a:
        g = build_yee_grid(sim)
        assert 200 + 24 <= g.nx <= 1500 + 24 and 200 + 24 <= g.ny <= 1500 + 24
        assert g.n_steps(sim.run_time) == 4000
        assert 1 <= len(sim.monitors) <= 8
        assert all(m.wavelengths.size == 11 if hasattr(m.wavelengths, "size")
                   else len(m.wavelengths) == 11 for m in sim.monitors)


def test_batch_geometry_matches_the_oracle():
    """Bends (quarter-ring sectors) and rings rasterise exactly as in the oracle."""
    for sim in uneven_batch(4, seed=3, lo=80, hi=170, steps=10):
        g, rg = build_yee_grid(sim), ref.make_grid(sim)
        eps, reps = permittivity_on_nodes(sim, g), ref.permittivity(sim, rg)
        assert bitwise_equal(eps["Ex"], reps["Ex"]) and bitwise_equal(eps["Ey"], reps["Ey"])
        assert np.any(eps["Ex"] > 10.0)  # the silicon is on the grid


def test_batch_rejects_mixed_precision_and_step_counts():
    a = uneven_batch(1, seed=1, lo=80, hi=100, steps=30)
    b = uneven_batch(1, seed=1, lo=80, hi=100, steps=30, precision="double")
    c = uneven_batch(1, seed=1, lo=80, hi=100, steps=31)
    with pytest.raises(ValueError, match="one precision"):
        BatchSolver(a + b)
    with pytest.raises(ValueError, match="lockstep"):
        BatchSolver(a + c)
    with pytest.raises(ValueError, match="at least one"):
        BatchSolver([])


def test_batch_adjoint_needs_one_step_count():
    from etalon_amd.adjoint import AdjointProblem
    from etalon_amd.batch_adjoint import BatchAdjointProblem
    from tests.adjoint_sims import design_problem

    a, _ = design_problem("single", steps=200)
    b, _ = design_problem("single", steps=260)
    with pytest.raises(ValueError, match="step count"):
        BatchAdjointProblem([AdjointProblem(a, ["out"]), AdjointProblem(b, ["out"])])
    with pytest.raises(ValueError, match="tile kernels"):
        BatchAdjointProblem([AdjointProblem(a, ["out"], kernel="plain")])


def test_cylinder_lattice_painting_equals_the_oracle():
    """Runs of cylinders are painted together; the result equals painting them one by
    one (the oracle), also with holes across the grid edge, overlapping holes and
    other structures between runs."""
    import etalon_amd as et
    from etalon_amd.workloads import phc_holes

    holes = phc_holes(2.2, keep_out=0.5)
    box = et.Box(center=(0.3, 0.2, 0.0), size=(1.0, 0.4, 0.0), medium=et.Medium(n=2.0))
    big = et.Cylinder(center=(-1.0, 0.5, 0.0), radius=0.5, medium=et.Medium(n=1.0))
    sim = et.Simulation(
        size=(4.0, 3.6, 0.0), grid=et.UniformGrid(dl=0.05), medium=et.Medium(n=3.48),
        structures=[*holes[:40], box, big, *holes[40:]], boundary=et.PML(layers=8),
        sources=[et.PointSource(center=(-1.5, 0.0, 0.0), component="Hz",
                                pulse=et.GaussianPulse(wavelength=1.55, bandwidth=0.2))],
        run_time=0.05, precision="double", backend="hip")
    g = build_yee_grid(sim)
    mine = permittivity_on_nodes(sim, g)
    theirs = ref.permittivity(sim, ref.make_grid(sim))
    for comp in ("Ex", "Ey"):
        assert np.array_equal(mine[comp], theirs[comp]), comp
        assert (mine[comp] == 1.0).sum() > 100, "holes painted"


def test_batch_adjoint_needs_one_mode():
    from etalon_amd.adjoint import AdjointProblem
    from etalon_amd.batch_adjoint import BatchAdjointProblem
    from tests.adjoint_sims import design_problem

    a, _ = design_problem("single", steps=200)
    with pytest.raises(ValueError, match="mode"):
        BatchAdjointProblem([AdjointProblem(a, ["out"]), AdjointProblem(a, ["out"], mode="store")])
    assert BatchAdjointProblem([AdjointProblem(a, ["out"], mode="store")]).mode == "store"
