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
