"""Batched adjoint: every candidate's J and gradient from one BatchAdjointProblem
(one forward plan, one co-plan sweep over all candidates) are bit-identical to the
candidate's own AdjointProblem."""

import numpy as np
import pytest

from etalon_amd.adjoint import AdjointProblem
from etalon_amd.batch_adjoint import BatchAdjointProblem
from etalon_amd.hip.plan import Tune
from tests.adjoint_sims import design_problem

pytestmark = pytest.mark.gpu

# uneven candidates: grid size, design region, pixel grid and CPML depth differ
SPECS = [dict(seed=1, size=(3.0, 2.0), pixels=(6, 6)),
         dict(seed=2, size=(4.0, 2.5), pixels=(7, 5), region=0.8),
         dict(seed=3, size=(3.2, 3.0), pixels=(5, 5), pml=10)]


def candidates(precision, steps=301):
    out = [design_problem(precision, steps=steps, **spec) for spec in SPECS]
    problems = [AdjointProblem(sim, ["out"], checkpoint_every=64) for sim, _ in out]
    return problems, [np.asarray(design.permittivity) for _, design in out]


@pytest.mark.parametrize("precision", ["single", "double"])
def test_batch_gradients_equal_each_candidate_alone(hip, precision):
    problems, eps = candidates(precision)
    print(f"{len(problems)} candidates, {precision}", flush=True)
    Js, grads = BatchAdjointProblem(problems).run(eps)
    for k, (p, e) in enumerate(zip(problems, eps)):
        J1, g1 = p.run(e)
        assert Js[k] == J1, f"candidate {k}: J"
        assert np.array_equal(grads[k], g1), f"candidate {k}: gradient"


@pytest.mark.parametrize("depth", [1, 2])
def test_batch_at_other_depths(hip, depth):
    problems, eps = candidates("single", steps=203)
    tune = Tune(temporal=depth)
    for p in problems:
        p.tune = tune
    print(f"{len(problems)} candidates, depth {depth}", flush=True)
    Js, grads = BatchAdjointProblem(problems, tune=tune).run(eps)
    for k, (p, e) in enumerate(zip(problems, eps)):
        J1, g1 = p.run(e)
        assert Js[k] == J1 and np.array_equal(grads[k], g1), f"candidate {k}"


def test_batch_with_split_launches(hip):
    """Launch mode 3 (large plans): per step block the interior tiles of both plans in
    one launch, then their strip and source/monitor tiles; the same bits."""
    problems, eps = candidates("single", steps=203)
    print(f"{len(problems)} candidates, launch 3", flush=True)
    Js, grads = BatchAdjointProblem(problems, tune=Tune(launch=3)).run(eps)
    for k, (p, e) in enumerate(zip(problems, eps)):
        J1, g1 = p.run(e)
        assert Js[k] == J1 and np.array_equal(grads[k], g1), f"candidate {k}"
