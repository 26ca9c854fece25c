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


@pytest.mark.parametrize("depth", [1, 2, 6])
def test_batch_at_other_depths(hip, depth):
    problems, eps = candidates("single", steps=203)
    tune = Tune(temporal=depth)
    for p in problems:
        p.tune = p.rtune = tune
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


@pytest.mark.parametrize("precision", ["single", "double"])
def test_batch_design_loop_session(hip, precision):
    """A design loop on one BatchAdjointProblem (the batch, plans, graphs and buffers of
    the first gradient kept; each candidate's cb replaced in place) gives the same bits
    as a fresh batch and as each candidate alone, for a new permittivity and for the
    first one again."""
    problems, eps = candidates(precision, steps=203)
    rng = np.random.default_rng(11)
    eps2 = [e + rng.uniform(-0.5, 0.5, size=e.shape) for e in eps]
    loop = BatchAdjointProblem(problems)
    print(f"{len(problems)} candidates, {precision}, design loop", flush=True)
    first = loop.run(eps)
    second = loop.run(eps2)
    third = loop.run(eps)
    assert loop._session is not None and loop._session.rev is not None
    fresh_J, fresh_g = BatchAdjointProblem(problems, reuse=False).run(eps2)
    for k, (p, e) in enumerate(zip(problems, eps2)):
        alone = AdjointProblem(p.sim, ["out"], checkpoint_every=64, reuse=False)
        J1, g1 = alone.run(e)
        assert second[0][k] == fresh_J[k] == J1, f"candidate {k}: J"
        assert np.array_equal(second[1][k], fresh_g[k]), f"candidate {k}: gradient vs fresh batch"
        assert np.array_equal(second[1][k], g1), f"candidate {k}: gradient vs alone"
        assert third[0][k] == first[0][k], f"candidate {k}: J after the loop"
        assert np.array_equal(third[1][k], first[1][k]), f"candidate {k}: gradient after the loop"
    loop.close()


def test_batch_store_mode_equals_checkpoint_mode(hip):
    """Store mode (each design box's Hz kept for every step of the forward pass, no
    recompute) gives the same bits as each candidate alone in store mode, also on the
    second gradient of a design loop, and the checkpointed gradient to float64
    rounding (the two sweeps group the reverse steps into different temporal blocks,
    so the float64 sums of dJ/dcb are grouped differently)."""
    problems, eps = candidates("single", steps=203)
    stored = [AdjointProblem(p.sim, ["out"], checkpoint_every=64, mode="store")
              for p in problems]
    rng = np.random.default_rng(5)
    eps2 = [e + rng.uniform(-0.5, 0.5, size=e.shape) for e in eps]
    loop = BatchAdjointProblem(stored)
    assert loop.mode == "store"
    print(f"{len(problems)} candidates, store mode", flush=True)
    loop.run(eps)
    Js, grads = loop.run(eps2)
    Jc, gc = BatchAdjointProblem(problems, reuse=False).run(eps2)
    for k, (p, e) in enumerate(zip(stored, eps2)):
        J1, g1 = p.run(e)
        assert Js[k] == Jc[k] == J1, f"candidate {k}: J"
        assert np.array_equal(grads[k], g1), f"candidate {k}: store batch vs alone"
        scale = np.abs(gc[k]).max()
        assert np.abs(grads[k] - gc[k]).max() <= 1e-12 * scale, f"candidate {k}: vs checkpoint"
    loop.close()


def test_batch_with_16_row_forward_tiles(hip):
    """16-row forward tiles in the forward plan and in the recompute half of the
    co-plan's joint launches: the same bits as each candidate alone at the default
    8-row shape (tile shapes do not change the arithmetic)."""
    problems, eps = candidates("single", steps=203)
    print(f"{len(problems)} candidates, band 16", flush=True)
    Js, grads = BatchAdjointProblem(problems, tune=Tune(band_rows=16), reuse=False).run(eps)
    for k, (p, e) in enumerate(zip(problems, eps)):
        J1, g1 = p.run(e)
        assert Js[k] == J1 and np.array_equal(grads[k], g1), f"candidate {k}"
