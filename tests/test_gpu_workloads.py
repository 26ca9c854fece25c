"""The measured workloads at their own sizes, against the oracle.

* Batched uneven runs: the 64 seeded simulations of the batched benchmark
  (interior 200-1500 cells per side, bends and rings, 1-8 monitors at 11
  wavelengths) advanced together; the smallest, the median and the largest equal
  the NumPy oracle bit for bit — fields, psi and every monitor accumulator — in
  float32, and in float64 as a batch of their own.
* Adjoint gradient: the 424 x 424 design problem (40 x 40 pixels, 4000 steps,
  checkpoints every 64 steps): the float64 HIP gradient equals the float64 NumPy
  adjoint gradient (tests/golden/adjoint_cfg4.npz, written by
  tests/golden/make_adjoint_golden.py) to float64 rounding, and the float32
  gradient agrees with float64 central differences on 16 random pixels to 1e-3.
* Batched adjoint: photonic-crystal candidates at 1024 x 1024 (80 x 80 pixels,
  2048 steps, checkpoints every 64): the batch's gradients equal each candidate's
  own AdjointProblem bit for bit; candidate 0's float64 gradient equals the NumPy
  adjoint (tests/golden/adjoint_cfg5.npz) and float64 central differences, and its
  float32 gradient the float64 one to 1e-3.  At the benchmark's full size (32
  candidates, 8000 steps) the first and last candidates equal their own problems bit
  for bit and candidate 0 agrees with float64 central differences.
"""

from pathlib import Path

import numpy as np
import pytest

from oracle import te_numpy as ref
from tests.sims import bitwise_equal

pytestmark = pytest.mark.gpu
GOLDEN = Path(__file__).resolve().parent / "golden"
STATE = ("ex", "ey", "hz", "psi_ex_y", "psi_ey_x", "psi_hz_x", "psi_hz_y")
BATCH_STEPS = 240


def _picks(sims):
    from etalon_amd.grid import build_yee_grid

    cells = [build_yee_grid(s).n_cells for s in sims]
    order = np.argsort(cells)
    return [int(order[0]), int(order[len(order) // 2]), int(order[-1])]


def _same_as_oracle(solver, sim, steps):
    o = ref.Solver(sim)
    for _ in range(steps):
        o.step()
    for name in STATE:
        assert bitwise_equal(getattr(solver.state, name).get(), getattr(o.s, name)), name
    nonzero = 0
    for k, (mon, acc) in enumerate(zip(solver.monitors, o.acc)):
        assert bitwise_equal(mon.acc.get(), acc), f"monitor {k}"
        nonzero += bool(np.abs(acc).max() > 0)
    return nonzero, len(o.acc)


@pytest.mark.parametrize("precision", ["single", "double"])
def test_batched_workload_equals_the_oracle(hip, precision):
    from etalon_amd.batch import BatchSolver
    from etalon_amd.workloads import uneven_batch

    sims = uneven_batch(64, seed=0, steps=4000)
    picks = _picks(sims)
    if precision == "double":
        sims = [uneven_batch(64, seed=0, steps=4000, precision="double")[k] for k in picks]
        picks = [0, 1, 2]
    batch = BatchSolver(sims)
    batch.advance(BATCH_STEPS)
    for k in picks:
        nonzero, total = _same_as_oracle(batch.solvers[k], sims[k], BATCH_STEPS)
        g = batch.solvers[k].grid
        print(f"{precision} simulation {k} ({g.nx} x {g.ny}): bitwise, {nonzero}/{total} "
              f"monitor components nonzero", flush=True)


@pytest.fixture(scope="module")
def cfg4():
    from etalon_amd.workloads import adjoint_design_sim

    golden = np.load(GOLDEN / "adjoint_cfg4.npz")
    sim32, design, names = adjoint_design_sim("single", seed=int(golden["seed"]))
    sim64, _, _ = adjoint_design_sim("double", seed=int(golden["seed"]))
    return sim32, sim64, np.asarray(design.permittivity), names, golden


def test_adjoint_workload_float64_equals_the_numpy_adjoint(hip, cfg4):
    from etalon_amd.adjoint import AdjointProblem

    _, sim64, eps, names, golden = cfg4
    prob = AdjointProblem(sim64, names)
    assert (prob.grid.nx, prob.grid.ny, prob.n_steps, prob.K) == (424, 424, 4000, 64)
    assert eps.shape == (40, 40) and int(golden["steps"]) == 4000
    J, g = prob.run(eps)
    g_ref = golden["gradient"]
    err = np.abs(g - g_ref).max() / np.abs(g_ref).max()
    print(f"J {J:.12e} (NumPy {float(golden['J']):.12e}); max |g - g_numpy| / max = {err:.2e}")
    assert J == pytest.approx(float(golden["J"]), rel=1e-11)
    assert err < 1e-9


def test_adjoint_workload_float32_against_central_differences(hip, cfg4):
    from etalon_amd.adjoint import AdjointProblem

    sim32, sim64, eps, names, golden = cfg4
    _, g = AdjointProblem(sim32, names).run(eps)
    prob64 = AdjointProblem(sim64, names)
    scale = np.abs(golden["gradient"]).max()
    rng = np.random.default_rng(1)
    worst = 0.0
    for flat in rng.choice(eps.size, size=16, replace=False):
        p, q = divmod(int(flat), eps.shape[1])
        h = 1e-4 * eps[p, q]
        plus, minus = eps.copy(), eps.copy()
        plus[p, q] += h
        minus[p, q] -= h
        fd = (prob64.forward(plus).J - prob64.forward(minus).J) / (2 * h)
        rel = abs(g[p, q] - fd) / scale
        worst = max(worst, rel)
        assert rel < 1e-3, (p, q, g[p, q], fd)
    print(f"16 pixels: max |g32 - fd64| / max |g| = {worst:.2e}")


@pytest.fixture(scope="module")
def cfg5():
    from etalon_amd.workloads import phc_adjoint_candidates

    golden = np.load(GOLDEN / "adjoint_cfg5.npz")
    steps, seed = int(golden["steps"]), int(golden["seed"])
    sims, eps, names = phc_adjoint_candidates(n=3, seed=seed, steps=steps)
    sims64, eps64, _ = phc_adjoint_candidates(n=1, seed=seed, steps=steps, precision="double")
    assert np.array_equal(eps[0], eps64[0])
    return sims, eps, sims64[0], names, golden


def test_batched_adjoint_workload_equals_each_candidate_alone(hip, cfg5):
    from etalon_amd.adjoint import AdjointProblem
    from etalon_amd.batch_adjoint import BatchAdjointProblem

    from etalon_amd.hip.plan import Tune

    sims, eps, _, names, golden = cfg5
    # the same tuning on both sides (the batch alone would pick depth 6 and 16-row
    # tiles, one 1024^2 candidate depth 4: the float64 dJ/dcb sums then group the
    # reverse steps differently and agree to rounding only)
    tune = Tune(temporal=6, band_rows=16)
    problems = [AdjointProblem(s, names, tune=tune) for s in sims]
    assert (problems[0].grid.nx, problems[0].K, eps[0].shape) == (1024, 64, (80, 80))
    Js, grads = BatchAdjointProblem(problems, tune=tune).run(eps)
    for k, (sim, e) in enumerate(zip(sims, eps)):
        J, g = AdjointProblem(sim, names, tune=tune).run(e)
        assert Js[k] == J, k
        assert np.array_equal(grads[k], g), k
    g_ref = golden["gradient"]
    err = np.abs(grads[0] - g_ref).max() / np.abs(g_ref).max()
    print(f"candidate 0: float32 against the float64 NumPy adjoint {err:.2e}")
    assert err < 1e-3


def test_batched_adjoint_workload_float64_candidate(hip, cfg5):
    """Candidate 0 in float64: the NumPy adjoint's gradient, and central differences on
    8 random pixels (the 16 perturbed runs advanced as one float64 batch)."""
    from etalon_amd.adjoint import AdjointProblem
    from etalon_amd.batch import BatchSolver

    _, _, sim64, names, golden = cfg5
    prob = AdjointProblem(sim64, names)
    eps = np.asarray(sim64.structures[-1].permittivity)
    J, g = prob.run(eps)
    g_ref = golden["gradient"]
    scale = np.abs(g_ref).max()
    err = np.abs(g - g_ref).max() / scale
    print(f"J {J:.12e} (NumPy {float(golden['J']):.12e}); max |g - g_numpy| / max = {err:.2e}")
    assert J == pytest.approx(float(golden["J"]), rel=1e-11)
    assert err < 1e-9
    rng = np.random.default_rng(2)
    picks = rng.choice(eps.size, size=8, replace=False)
    runs = []
    for flat in picks:
        p, q = divmod(int(flat), eps.shape[1])
        for sign in (1.0, -1.0):
            e = eps.copy()
            e[p, q] += sign * 1e-4 * eps[p, q]
            runs.append(prob.with_permittivity(e))
    fd_batch = BatchSolver(runs)
    fd_batch.run()
    Jfd = [prob._objective(s) for s in fd_batch.solvers]
    for k, flat in enumerate(picks):
        p, q = divmod(int(flat), eps.shape[1])
        fd = (Jfd[2 * k] - Jfd[2 * k + 1]) / (2 * 1e-4 * eps[p, q])
        assert abs(g[p, q] - fd) / scale < 1e-3, (p, q, g[p, q], fd)


def test_batched_adjoint_workload_at_full_size(hip):
    """The benchmark's batch as it runs: 32 candidates, 8000 steps (125 checkpoint
    segments, window checkpoints).  The first and last candidates' gradients equal their
    own AdjointProblem's bit for bit (same tune), and candidate 0's float32 gradient
    agrees with float64 central differences on 8 random pixels to 1e-3 of its largest
    entry (the 16 perturbed float64 runs advanced as one batch)."""
    from etalon_amd.adjoint import AdjointProblem
    from etalon_amd.batch import BatchSolver
    from etalon_amd.batch_adjoint import BatchAdjointProblem
    from etalon_amd.hip.plan import Tune
    from etalon_amd.workloads import phc_adjoint_candidates

    sims, eps, names = phc_adjoint_candidates(n=32, seed=0, steps=8000)
    tune = Tune(temporal=6, band_rows=16)
    batch = BatchAdjointProblem([AdjointProblem(s, names, tune=tune) for s in sims], tune=tune)
    assert (batch.n_steps, batch.K) == (8000, 64)
    Js, grads = batch.run(eps)
    batch.close()
    for k in (0, 31):
        J, g = AdjointProblem(sims[k], names, tune=tune, reuse=False).run(eps[k])
        assert Js[k] == J, k
        assert np.array_equal(grads[k], g), k
    sims64, eps64, _ = phc_adjoint_candidates(n=1, seed=0, steps=8000, precision="double")
    assert np.array_equal(eps64[0], eps[0])
    prob64 = AdjointProblem(sims64[0], names)
    scale = np.abs(grads[0]).max()
    rng = np.random.default_rng(3)
    picks = rng.choice(eps[0].size, size=8, replace=False)
    runs = []
    for flat in picks:
        p, q = divmod(int(flat), eps[0].shape[1])
        for sign in (1.0, -1.0):
            e = eps[0].copy()
            e[p, q] += sign * 1e-4 * eps[0][p, q]
            runs.append(prob64.with_permittivity(e))
    fd_batch = BatchSolver(runs)
    fd_batch.run()
    Jfd = [prob64._objective(s) for s in fd_batch.solvers]
    worst = 0.0
    for k, flat in enumerate(picks):
        p, q = divmod(int(flat), eps[0].shape[1])
        fd = (Jfd[2 * k] - Jfd[2 * k + 1]) / (2 * 1e-4 * eps[0][p, q])
        rel = abs(grads[0][p, q] - fd) / scale
        worst = max(worst, rel)
        assert rel < 1e-3, (p, q, grads[0][p, q], fd)
    print(f"32 candidates x 8000 steps: max |g32 - fd64| / max |g| = {worst:.2e}", flush=True)
