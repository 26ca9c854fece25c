"""Batched plans: every simulation of a batch of uneven grids must come out bit for
bit as when it runs alone — fields, ψ and monitor accumulators — whatever the tile
shape or workgroup size, with or without the autotuner, and float64 must equal the
NumPy oracle."""

import numpy as np
import pytest

from etalon_amd.batch import BatchSolver, run_batch
from etalon_amd.hip.plan import Tune
from etalon_amd.solver import FDTDSolver
from etalon_amd.workloads import uneven_batch
from oracle import te_numpy as ref
from tests.sims import bitwise_equal

pytestmark = pytest.mark.gpu
STATE = ("ex", "ey", "hz", "psi_ex_y", "psi_ey_x", "psi_hz_x", "psi_hz_y")


def small_batch(precision="single", n=5, steps=300, seed=3):
    return uneven_batch(n=n, seed=seed, lo=80, hi=170, steps=steps, precision=precision)


def snapshot(solver):
    out = {name: getattr(solver.state, name).get() for name in STATE}
    for k, mon in enumerate(solver.monitors):
        out[f"acc{k}"] = mon.acc.get()
    return out


def alone(sim):
    s = FDTDSolver(sim)
    s.advance(s.n_steps)
    return snapshot(s)


def assert_same(a, b, what):
    assert a.keys() == b.keys()
    for key in a:
        assert bitwise_equal(a[key], b[key]), f"{what}: {key} differs"


@pytest.mark.parametrize("precision", ["single", "double"])
@pytest.mark.parametrize("shape", [(1, 4, 8, 1, 3), (1, 4, 8, 4, 3), (1, 2, 16, 2, 3),
                                   (1, 1, 8, 1, 3), (1, 4, 8, 1, 1), (1, 1, 8, 1, 1),
                                   (1, 4, 16, 1, 1), (1, 6, 16, 1, 1), (1, 6, 8, 1, 1)])
def test_batch_equals_each_simulation_alone(hip, precision, shape):
    vec, depth, band, waves, launch = shape
    sims = small_batch(precision)
    print(f"batch of {len(sims)} {precision}, shape vec={vec} T={depth} band={band} waves={waves} "
          f"launch={launch}", flush=True)
    batch = BatchSolver(sims, tune=Tune(band_rows=band, vec=vec, temporal=depth, waves=waves,
                                        launch=launch))
    batch.advance(batch.n_steps)
    for k, (sim, solver) in enumerate(zip(sims, batch.solvers)):
        assert_same(snapshot(solver), alone(sim), f"simulation {k}")


def test_batch_in_chunks_and_with_autotune(hip, tmp_path, monkeypatch):
    monkeypatch.setenv("ETALON_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    sims = small_batch("single", n=4, steps=257, seed=11)
    batch = BatchSolver(sims, autotune=True, autotune_steps=32)
    assert batch.tune_report and not batch.tune_report["cached"]
    for count in (3, 64, 100, 90):  # tails, whole chunks, graph replays
        batch.advance(count)
    for k, (sim, solver) in enumerate(zip(sims, batch.solvers)):
        assert_same(snapshot(solver), alone(sim), f"simulation {k}")
    again = BatchSolver(sims, autotune=True)  # now from the cache
    assert again.tune_report["cached"]


def test_double_batch_equals_the_oracle(hip):
    sims = small_batch("double", n=3, steps=120, seed=5)
    batch = BatchSolver(sims, tune=Tune(band_rows=8, vec=1, temporal=4, waves=2))
    batch.advance(batch.n_steps)
    for sim, solver in zip(sims, batch.solvers):
        o = ref.Solver(sim)
        for _ in range(120):
            o.step()
        for name in STATE:
            assert bitwise_equal(getattr(solver.state, name).get(), getattr(o.s, name)), name


def test_run_batch_groups_by_step_count(hip):
    a = small_batch("single", n=2, steps=80, seed=1)
    b = small_batch("single", n=2, steps=50, seed=2)
    results = run_batch([a[0], b[0], a[1], b[1]], autotune=False)
    assert [r.metadata["n_steps"] for r in results] == [80, 50, 80, 50]
    solo = FDTDSolver(b[1]).run()
    for name in solo.monitors:
        for comp, data in solo[name].items():
            assert bitwise_equal(results[3][name][comp].values, data.values)


def _probe_run(solvers, probes, steps):
    """Advance ``solvers`` (each its own simulation) in one fused plan for ``steps``
    steps with the time-domain probe ``probes[k]`` = (component, flat nodes) on
    simulation k; returns each probe's (steps, nodes) rows."""
    from etalon_amd.hip.backend import zeros
    from etalon_amd.hip.plan import FusedPlan

    plan = FusedPlan([s.state for s in solvers])
    outs = []
    for k, (s, (comp, flat)) in enumerate(zip(solvers, probes)):
        ids = s.register(plan, k)
        delta, phasors = s._tables(0, steps)
        if delta is not None:
            plan.source_table(0, delta, sim=k)
        for mon_id, table in zip(ids, phasors):
            if table is not None:
                plan.phasor_table(mon_id, 0, table, sim=k)
        out = zeros((steps, flat.size), s.real)
        plan.add_probe(comp, flat, out, sim=k)
        outs.append(out)
    plan.advance(0, steps)
    rows = [o.get() for o in outs]
    plan.close()
    return rows


@pytest.mark.parametrize("precision", ["single", "double"])
def test_probes_of_every_simulation_in_one_launch(hip, precision):
    """Time-domain probes of different sizes on three simulations of one plan — one
    launch copies all of them after each trace chunk, graph-replayed chunks and the
    directly launched tail alike — hold, row for row and bit for bit, what the same
    probe records on each simulation's own plan."""
    sims = small_batch(precision, n=3, steps=400)
    probes = [("Hz", np.arange(37, dtype=np.int64) * 3 + 200),
              ("Ex", np.arange(5, dtype=np.int64) * 11 + 300),
              ("Ey", np.arange(61, dtype=np.int64) + 1000)]
    steps = 400  # trace chunks of 64 (graphs) and a tail of 16
    together = _probe_run([FDTDSolver(s) for s in sims], probes, steps)
    for k, sim in enumerate(sims):
        alone_rows = _probe_run([FDTDSolver(sim)], [probes[k]], steps)[0]
        assert np.abs(alone_rows).max() > 0, f"probe {k} saw nothing"
        assert bitwise_equal(together[k], alone_rows), f"probe {k} differs"

