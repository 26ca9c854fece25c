"""Drive libetalon_hip's host code through the package the way the benchmarks do, with
the library built against the HIP runtime stub (build.py) under AddressSanitizer.

    ETALON_AMD_LIB=<stub library> LD_PRELOAD=<asan runtime> python lifecycle_workload.py

No kernel runs (the stub launches nothing), so the numbers are meaningless and a plan
whose creation checks read flags that no kernel wrote may fall back to the per-step
primitives (another host path, also exercised); what runs is
every host path of a plan's life — creation in every launch mode and shape the
autotuner tries, graph capture and replay, table refills, timing with events, teardown
between candidates — and of the adjoint's sessions, checkpoints, reverse plans and
co-plans, the batched adjoint and the observers.  The process exits 0 only if nothing
faulted; AddressSanitizer aborts it with a report on the first bad host access.
"""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))


def step(name: str) -> None:
    print(f"lifecycle: {name}", flush=True)


def batched(precision: str) -> None:
    """The batched benchmark's host sequence (bench.py run_batched): an autotuned batch
    plan (every candidate created, timed and closed), the batch run, the same
    simulations one by one on the fused path and with the per-step primitives."""
    from etalon_amd.batch import BatchSolver
    from etalon_amd.solver import FDTDSolver
    from etalon_amd.workloads import uneven_batch

    sims = uneven_batch(3, seed=3, lo=80, hi=120, steps=300, precision=precision)
    batch = BatchSolver(sims, autotune=True)
    batch.advance(300)
    batch.backend.synchronize()
    for s in batch.solvers:
        s.fields()
    del batch
    for sim in sims:
        solver = FDTDSolver(sim)
        solver.advance(211)
        solver.advance(89)
        solver.result()
        for _ in range(3):
            solver.step()
        del solver


def single() -> None:
    """One solver: short and long advances (graph chunks, tails), reset and rerun, host
    writes that rebuild the plan, the event-timed launches, observers."""
    from etalon_amd import observe
    from etalon_amd.hip.plan import Tune
    from etalon_amd.solver import FDTDSolver
    from etalon_amd.workloads import throughput_sim

    sim = throughput_sim("single", 300)
    for tune in (None, Tune(launch=3, band_rows=8, temporal=4), Tune(temporal=1, use_graph=0)):
        solver = FDTDSolver(sim, tune=tune)
        for count in (5, 20, 193, 7):
            solver.advance(count)
        plan = solver._plan
        if plan is None:  # the plan's checks read flags no stub kernel wrote: steps
            continue
        n = solver.n
        delta, phasors = solver._tables(n, 192)
        if delta is not None:
            plan.source_table(n, delta)
        for mon_id, table in zip(solver._plan_ids, phasors):
            if table is not None:
                plan.phasor_table(mon_id, n, table)
        plan.time(n, 192)
        solver._plan_tables = (n, n + 192)
        plan.bind()
        solver.n += 192
        solver.state.hz.set(np.zeros(solver.state.hz.shape, solver.state.hz.dtype))
        solver.advance(30)
        solver.reset()
        solver.advance(64)
        observe.field_energy(solver)
        observe.snapshot(solver, 4096)
        observe.energy_sums(solver.state)
        del solver


def adjoint() -> None:
    """A design loop on a small adjoint problem: session reuse (checkpointed, joint
    schedule), a cold problem, store mode, the sequential schedule, a stale tape."""
    from etalon_amd.adjoint import AdjointProblem
    from etalon_amd.workloads import adjoint_design_sim

    sim, design, names = adjoint_design_sim("single", steps=300, interior=120, region=1.0,
                                            pixels=10)
    eps = np.asarray(design.permittivity)
    prob = AdjointProblem(sim, names, checkpoint_every=32)
    for k in range(3):
        tape = prob.forward(eps + 0.01 * k)
        prob.backward(tape)
    old = prob.forward(eps)
    prob.forward(eps + 0.02)
    try:
        prob.backward(old)  # a tape of an earlier forward: refused or replayed
    except Exception:
        pass
    for kw in ({"reuse": False}, {"mode": "store"}, {"schedule": "sequential"}):
        p = AdjointProblem(sim, names, checkpoint_every=32, **kw)
        p.backward(p.forward(eps))
        del p


def batched_adjoint() -> None:
    """The batched adjoint's forward and backward passes over three candidates, twice
    (the session kept), then store mode."""
    from etalon_amd.adjoint import AdjointProblem
    from etalon_amd.batch_adjoint import BatchAdjointProblem
    from etalon_amd.workloads import phc_adjoint_candidates

    sims, eps, names = phc_adjoint_candidates(n=3, steps=200, interior=200, region=1.0,
                                              pixels=10)
    batch = BatchAdjointProblem([AdjointProblem(s, names) for s in sims])
    for _ in range(2):
        batch.backward(batch.forward(eps))
    batch.close()
    stored = BatchAdjointProblem([AdjointProblem(s, names, mode="store") for s in sims])
    stored.backward(stored.forward(eps))
    stored.close()


if __name__ == "__main__":
    import os

    os.environ.setdefault("ETALON_AMD_TUNE_CACHE", "/nonexistent/tune.json")
    for name, fn in (("single", single), ("batched single", lambda: batched("single")),
                     ("batched double", lambda: batched("double")), ("adjoint", adjoint),
                     ("batched adjoint", batched_adjoint)):
        step(name)
        fn()
    step("done")
