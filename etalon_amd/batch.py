"""Many independent simulations advanced together on one GPU.

``BatchSolver(sims)`` builds every simulation's state exactly as ``FDTDSolver``
does and advances all of them with one fused plan: per T steps one launch of each
tile kernel covers the tiles of every grid (a work-list of tile descriptors, each
naming its simulation), and one DFT fold covers every monitor.  Small grids that
could not fill the GPU alone fill it together, and the per-launch overhead is paid
once for the whole batch.  Every simulation's fields and monitors are bit-identical
to running it alone (``FDTDSolver(sim).run()``), which is bit-identical to the
reference's ``step()`` loop.

The simulations must share the precision and the number of steps (they advance in
lockstep); ``run_batch`` groups arbitrary simulations by step count.  The reference
has no counterpart: it runs one ``FDTDSolver`` per simulation
(etalon/fdtd/solver.py:49-165).
"""

from __future__ import annotations

import time
from typing import Optional, Sequence

from etalon_amd.backends import get_backend
from etalon_amd.grid import build_yee_grid
from etalon_amd.solver import _FIRST_TABLE_ROWS, _PLAN_TABLE_ROWS, FDTDSolver


class BatchSolver:
    def __init__(self, sims: Sequence, tune=None, autotune: bool = False,
                 autotune_steps: int = 128) -> None:
        if not sims:
            raise ValueError("BatchSolver needs at least one simulation")
        precisions = {s.precision for s in sims}
        if len(precisions) != 1:
            raise ValueError(f"a batch runs one precision; got {sorted(precisions)}")
        steps = {build_yee_grid(sim).n_steps(sim.run_time) for sim in sims}
        if len(steps) != 1:
            raise ValueError(f"a batch advances in lockstep; step counts differ: {sorted(steps)} "
                             "(use run_batch to group them)")
        backend = get_backend("hip", sims[0].precision)
        self.solvers = [FDTDSolver(sim, backend=backend) for sim in sims]
        for s in self.solvers:
            if not s.fused_supported():
                raise ValueError("a batched simulation takes at most 16 source nodes")
        self.backend = backend
        self.n_steps = self.solvers[0].n_steps
        self.n = 0
        self.tune_report: Optional[dict] = None
        if tune is None and autotune:
            from etalon_amd.autotune import autotune as _autotune

            tune, self.tune_report = _autotune(self.solvers, steps=autotune_steps)
        self.tune = tune
        self._plan = None
        self._ids: list[list[int]] = []
        # A batch always runs on the fused plan: build it (second field buffers,
        # tile work-list, monitor layout) with the states rather than in the first
        # advance.
        self._ensure_plan()

    @property
    def cells(self) -> int:
        """Cells of all grids together (PML included)."""
        return sum(s.grid.nx * s.grid.ny for s in self.solvers)

    def _ensure_plan(self):
        if self._plan is None:
            from etalon_amd.hip.plan import FusedPlan

            plan = FusedPlan([s.state for s in self.solvers], self.tune)
            self._ids = [s.register(plan, k) for k, s in enumerate(self.solvers)]
            plan.prepare()
            self._plan = plan
        return self._plan

    def advance(self, count: int) -> None:
        """Advance every simulation ``count`` steps."""
        if count <= 0:
            return
        plan = self._ensure_plan()
        done = 0
        grow = _FIRST_TABLE_ROWS  # tables of the next rows built while the GPU runs these
        while done < count:
            rows = min(count - done, grow)
            grow = min(4 * grow, _PLAN_TABLE_ROWS)
            for k, (s, ids) in enumerate(zip(self.solvers, self._ids)):
                delta, phasors = s._tables(self.n, rows)
                if delta is not None:
                    plan.source_table(self.n, delta, sim=k)
                for mon_id, table in zip(ids, phasors):
                    if table is not None:
                        plan.phasor_table(mon_id, self.n, table, sim=k)
            plan.advance(self.n, rows)
            plan.bind()
            for s in self.solvers:
                s.n += rows
            self.n += rows
            done += rows

    def run(self) -> list:
        """Run every simulation to its run_time; one SimulationResult each (the wall
        time in each result's metadata is the batch's)."""
        start = time.perf_counter()
        self.advance(self.n_steps - self.n)
        self.backend.synchronize()
        wall = time.perf_counter() - start
        return [s.result(wall) for s in self.solvers]


def run_batch(sims: Sequence, autotune: bool = True) -> list:
    """Results of ``sims`` (any step counts and sizes), batched by step count."""
    groups: dict[tuple, list[int]] = {}
    for k, sim in enumerate(sims):
        steps = build_yee_grid(sim).n_steps(sim.run_time)
        groups.setdefault((sim.precision, steps), []).append(k)
    results: list = [None] * len(sims)
    for members in groups.values():
        batch = BatchSolver([sims[k] for k in members], autotune=autotune)
        for k, res in zip(members, batch.run()):
            results[k] = res
    return results
