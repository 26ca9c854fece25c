"""Permittivity gradients of many design candidates at once.

``BatchAdjointProblem(problems)`` takes one ``AdjointProblem`` per candidate (any
grid sizes, sources, monitors and design regions; one precision, one step count)
and computes every candidate's J and ∂J/∂ε with the same checkpointed discrete
adjoint as ``AdjointProblem``, batched:

1. forward: one fused plan advances all candidates (one tile work-list, one DFT
   fold), saving every candidate's state at each segment start;
2. backward: a recompute plan over all candidates (sources and the Hz probe of
   each design box) and an adjoint plan over all candidates form a co-plan; its
   sweep restores every candidate's checkpoint of segment k - 1 and then, per step
   block, ONE launch of the joint kernel advances the recompute of segment k - 1
   and the reverse steps of segment k of every candidate.

Each candidate's gradient is bit-identical to its own ``AdjointProblem`` (tile
kernels, same tuning).  No reference counterpart (the reference has no adjoint).
"""

from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from etalon_amd.adjoint import AdjointProblem, _Reverse, recompute_tune
from etalon_amd.batch import BatchSolver
from etalon_amd.hip import _lib
from etalon_amd.hip.backend import zeros
from etalon_amd.state import FIELDS


@dataclass
class BatchTape:
    batch: BatchSolver
    J: list = field(default_factory=list)
    checkpoints: list = field(default_factory=list)  # per candidate: field -> (n_ck, size)
    forward_s: float = 0.0


class BatchAdjointProblem:
    def __init__(self, problems: Sequence[AdjointProblem], tune=None) -> None:
        if not problems:
            raise ValueError("BatchAdjointProblem needs at least one candidate")
        p0 = problems[0]
        for p in problems:
            if p.kernel != "tile" or p.mode != "checkpoint":
                raise ValueError("batched candidates use the tile kernels in checkpoint mode")
            if p.n_steps != p0.n_steps or p.K != p0.K or p.sim.precision != p0.sim.precision:
                raise ValueError("batched candidates share the step count, the checkpoint "
                                 "interval and the precision")
        self.problems = list(problems)
        self.tune = tune
        self.n_steps, self.K = p0.n_steps, p0.K

    def forward(self, eps_list: Sequence[np.ndarray]) -> BatchTape:
        """Run every candidate with its design permittivity; J per candidate and the tape."""
        if len(eps_list) != len(self.problems):
            raise ValueError("one permittivity array per candidate")
        sims = [p.with_permittivity(e) for p, e in zip(self.problems, eps_list)]
        batch = BatchSolver(sims, tune=self.tune)
        for p, s in zip(self.problems, batch.solvers):
            p._check_box_off_cpml(s)
        N, K = self.n_steps, self.K
        t0 = time.perf_counter()
        plan = batch._ensure_plan()
        tape = BatchTape(batch=batch)
        n_ck = (N + K - 1) // K
        for k, (s, ids) in enumerate(zip(batch.solvers, batch._ids)):
            delta, phasors = s._tables(0, N)
            if delta is not None:
                plan.source_table(0, delta, sim=k)
            for mon_id, table in zip(ids, phasors):
                if table is not None:
                    plan.phasor_table(mon_id, 0, table, sim=k)
            ck = {name: zeros((n_ck, getattr(s.state, name).size), s.real) for name in FIELDS}
            plan.set_checkpoints([ck[name] for name in FIELDS], n_ck, sim=k)
            tape.checkpoints.append(ck)
        plan.advance_checkpointed(0, N, K)  # slot k = every candidate's state at step k K
        plan.bind()
        for s in batch.solvers:
            s.n = N
        batch.n = N
        batch.backend.synchronize()
        tape.forward_s = time.perf_counter() - t0
        tape.J = [p._objective(s) for p, s in zip(self.problems, batch.solvers)]
        return tape

    def backward(self, tape: BatchTape) -> list:
        """∂J/∂ε per design pixel (float64) of every candidate."""
        from etalon_amd.hip.plan import FusedPlan

        solvers = tape.batch.solvers
        N, K = self.n_steps, self.K
        hists = [zeros((2 * K, p._box_hz_nodes().size), s.real)
                 for p, s in zip(self.problems, solvers)]
        bufs = [p._adjoint_buffers(s, h) for p, s, h in zip(self.problems, solvers, hists)]
        rev = _Reverse("tile", [b.desc for b in bufs], self.tune)
        plan: Optional[FusedPlan] = None
        co = C.c_void_p()
        try:
            for k, (p, s) in enumerate(zip(self.problems, solvers)):
                comp, flat, table = p._sources(s)
                rev.sources(comp, flat, table, sim=k)
            plan = FusedPlan([s.state for s in solvers], recompute_tune(self.tune))
            for k, (p, s, h) in enumerate(zip(self.problems, solvers, hists)):
                p._configure_recompute(plan, s, k, h)
                ck = tape.checkpoints[k]
                plan.set_checkpoints([ck[name] for name in FIELDS], ck[FIELDS[0]].shape[0], sim=k)
            _lib.call("etl_coplan_create", plan.handle, rev.handle, C.byref(co))
            _lib.call("etl_coplan_sweep", co, N, K)
            tape.batch.backend.synchronize()
        finally:
            if co:
                _lib.call("etl_coplan_destroy", co)
            if plan is not None:
                plan.close()
            rev.close()
        return [p._pixel_gradient(s, b) for p, s, b in zip(self.problems, solvers, bufs)]

    def run(self, eps_list: Sequence[np.ndarray]) -> tuple[list, list]:
        """(J per candidate, ∂J/∂ε per candidate)."""
        tape = self.forward(eps_list)
        return tape.J, self.backward(tape)
