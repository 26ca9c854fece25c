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
kernels, same tuning).  In a design loop (``reuse``, the default) the batch, plans,
graphs and buffers of the first gradient are kept: later gradients give each
candidate its new permittivity in place and run the same plans again.  No reference
counterpart (the reference has no adjoint).
"""

from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from etalon_amd.adjoint import AdjointProblem, _Reverse, design_tune, recompute_tune
from etalon_amd.batch import BatchSolver
from etalon_amd.hip import _lib
from etalon_amd.hip.backend import zeros
from etalon_amd.state import FIELDS


@dataclass
class BatchTape:
    batch: BatchSolver
    J: list = field(default_factory=list)
    checkpoints: list = field(default_factory=list)  # per candidate: field -> (n_ck, size)
    histories: list = field(default_factory=list)    # store mode: per candidate (N, box nodes)
    forward_s: float = 0.0
    eps: list = field(default_factory=list)          # the design permittivities it ran
    generation: int = 0                              # the session forward that made it


@dataclass
class _BatchSession:
    """What a design loop keeps between batched gradients of one set of geometries:
    the batch (states, forward plan, tables, checkpoint storage) and, after the first
    backward pass, the reverse plan, the recompute plan and their co-plan.  Every
    forward on it overwrites the batch's state; ``generation`` counts those forwards,
    so a tape of an earlier one is known to be stale."""

    batch: BatchSolver
    checkpoints: list
    histories: list
    generation: int = 1
    bufs: list = field(default_factory=list)
    rev: Optional[_Reverse] = None
    rplan: Optional[object] = None
    windows: list = field(default_factory=list)  # per candidate: its recompute window
    co: C.c_void_p = field(default_factory=C.c_void_p)

    def close(self) -> None:
        if self.co:
            _lib.call("etl_coplan_destroy", self.co)
            self.co = C.c_void_p()
        if self.rplan is not None:
            self.rplan.close()
            self.rplan = None
        if self.rev is not None:
            self.rev.close()
            self.rev = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchAdjointProblem:
    def __init__(self, problems: Sequence[AdjointProblem], tune=None, reuse: bool = True) -> None:
        if not problems:
            raise ValueError("BatchAdjointProblem needs at least one candidate")
        p0 = problems[0]
        for p in problems:
            if p.kernel != "tile":
                raise ValueError("batched candidates use the tile kernels")
            if (p.n_steps != p0.n_steps or p.K != p0.K or p.sim.precision != p0.sim.precision
                    or p.mode != p0.mode or p.window != p0.window):
                raise ValueError("batched candidates share the step count, the checkpoint "
                                 "interval, the precision, the mode and the window choice")
        self.problems = list(problems)
        # "checkpoint": segments recomputed beside the reverse steps (co-plan);
        # "store": the forward pass keeps every step's Hz on each design box and the
        # reverse steps read it (no recompute; N x box nodes values per candidate)
        self.mode = p0.mode
        self.tune = design_tune(tune)  # the forward pass
        self.rtune = recompute_tune(self.tune, p0.sim.precision)  # the recompute and reverse
        self.n_steps, self.K = p0.n_steps, p0.K
        # a design loop's gradients keep the batch, plans, graphs and buffers of the
        # first one: new permittivities only replace each candidate's cb_ex and cb_ey.
        # A tape is then valid until the next forward.
        self.reuse = bool(reuse)
        self._session: Optional[_BatchSession] = None

    def close(self) -> None:
        """Release the session (plans, graphs, buffers) of a design loop."""
        if self._session is not None:
            self._session.close()
            self._session = None

    def forward(self, eps_list: Sequence[np.ndarray]) -> BatchTape:
        """Run every candidate with its design permittivity; J per candidate and the tape."""
        if len(eps_list) != len(self.problems):
            raise ValueError("one permittivity array per candidate")
        if self.reuse and self._session is not None:
            return self._forward_again(eps_list)
        return self._forward_new(eps_list, session=self.reuse)

    def _forward_new(self, eps_list: Sequence[np.ndarray], session: bool) -> BatchTape:
        """The forward pass on a new batch; ``session``: keep it for a design loop."""
        sims = [p.with_permittivity(e) for p, e in zip(self.problems, eps_list)]
        batch = BatchSolver(sims, tune=self.tune)
        for p, s in zip(self.problems, batch.solvers):
            p._check_box_off_cpml(s)
        N, K = self.n_steps, self.K
        t0 = time.perf_counter()
        plan = batch._ensure_plan()
        tape = BatchTape(batch=batch, eps=[np.array(e, dtype=np.float64) for e in eps_list])
        n_ck = (N + K - 1) // K
        for k, (p, s, ids) in enumerate(zip(self.problems, batch.solvers, batch._ids)):
            delta, phasors = s._tables(0, N)
            if delta is not None:
                plan.source_table(0, delta, sim=k)
            for mon_id, table in zip(ids, phasors):
                if table is not None:
                    plan.phasor_table(mon_id, 0, table, sim=k)
            if self.mode == "store":
                nodes = p._box_hz_nodes()
                hist = zeros((N, nodes.size), s.real)
                plan.add_probe("Hz", nodes, hist, sim=k)
                tape.histories.append(hist)
            else:
                ck = p._checkpoint_storage(s, n_ck)
                p._set_forward_checkpoints(plan, ck, n_ck, k)
                tape.checkpoints.append(ck)
        self._advance_forward(plan)
        for s in batch.solvers:
            s.n = N
        batch.n = N
        batch.backend.synchronize()
        tape.forward_s = time.perf_counter() - t0
        tape.J = [p._objective(s) for p, s in zip(self.problems, batch.solvers)]
        if session:
            self._session = _BatchSession(batch=batch, checkpoints=tape.checkpoints,
                                          histories=tape.histories)
            tape.generation = self._session.generation
        return tape

    def _advance_forward(self, plan) -> None:
        if self.mode == "store":
            plan.advance(0, self.n_steps)  # probe row n = every candidate's box at step n
        else:
            # slot k = every candidate's state at step k K
            plan.advance_checkpointed(0, self.n_steps, self.K)
        plan.bind()

    def _forward_again(self, eps_list: Sequence[np.ndarray]) -> BatchTape:
        """The forward pass on the session's batch: every candidate takes its new design
        permittivity (``AdjointProblem.set_design``), and the same plan, tables and
        checkpoint storage run again."""
        ses = self._session
        batch = ses.batch
        t0 = time.perf_counter()
        for p, s, e in zip(self.problems, batch.solvers, eps_list):
            p.set_design(s, e)
        ses.generation += 1
        self._advance_forward(batch._ensure_plan())
        for s in batch.solvers:
            s.n = self.n_steps
        batch.n = self.n_steps
        batch.backend.synchronize()
        tape = BatchTape(batch=batch, checkpoints=ses.checkpoints, histories=ses.histories,
                         eps=[np.array(e, dtype=np.float64) for e in eps_list],
                         generation=ses.generation)
        tape.forward_s = time.perf_counter() - t0
        tape.J = [p._objective(s) for p, s in zip(self.problems, batch.solvers)]
        return tape

    def backward(self, tape: BatchTape) -> list:
        """∂J/∂ε per design pixel (float64) of every candidate.

        A tape of the design-loop session whose state a later forward has overwritten
        is replayed: its forward runs again at its own permittivities on a new batch,
        and the backward pass runs on that."""
        ses = self._session
        if self.reuse and ses is not None and tape.batch is ses.batch:
            if tape.generation != ses.generation:
                return self.backward(self._forward_new(tape.eps, session=False))
            return self._backward_session(tape)
        solvers = tape.batch.solvers
        rev, rplan, co, bufs = self._reverse_plans(tape)
        try:
            self._sweep(rev, co)
            tape.batch.backend.synchronize()
        finally:
            if co:
                _lib.call("etl_coplan_destroy", co)
            if rplan is not None:
                rplan.close()
            rev.close()
        return [p._pixel_gradient(s, b) for p, s, b in zip(self.problems, solvers, bufs)]

    def _sweep(self, rev, co) -> None:
        """Every reverse step: from the stored histories, or with the recompute of
        segment k - 1 beside the reverse steps of segment k (co-plan sweep)."""
        if self.mode == "store":
            rev.run(0, self.n_steps)
        else:
            _lib.call("etl_coplan_sweep", co, self.n_steps, self.K)

    def _backward_session(self, tape: BatchTape) -> list:
        """The backward pass on the session's plans: built on the first gradient, reset
        (λ and ∂J/∂cb to zero, the new adjoint-source tables in place) on the next."""
        ses = self._session
        solvers = tape.batch.solvers
        if ses.rev is None:
            ses.rev, ses.rplan, ses.co, ses.bufs = self._reverse_plans(tape)
            ses.windows = self._windows
        else:
            _lib.call("etl_adjplan_reset", ses.rev.handle)
            for k, (p, s, b) in enumerate(zip(self.problems, solvers, ses.bufs)):
                b.g_ex.fill_zero()
                b.g_ey.fill_zero()
                comp, flat, table = p._sources(s)
                ses.rev.sources(comp, flat, table, sim=k)
                if ses.windows and ses.windows[k] is not None:
                    p._refresh_window(s, ses.windows[k])
        self._sweep(ses.rev, ses.co)
        tape.batch.backend.synchronize()
        return [p._pixel_gradient(s, b) for p, s, b in zip(self.problems, solvers, ses.bufs)]

    def _reverse_plans(self, tape: BatchTape) -> tuple:
        """The adjoint plan over every candidate (λ, ∂J/∂cb, adjoint sources) and, in
        checkpoint mode, the recompute plan over every candidate (sources, the Hz probe
        of each design box into its history, the tape's checkpoints) and their co-plan.
        Returns (adjoint plan, recompute plan or None, co-plan or null, buffers)."""
        from etalon_amd.hip.plan import FusedPlan

        solvers = tape.batch.solvers
        K = self.K
        self._windows = []
        if self.mode == "store":
            hists = tape.histories
        else:
            hists = [zeros((2 * K, p._box_hz_nodes().size), s.real)
                     for p, s in zip(self.problems, solvers)]
        bufs = [p._adjoint_buffers(s, h) for p, s, h in zip(self.problems, solvers, hists)]
        rev = _Reverse("tile", [b.desc for b in bufs], self.rtune)
        plan: Optional[FusedPlan] = None
        co = C.c_void_p()
        try:
            for k, (p, s) in enumerate(zip(self.problems, solvers)):
                comp, flat, table = p._sources(s)
                rev.sources(comp, flat, table, sim=k)
            if self.mode == "store":
                return rev, None, co, bufs
            recompute = [p._recompute_state(s, rev, self.rtune)
                         for p, s in zip(self.problems, solvers)]
            windows = [w for _, _, w in recompute]
            plan = FusedPlan([st for st, _, _ in recompute], recompute[0][1])
            for k, (p, s, h, w) in enumerate(zip(self.problems, solvers, hists, windows)):
                p._configure_recompute(plan, s, k, h, w)
                ck = tape.checkpoints[k]
                p._set_recompute_checkpoints(plan, ck, ck[FIELDS[0]].shape[0], k, w)
            _lib.call("etl_coplan_create", plan.handle, rev.handle, C.byref(co))
        except BaseException:
            if plan is not None:
                plan.close()
            rev.close()
            raise
        self._windows = windows
        return rev, plan, co, bufs

    def run(self, eps_list: Sequence[np.ndarray]) -> tuple[list, list]:
        """(J per candidate, ∂J/∂ε per candidate)."""
        tape = self.forward(eps_list)
        return tape.J, self.backward(tape)
