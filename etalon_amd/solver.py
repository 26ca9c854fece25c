"""The 2D TE time loop on the HIP backend (etalon/fdtd/solver.py:1-165).

One step n (E^n, H^(n−½) → E^(n+1), H^(n+½)):

    1. H update; magnetic sources injected with M sampled at t = nΔt
    2. DFT of H with t = (n + ½)Δt
    3. E update; electric sources injected with J sampled at t = (n + ½)Δt
    4. DFT of E with t = (n + 1)Δt

``step()`` performs exactly these launches through the backend protocol (the same
call sequence as the reference solver).  ``advance(n)`` and ``run()`` replay the
same arithmetic from a fused plan: one kernel per step plus a DFT kernel per
trace chunk, bit-identical to n calls of ``step()``.  Host setup is float64 and
SI; arrays are cast once to the run's precision.
"""

from __future__ import annotations

import functools
import operator
import time
from typing import Any, Optional

import numpy as np

from etalon_amd import __version__
from etalon_amd.backends import get_backend
from etalon_amd.capabilities import check_supported
from etalon_amd.cpml import cpml_axis
from etalon_amd.grid import build_yee_grid, permittivity_on_nodes
from etalon_amd.hip import backend as _hip_backend
from etalon_amd.result import SimulationResult, monitor_data
from etalon_amd.sampling import dft_phasors, monitor_samplings
from etalon_amd.sources import point_injection
from etalon_amd.state import FIELDS, build_te_state

_state_fields = operator.attrgetter(*FIELDS)  # a state's seven arrays, FIELDS order

# Steps per fused-plan table refill (source deltas and phasors): the first refill of an
# advance covers _FIRST_TABLE_ROWS steps and each next one 4x more, up to
# _PLAN_TABLE_ROWS.  The plan's launches are asynchronous, so the host builds the next
# tables while the GPU runs the steps of the current ones, and only the first (small)
# table is built with the GPU idle.  Both are multiples of every temporal depth (1, 2,
# 4, 6) and trace chunk (64, 96), so a refill splits no step block.
_FIRST_TABLE_ROWS = 192
_PLAN_TABLE_ROWS = 24576


@functools.lru_cache(maxsize=16)
def _phasor_table(n0: int, rows: int, dt: float, magnetic: bool, freqs: bytes,
                  dtype: str) -> np.ndarray:
    """Phasors of steps [n0, n0 + rows) at H times (n+½)Δt or E times (n+1)Δt, cast once.
    The same for every simulation of one Δt and frequency list (a batch of one grid
    step shares it); read-only."""
    steps = np.arange(n0, n0 + rows, dtype=np.float64)
    t = (steps + 0.5) * dt if magnetic else (steps + 1.0) * dt
    table = dft_phasors(np.frombuffer(freqs, np.float64), t, dt).astype(np.dtype(dtype))
    table.setflags(write=False)
    return table


class MonitorDFT:
    """One monitor component's running DFT on the device: the sampled nodes' flat
    indices, the complex accumulator (n_f, nodes), and, for the per-step path, the
    device copy of the phasor rows of the current block of steps."""

    def __init__(self, sampling, backend, complex_dtype) -> None:
        self.sampling = sampling
        self.index = backend.to_device(sampling.flat_index)
        self.acc = backend.to_device(
            np.zeros((sampling.wavelengths.size, sampling.flat_index.size), complex_dtype))
        self.freqs = np.asarray(sampling.frequencies, np.float64).tobytes()
        self.block: tuple = ()       # (first step, device rows) held for step()
        self.block_rows: Any = None


class FDTDSolver:
    """Builds the state of a simulation and advances it; ``run()`` does everything."""

    def __init__(self, sim, backend: Any = None, tune=None) -> None:
        check_supported(sim)
        self.sim = sim
        self.grid = build_yee_grid(sim)
        self.n_steps = self.grid.n_steps(sim.run_time)
        self.backend = backend if backend is not None else get_backend(sim.backend,
                                                                       sim.precision)
        self.real = np.float32 if sim.precision == "single" else np.float64
        self.complex = np.complex64 if sim.precision == "single" else np.complex128
        grid = self.grid
        self.eps_r = permittivity_on_nodes(sim, grid)
        background = sim.medium.permittivity
        cpml_x = cpml_axis(grid.nx, grid.dx, grid.dt, sim.boundary, background)
        cpml_y = cpml_axis(grid.ny, grid.dy, grid.dt, sim.boundary, background)
        self.state = build_te_state(grid, self.eps_r, cpml_x, cpml_y, self.real,
                                    self.backend.to_device)
        self.injections = [point_injection(s, grid, self.eps_r) for s in sim.sources]
        self._injection_index = [(self.backend.to_device(inj.i), self.backend.to_device(inj.j))
                                 for inj in self.injections]
        self.monitors = [MonitorDFT(sampling, self.backend, self.complex)
                         for monitor in sim.monitors
                         for sampling in monitor_samplings(monitor, grid)]
        self.n = 0
        self.tune = tune
        self._plan = None
        self._plan_ids: list[int] = []
        self._plan_marks: tuple = ()
        self._plan_tables = (0, 0)
        self._plan_epoch = -1  # host-write epoch at which the plan's buffers were checked

    # ---------------------------------------------------------------- one step
    def _inject(self, magnetic: bool) -> None:
        for inj, (i, j) in zip(self.injections, self._injection_index):
            if inj.is_magnetic != magnetic:
                continue
            delta = self.backend.to_device(inj.delta(self.n, self.grid.dt).astype(self.real))
            field = getattr(self.state, inj.component.lower())
            field[i, j] -= delta

    def _step_phasors(self, mon: MonitorDFT):
        """Phasor row of step n for step(): H samples at (n+½)Δt, E samples at (n+1)Δt —
        the fused path's cached tables (_phasor_table), one device copy per block of
        _FIRST_TABLE_ROWS steps, so both paths multiply by the same bits."""
        n0 = self.n - self.n % _FIRST_TABLE_ROWS
        if mon.block != (n0,):
            table = _phasor_table(n0, _FIRST_TABLE_ROWS, self.grid.dt, mon.sampling.is_magnetic,
                                  mon.freqs, np.dtype(self.complex).str)
            mon.block_rows = self.backend.to_device(table)
            mon.block = (n0,)
        return mon.block_rows[self.n - n0]

    def _accumulate(self, magnetic: bool) -> None:
        for mon in self.monitors:
            if mon.sampling.is_magnetic != magnetic:
                continue
            field = getattr(self.state, mon.sampling.component.lower())
            self.backend.accumulate_dft(mon.acc, field, mon.index, self._step_phasors(mon))

    def step(self) -> None:
        """Advance one time step with the per-half-step primitives."""
        self.backend.update_h_te(self.state)
        self._inject(magnetic=True)
        self._accumulate(magnetic=True)
        self.backend.update_e_te(self.state)
        self._inject(magnetic=False)
        self._accumulate(magnetic=False)
        self.n += 1

    # ---------------------------------------------------------------- fused path
    def _ensure_plan(self):
        from etalon_amd.hip.plan import FusedPlan

        plan = self._plan
        if (plan is not None
                and all(getattr(self.state, name) is plan.pairs[name][plan.parity]
                        for name in FIELDS)
                and self._pair_marks(plan) == self._plan_marks):
            return plan
        # New state arrays, or values written into them from the host since the plan
        # checked them: a new plan checks again that psi is +0 outside the CPML strips
        # and that no field holds -0 (else ETL_ESTATE, and advance() steps instead).
        self._plan = None
        if plan is not None:
            plan.close()
        plan = FusedPlan(self.state, self.tune)
        self._plan_ids = self.register(plan, 0)
        self._plan = plan
        self._plan_marks = self._pair_marks(plan)
        self._plan_tables = (0, 0)  # steps [a, b) the plan's tables cover
        return plan

    @staticmethod
    def _pair_marks(plan) -> tuple:
        """Host-write counts of both buffers of every field pair of ``plan`` (see
        DeviceArray.writes): the state lives in one of each pair, which one depends on
        the steps run, and a host write into it bumps its count."""
        return tuple(plan.pairs[name][b].writes for name in FIELDS for b in (0, 1))

    def register(self, plan, sim: int) -> list[int]:
        """Give simulation ``sim`` of a fused plan this solver's sources and monitors;
        returns the plan's monitor ids, in the order of ``self.monitors``."""
        comp, ii, jj = [], [], []
        code = {"Ex": 0, "Ey": 1, "Hz": 2}
        for inj in self.injections:
            comp += [code[inj.component]] * inj.i.size
            ii += list(inj.i)
            jj += list(inj.j)
        plan.set_sources(np.array(comp, np.int32), np.array(ii, np.int64),
                         np.array(jj, np.int64), sim=sim)
        ids = [plan.add_monitor(m.sampling.component, m.sampling.flat_index, m.acc, sim=sim)
               for m in self.monitors]
        for k, owner in enumerate(self._phasor_owner()):
            if owner is not None:
                plan.share_phasors(ids[k], ids[owner], sim=sim)
        return ids

    def _phasor_owner(self) -> list:
        """Per monitor: None if it owns its phasor table, else the index of the first
        monitor with the same table (same frequencies, and both H or both E samples:
        dft_phasors of the same times — equal bits)."""
        first: dict = {}
        owners = []
        for k, mon in enumerate(self.monitors):
            key = (mon.sampling.is_magnetic, mon.sampling.frequencies.tobytes())
            owners.append(first.get(key))
            first.setdefault(key, k)
        return owners

    def reset(self) -> None:
        """Back to step 0: fields, ψ and monitor accumulators zeroed (stream-ordered)."""
        for name in FIELDS:
            getattr(self.state, name).fill_zero()
        for mon in self.monitors:
            mon.acc.fill_zero()
        self.n = 0

    def _tables(self, n0: int, rows: int):
        dt = self.grid.dt
        delta = None
        if self.injections:
            delta = np.concatenate([inj.delta_table(n0, rows, dt) for inj in self.injections],
                                   axis=1).astype(self.real)
        phasors = []
        for mon, owner in zip(self.monitors, self._phasor_owner()):
            if owner is not None:  # the plan reads the owner's table
                phasors.append(None)
                continue
            smp = mon.sampling
            phasors.append(_phasor_table(n0, rows, dt, smp.is_magnetic, mon.freqs,
                                         np.dtype(self.complex).str))
        return delta, phasors

    def fused_supported(self) -> bool:
        """The fused plan takes at most 16 source nodes; larger source sets step."""
        from etalon_amd.hip.plan import MAX_SOURCE_NODES

        return sum(inj.i.size for inj in self.injections) <= MAX_SOURCE_NODES

    def advance(self, count: int) -> None:
        """Advance ``count`` steps with the fused kernel (same results as ``step()``).

        A simulation the plan cannot take (more than 16 source nodes, or ψ set by
        hand outside the CPML strips) is advanced with ``step()`` instead — still on
        the HIP kernels, just one launch per half-step.
        """
        if count <= 0:
            return
        plan = self._plan
        if plan is not None:
            # The common case of a step loop: the same plan, no host write into any
            # device buffer since it was checked, its tables covering the steps, the
            # state in its buffers — one library call.
            n, parity = self.n, plan.parity
            lo, hi = self._plan_tables
            if (self._plan_epoch == _hip_backend.host_write_epoch() and lo <= n
                    and n + count <= hi
                    and all(map(operator.is_, _state_fields(self.state), plan.arrays[parity]))):
                if plan.advance(n, count) != parity:
                    plan.bind()
                self.n = n + count
                return
        if not self.fused_supported():
            for _ in range(count):
                self.step()
            return
        from etalon_amd.hip._lib import HipError

        try:
            plan = self._ensure_plan()
        except HipError as error:
            if error.code != -3:  # ETL_ESTATE: psi not clean outside the strips
                raise
            for _ in range(count):
                self.step()
            return
        self._plan_epoch = _hip_backend.host_write_epoch()
        start, done, refills = self.n, 0, 0
        grow = _FIRST_TABLE_ROWS
        while done < count:
            have = self._plan_tables
            if have[0] <= self.n < have[1]:
                # the tables depend on the step numbers only: a run from step 0 again
                # (after reset) reuses what the plan holds
                rows = min(count - done, have[1] - self.n)
            else:
                # at least _FIRST_TABLE_ROWS rows: short advances that follow (a step
                # loop of advance(1), a timed region after a warm-up) then find their
                # tables on the device and upload nothing
                staged = min(grow, max(count - done, _FIRST_TABLE_ROWS))
                rows = min(count - done, staged)
                grow = min(4 * grow, _PLAN_TABLE_ROWS)
                self._stage_tables(plan, self.n, staged)
                refills += 1
            plan.advance(self.n, rows)
            plan.bind()
            self.n += rows
            done += rows
        if refills > 1 and count <= _PLAN_TABLE_ROWS:
            # Built while the GPU runs the last piece and uploaded before the plan's next
            # steps: a run over the same steps again (a design loop's next forward, after
            # reset) then finds all its tables in one piece.
            self._stage_tables(plan, start, count)

    def _stage_tables(self, plan, n0: int, rows: int) -> None:
        """Hand the plan the source and phasor tables of steps [n0, n0 + rows); it
        uploads them before its next steps."""
        delta, phasors = self._tables(n0, rows)
        if delta is not None:
            plan.source_table(n0, delta)
        for mon_id, table in zip(self._plan_ids, phasors):
            if table is not None:
                plan.phasor_table(mon_id, n0, table)
        self._plan_tables = (n0, n0 + rows)

    # ---------------------------------------------------------------- results
    def fields(self) -> dict[str, np.ndarray]:
        """Host copies of Ex, Ey, Hz (E at nΔt, H at (n−½)Δt)."""
        s = self.state
        return {name: self.backend.to_host(getattr(s, name.lower())) for name in ("Ex", "Ey", "Hz")}

    def result(self, wall_time: float = float("nan")) -> SimulationResult:
        data: dict[str, dict[str, Any]] = {m.name: {} for m in self.sim.monitors}
        for mon in self.monitors:
            acc = self.backend.to_host(mon.acc)
            data[mon.sampling.monitor][mon.sampling.component] = monitor_data(
                mon.sampling, acc, self.grid.z)
        metadata = {
            "etalon_version": __version__,
            "backend": self.backend.name,
            "device": self.backend.device,
            "precision": self.sim.precision,
            "n_steps": self.n,
            "dt_s": self.grid.dt,
            "grid_cells": (self.grid.nx, self.grid.ny),
            "wall_time_s": wall_time,
            "cell_updates_per_s": self.grid.n_cells * self.n / wall_time if wall_time > 0 else None,
        }
        return SimulationResult(self.sim, self.grid, data, metadata)

    def run(self, fused: Optional[bool] = None) -> SimulationResult:
        """Run to ``run_time``; the fused path unless ``fused=False``."""
        start = time.perf_counter()
        if fused is None or fused:
            self.advance(self.n_steps - self.n)
        else:
            while self.n < self.n_steps:
                self.step()
        self.backend.synchronize()
        return self.result(time.perf_counter() - start)


def run(sim) -> SimulationResult:
    """Run a simulation to ``run_time`` and collect its monitors."""
    return FDTDSolver(sim).run()
