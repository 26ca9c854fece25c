"""Permittivity gradients by the discrete adjoint, on the HIP path.

For a simulation whose last structure is a ``PixelGrid`` design region and an
objective J = Σ over named monitors of Σ |values|² (every component they record,
values as the result reports them), ``AdjointProblem`` computes J and ∂J/∂ε for
every design pixel:

1. forward: the fused plan advances the simulation in segments of K ≈ √N steps
   and saves the full state at every segment start (checkpoints), then J comes
   from the monitors;
2. backward: for each segment from the last, the state is restored from its
   checkpoint and the segment is recomputed forward with a probe on Hz around
   the design box (time reversal through the CPML is unstable, so the forward
   fields are recomputed, never run backwards), then the adjoint kernels run the
   segment's reverse steps, accumulating ∂J/∂cb on the box; with the tile
   kernels the recompute of segment k − 1 runs beside the reverse steps of
   segment k, in the same launches (``schedule="joint"``, a co-plan) or on a
   second stream (``"lanes"``), the two alternating between two history halves;
3. ∂J/∂εr = −(cb/εr) ∂J/∂cb, summed over the Ex/Ey nodes each pixel paints.

``mode="store"`` keeps the probe history of the whole forward run instead (no
recompute) when its memory is affordable.  ``PermittivityGradient`` exposes it
as a ``torch.autograd.Function``.  The reference has no adjoint; the float64
NumPy adjoint in the tests and central differences pin this one.
"""

from __future__ import annotations

import ctypes as C
import functools
import math
import time
from dataclasses import dataclass, field, replace
from typing import Optional, Sequence

import numpy as np

from etalon_amd import models as et
from etalon_amd.grid import build_yee_grid, pixel_node_map
from etalon_amd.hip import _lib
from etalon_amd.hip.backend import zeros
from etalon_amd.solver import FDTDSolver, _phasor_table
from etalon_amd.state import FIELDS, PROFILES, TEState, cb_arrays

_COMP = {"Ex": 0, "Ey": 1, "Hz": 2}


def checkpoint_interval(n_steps: int) -> int:
    """K ≈ √N, a multiple of 64 (the plan's graph chunk) and at least 64."""
    return max(64, 64 * round(math.sqrt(n_steps) / 64))


@functools.lru_cache(maxsize=8)
def _forward_phasors(n_steps: int, dt: float, magnetic: bool, freqs: bytes,
                     dtype: str) -> np.ndarray:
    """[Re P | Im P] (n_steps, 2 n_f) float64 of the forward pass's phasor table P of
    steps [0, n_steps), in its precision; shared by the candidates of one Δt and
    frequency list; read-only."""
    P = _phasor_table(0, n_steps, dt, magnetic, freqs, dtype).astype(np.complex128)
    Q = np.ascontiguousarray(np.concatenate([P.real, P.imag], axis=1))
    Q.setflags(write=False)
    return Q


def design_tune(tune):
    """The tune of the plans of a design loop, whose cb arrays every gradient rewrites:
    ``tune`` with uniform-material tiles off (each rewrite would lay the tiles out again)."""
    from etalon_amd.hip.plan import Tune

    return replace(tune if tune is not None else Tune(), uniform_cb=-1)


def recompute_tune(tune, precision: str):
    """The tune of an adjoint problem's recompute and reverse plans: the forward tune,
    in double with an unset temporal depth 4 and unset tile rows 8 (large double forward
    plans default to depth 6 and 16-row tiles; the double reverse kernels, and the
    double kernels pairing a recompute with reverse steps, are compiled up to depth 4
    with 8-row forward tiles)."""
    from etalon_amd.hip.plan import Tune

    if precision != "double":
        return tune
    t = tune if tune is not None else Tune()
    return replace(t, temporal=t.temporal or 4, band_rows=t.band_rows or 8)


def _weights(sampling) -> np.ndarray:
    """Interpolation matrix W (n_values, n_nodes) of a monitor component's labelled
    values (monitor_data: size-0 axes interpolated in float64)."""
    nxs, nys = sampling.x.index.size, sampling.y.index.size
    wx = np.eye(nxs) if sampling.x.weights is None else np.asarray(sampling.x.weights)[None, :]
    wy = np.eye(nys) if sampling.y.weights is None else np.asarray(sampling.y.weights)[None, :]
    return np.kron(wx, wy)


@dataclass
class Tape:
    """What the backward pass needs from the forward pass."""

    solver: FDTDSolver
    J: float
    checkpoints: dict = field(default_factory=dict)  # field name -> (n_ckpt, size) DeviceArray
    starts: list = field(default_factory=list)       # segment starts
    history: Optional[object] = None                 # store mode: (N, box) probe output
    forward_s: float = 0.0
    eps: Optional[np.ndarray] = None                 # the design permittivity it ran
    generation: int = 0                              # the session forward that made it


@dataclass
class _Window:
    """The region a recompute needs: Hz nodes [i0, i0 + nx) x [j0, j0 + ny) of the grid
    (Ex, Ey of the same rows and columns, one more column / row), the design box's
    nodes and every node within K + 2 of them.  Over a segment of K steps from a
    checkpoint nothing outside it reaches the box (one step moves an influence one
    node), so the box's history recomputed on the window alone — its edges held as
    the window grid's boundary — is the full grid's, bit for bit.  ``state`` is the
    window's TEState (fields restored from the full checkpoints, cb and CPML profiles
    the grid's there)."""

    i0: int
    j0: int
    nx: int
    ny: int
    state: object


@dataclass
class _Session:
    """What a design loop keeps between gradients of one geometry.  Every forward on
    it overwrites the solver's state, monitors and checkpoints; ``generation`` counts
    those forwards, so a tape of an earlier one is known to be stale."""

    solver: FDTDSolver
    checkpoints: dict
    starts: list
    generation: int = 1
    bufs: Optional["_AdjointBuffers"] = None
    rev: Optional["_Reverse"] = None
    rplan: Optional[object] = None
    window: Optional[_Window] = None
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


@dataclass
class _AdjointBuffers:
    """λ, ∂J/∂cb and history arrays of one simulation's reverse run."""

    lam: dict
    g_ex: object
    g_ey: object
    hist: object
    desc: object


class AdjointProblem:
    def __init__(self, sim, objective: Sequence[str], checkpoint_every: Optional[int] = None,
                 mode: str = "checkpoint", tune=None, kernel: str = "tile",
                 schedule: str = "joint", reuse: bool = True, window: bool = True):
        if not sim.structures or sim.structures[-1].type != "PixelGrid":
            raise ValueError("the design region must be the simulation's last structure, a "
                             "PixelGrid")
        names = {m.name for m in sim.monitors}
        missing = set(objective) - names
        if missing:
            raise ValueError(f"objective monitors not in the simulation: {sorted(missing)}")
        if mode not in ("checkpoint", "store"):
            raise ValueError("mode is 'checkpoint' or 'store'")
        if kernel not in ("tile", "plain"):
            raise ValueError("kernel is 'tile' (temporally blocked) or 'plain' (one launch per "
                             "reverse half-step)")
        if schedule not in ("joint", "lanes", "sequential"):
            raise ValueError("schedule is 'joint', 'lanes' or 'sequential'")
        self.kernel = kernel
        # checkpoint mode: recompute each segment on the window of nodes the design box
        # depends on over it (False: on the whole grid; the same bits)
        self.window = bool(window)
        # checkpoint mode: how the recompute of segment k - 1 and the reverse steps of
        # segment k share the GPU (the plain kernels run them one after the other)
        self.schedule = schedule if kernel == "tile" else "sequential"
        # a design loop's gradients (checkpoint mode, tile kernels, joint schedule) keep
        # the solver, plans, graphs and buffers of the first one: a new permittivity
        # only replaces cb_ex and cb_ey.  A tape is then valid until the next forward.
        self.reuse = bool(reuse) and mode == "checkpoint" and self.schedule == "joint"
        self._session: Optional[_Session] = None
        self.sim = sim
        self.objective = tuple(objective)
        self.mode = mode
        # the plans' tune (forward, recompute and reverse).  By default every float
        # plan over this grid picks the same temporal depth, so a co-plan pairs its step
        # blocks; double reverse kernels run at depth 4 at most, and the recompute
        # follows them.
        self.tune = design_tune(tune)
        self.rtune = recompute_tune(self.tune, sim.precision)
        grid = build_yee_grid(sim)
        self.grid = grid
        self.n_steps = grid.n_steps(sim.run_time)
        self.K = checkpoint_every or checkpoint_interval(self.n_steps)
        design = sim.structures[-1]
        self.shape = np.asarray(design.permittivity).shape
        self.pixmap = pixel_node_map(grid, design.center, design.size, self.shape)
        ii = np.concatenate([self.pixmap[c][0] for c in ("Ex", "Ey")])
        jj = np.concatenate([self.pixmap[c][1] for c in ("Ex", "Ey")])
        if ii.size == 0:
            raise ValueError("the design region covers no E node")
        self.box = (int(ii.min()), int(jj.min()), int(ii.max() - ii.min() + 1),
                    int(jj.max() - jj.min() + 1))
        self._box_checked = False
        for src in sim.sources:
            if src.component != "Hz":
                raise ValueError("adjoint problems take Hz sources (an E source's strength "
                                 "depends on the permittivity at its nodes)")

    # ------------------------------------------------------------------ helpers
    def with_permittivity(self, eps: np.ndarray):
        eps = np.asarray(eps, dtype=np.float64)
        if eps.shape != self.shape:
            raise ValueError(f"design permittivity must be {self.shape}, got {eps.shape}")
        design = self.sim.structures[-1].model_copy(update={"permittivity": eps})
        return self.sim.model_copy(update={"structures": (*self.sim.structures[:-1], design)})

    def _check_box_off_cpml(self, solver) -> None:
        if self._box_checked:  # the CPML profiles depend on the grid only
            return
        i0, j0, ni, nj = self.box
        prof = self._profiles(solver)
        for name, lo, hi in (("bx_e", i0 - 1, i0 + ni + 1), ("bx_h", i0 - 1, i0 + ni + 1),
                             ("by_e", j0 - 1, j0 + nj + 1), ("by_h", j0 - 1, j0 + nj + 1)):
            b = prof[name][lo:hi]
            c = prof[name.replace("b", "c", 1)][lo:hi]
            ik = prof[name.replace("b", "ik", 1)][lo:hi]
            if np.any(b != 1) or np.any(c != 0) or np.any(ik != 1):
                raise ValueError("the design region must lie outside the CPML")
        self._box_checked = True

    @staticmethod
    def _profiles(solver) -> dict:
        """The CPML profiles of ``solver``'s state on the host (PROFILES names)."""
        st = solver.state
        if st.host_profiles is not None:
            return st.host_profiles
        return {name: getattr(st, name).get() for name in PROFILES}

    def _box_hz_nodes(self) -> np.ndarray:
        i0, j0, ni, nj = self.box
        ny = self.grid.ny
        ii, jj = np.meshgrid(np.arange(i0 - 1, i0 + ni), np.arange(j0 - 1, j0 + nj), indexing="ij")
        return (ii * ny + jj).ravel().astype(np.int64)

    # ------------------------------------------------------------------ forward
    def forward(self, eps: np.ndarray) -> Tape:
        """Run the simulation with design permittivity ``eps``; J and the tape."""
        if self.reuse and self._session is not None:
            return self._forward_again(eps)
        return self._forward_new(eps, session=self.reuse)

    def _forward_new(self, eps: np.ndarray, session: bool) -> Tape:
        """The forward pass on a new solver; ``session``: keep it for a design loop."""
        sim = self.with_permittivity(eps)
        solver = FDTDSolver(sim, tune=self.tune)
        self._check_box_off_cpml(solver)
        backend = solver.backend
        N, K = self.n_steps, self.K
        t0 = time.perf_counter()
        plan = solver._ensure_plan()
        delta, phasors = solver._tables(0, N)
        if delta is not None:
            plan.source_table(0, delta)
        for mon_id, table in zip(solver._plan_ids, phasors):
            if table is not None:
                plan.phasor_table(mon_id, 0, table)
        solver._plan_tables = (0, N)
        tape = Tape(solver=solver, J=0.0, eps=np.array(eps, dtype=np.float64))
        if self.mode == "store":
            tape.history = zeros((N, self._box_hz_nodes().size), solver.real)
            plan.add_probe("Hz", self._box_hz_nodes(), tape.history)
            plan.advance(0, N)
            plan.bind()
        else:
            tape.starts = list(range(0, N, K))
            n_ck = len(tape.starts)
            tape.checkpoints = self._checkpoint_storage(solver, n_ck)
            self._set_forward_checkpoints(plan, tape.checkpoints, n_ck, 0)
            plan.advance_checkpointed(0, N, K)  # slot k = the state at step k K
            plan.bind()
            if session:
                self._session = _Session(solver=solver, checkpoints=tape.checkpoints,
                                         starts=tape.starts)
                tape.generation = self._session.generation
        solver.n = N
        backend.synchronize()
        tape.forward_s = time.perf_counter() - t0
        tape.J = self._objective(solver)
        return tape

    def _forward_again(self, eps: np.ndarray) -> Tape:
        """The forward pass on the session's solver: the design's nodes take the new
        permittivity (the PixelGrid rule, the last structure wins there), cb_ex and
        cb_ey are recomputed and uploaded, fields and monitors go back to zero, and the
        same plan, tables and checkpoint storage run again."""
        ses = self._session
        solver = ses.solver
        t0 = time.perf_counter()
        self.set_design(solver, eps)
        ses.generation += 1
        plan = solver._ensure_plan()
        plan.advance_checkpointed(0, self.n_steps, self.K)
        plan.bind()
        solver.n = self.n_steps
        solver.backend.synchronize()
        tape = Tape(solver=solver, J=0.0, checkpoints=ses.checkpoints, starts=ses.starts,
                    eps=np.array(eps, dtype=np.float64), generation=ses.generation)
        tape.forward_s = time.perf_counter() - t0
        tape.J = self._objective(solver)
        return tape

    def set_design(self, solver, eps: np.ndarray) -> None:
        """Give ``solver`` (a solver of this problem's geometry) the design permittivity
        ``eps`` and take it back to step 0: the design's nodes take the new values (the
        PixelGrid rule, the last structure wins there), cb_ex and cb_ey are recomputed
        from the whole permittivity map and uploaded in place (their device addresses,
        which plans hold, stay), and fields and monitors go back to zero."""
        eps = np.asarray(eps, dtype=np.float64)
        if eps.shape != self.shape:
            raise ValueError(f"design permittivity must be {self.shape}, got {eps.shape}")
        flat = eps.ravel()
        for c in ("Ex", "Ey"):
            ii, jj, p = self.pixmap[c]
            solver.eps_r[c][ii, jj] = flat[p]
        # cb is elementwise in εr, and εr changed only on the design box's rows
        i0, _, ni, _ = self.box
        rows = {c: solver.eps_r[c][i0:i0 + ni] for c in ("Ex", "Ey")}
        cb = cb_arrays(solver.grid, rows)
        solver.state.cb_ex.rows(i0, i0 + ni).set(cb["cb_ex"].astype(solver.real))
        solver.state.cb_ey.rows(i0, i0 + ni).set(cb["cb_ey"].astype(solver.real))
        solver.sim = self.with_permittivity(eps)
        solver.reset()

    def close(self) -> None:
        """Release the session (plans, graphs, buffers) of a design loop."""
        if self._session is not None:
            self._session.close()
            self._session = None

    def _objective(self, solver) -> float:
        J = 0.0
        for mon in solver.monitors:
            if mon.sampling.monitor in self.objective:
                acc = mon.acc.get().astype(np.complex128)
                G = acc @ _weights(mon.sampling).T
                J += float(np.sum(np.abs(G) ** 2))
        return J

    def _sources(self, solver):
        """Distinct adjoint-source nodes (comp, flat) and the table (N, n) of dJ/dv.

        Column k of a monitor's block is S_k = 2 Re(P conj(A_k)) over the forward
        pass's phasors P (N, n_f) and A = (acc Wᵀ) W, computed as one real product
        [Re P | Im P] [Re A; Im A]; columns of one node sum in monitor order."""
        N, dt = self.n_steps, self.grid.dt
        blocks = []
        for mon in solver.monitors:
            smp = mon.sampling
            if smp.monitor not in self.objective:
                continue
            W = _weights(smp)
            acc = mon.acc.get().astype(np.complex128)        # (n_f, n_nodes)
            A = (acc @ W.T) @ W                              # (n_f, n_nodes)
            Q = _forward_phasors(N, dt, smp.is_magnetic,
                                 np.asarray(smp.frequencies, np.float64).tobytes(),
                                 np.dtype(solver.complex).str)
            keys = [(_COMP[smp.component], int(f)) for f in smp.flat_index]
            blocks.append((Q, A, keys))
        if not blocks:
            return np.zeros(0, np.int32), np.zeros(0, np.int64), np.zeros((N, 0), solver.real)

        def columns(Q, A):
            S = Q @ np.concatenate([A.real, A.imag], axis=0)  # (N, n_nodes)
            S *= 2.0
            return S

        keys = blocks[0][2]
        if len(blocks) == 1 and len(set(keys)) == len(keys):
            # one monitor, distinct nodes: its columns in node order (0.0 + S as the
            # sums below start from)
            order = sorted(range(len(keys)), key=keys.__getitem__)
            Q, A, _ = blocks[0]
            S = columns(Q, A[:, order])
            S += 0.0
            keys = [keys[k] for k in order]
            table = S.astype(solver.real)
        else:
            cols: dict = {}
            for Q, A, bkeys in blocks:
                S = columns(Q, A)
                for k, key in enumerate(bkeys):
                    cols[key] = cols.get(key, 0.0) + S[:, k]
            keys = sorted(cols)
            table = np.ascontiguousarray(np.stack([cols[k] for k in keys], axis=1),
                                         dtype=solver.real)
        comp = np.array([k[0] for k in keys], np.int32)
        flat = np.array([k[1] for k in keys], np.int64)
        return comp, flat, table

    # ------------------------------------------------------------------ backward
    def backward(self, tape: Tape) -> np.ndarray:
        """∂J/∂ε per design pixel (float64), from the forward tape.

        A tape of the design-loop session whose state a later forward has overwritten
        (J1 = objective(a); J2 = objective(b); (J1 - J2).backward(), a line search
        between forward and backward) is replayed: its forward runs again at its own
        permittivity on a new solver, and the backward pass runs on that."""
        ses = self._session
        if self.reuse and ses is not None and tape.solver is ses.solver:
            if tape.generation != ses.generation:
                return self.backward(self._forward_new(tape.eps, session=False))
            return self._backward_session(tape)
        solver = tape.solver
        N, K = self.n_steps, self.K
        if self.mode == "store":
            hist = tape.history
        else:
            # two halves when recompute and reverse steps overlap
            rows = 2 * K if self.schedule != "sequential" else K
            hist = zeros((rows, self._box_hz_nodes().size), solver.real)
        bufs = self._adjoint_buffers(solver, hist)
        rev = _Reverse(self.kernel, bufs.desc, self.rtune)
        try:
            comp, flat, table = self._sources(solver)
            rev.sources(comp, flat, table)
            if self.mode == "store":
                rev.run(0, N)
            else:
                self._reverse_segments(tape, rev, hist)
            solver.backend.synchronize()
        finally:
            rev.close()
        return self._pixel_gradient(solver, bufs)

    def _backward_session(self, tape: Tape) -> np.ndarray:
        """The joint-schedule backward pass on the session's plans: built on the first
        gradient, reset (λ and ∂J/∂cb to zero, the new adjoint-source table in place)
        on the next ones."""
        from etalon_amd.hip.plan import FusedPlan

        ses = self._session
        solver = tape.solver
        N, K = self.n_steps, self.K
        comp, flat, table = self._sources(solver)
        if ses.rev is None:
            hist = zeros((2 * K, self._box_hz_nodes().size), solver.real)
            ses.bufs = self._adjoint_buffers(solver, hist)
            ses.rev = _Reverse("tile", ses.bufs.desc, self.rtune)
            ses.rev.sources(comp, flat, table)
            state, tune, ses.window = self._recompute_state(solver, ses.rev)
            ses.rplan = FusedPlan(state, tune)
            self._configure_recompute(ses.rplan, solver, 0, hist, ses.window)
            self._set_recompute_checkpoints(ses.rplan, ses.checkpoints, len(ses.starts), 0,
                                            ses.window)
            _lib.call("etl_coplan_create", ses.rplan.handle, ses.rev.handle, C.byref(ses.co))
        else:
            _lib.call("etl_adjplan_reset", ses.rev.handle)
            ses.bufs.g_ex.fill_zero()
            ses.bufs.g_ey.fill_zero()
            ses.rev.sources(comp, flat, table)
            if ses.window is not None:
                self._refresh_window(solver, ses.window)
        _lib.call("etl_coplan_sweep", ses.co, N, K)
        solver.backend.synchronize()
        return self._pixel_gradient(solver, ses.bufs)

    # ------------------------------------------------------------------ pieces
    def _adjoint_buffers(self, solver, hist) -> "_AdjointBuffers":
        """Zeroed λ and ∂J/∂cb arrays of one simulation and its adjoint descriptor."""
        state, real = solver.state, solver.real
        i0, j0, ni, nj = self.box
        b = _AdjointBuffers(lam={name: zeros(getattr(state, name).shape, real) for name in FIELDS},
                            g_ex=zeros((ni, nj), np.float64), g_ey=zeros((ni, nj), np.float64),
                            hist=hist, desc=_lib.EtlAdjointDesc())
        d = b.desc
        d.precision = 0 if real == np.float32 else 1
        d.nx, d.ny = int(state.nx), int(state.ny)
        d.db, d.inv_dx, d.inv_dy = float(state.db), float(state.inv_dx), float(state.inv_dy)
        for k, name in enumerate(FIELDS):
            d.lam[k] = b.lam[name].ptr
        d.cb_ex, d.cb_ey = state.cb_ex.ptr, state.cb_ey.ptr
        for name in PROFILES:
            setattr(d, name, getattr(state, name).ptr)
        d.box_i0, d.box_j0, d.box_ni, d.box_nj = i0, j0, ni, nj
        d.grad_ex, d.grad_ey = b.g_ex.ptr, b.g_ey.ptr
        d.hist, d.hist_rows = hist.ptr, hist.shape[0]
        return b

    def _pixel_gradient(self, solver, bufs) -> np.ndarray:
        """∂J/∂εr = −(cb/εr) ∂J/∂cb at the nodes, summed per pixel."""
        i0, j0, ni, _ = self.box
        state = solver.state
        eps = solver.eps_r
        # the design box's rows of cb (every design node lies in them)
        cb = {"Ex": state.cb_ex.rows(i0, i0 + ni).get(), "Ey": state.cb_ey.rows(i0, i0 + ni).get()}
        gcb = {"Ex": bufs.g_ex.get(), "Ey": bufs.g_ey.get()}
        out = np.zeros(self.shape[0] * self.shape[1])
        for c in ("Ex", "Ey"):
            ii, jj, p = self.pixmap[c]
            g = -(cb[c][ii - i0, jj].astype(np.float64) / eps[c][ii, jj]) * gcb[c][ii - i0, jj - j0]
            np.add.at(out, p, g)
        return out.reshape(self.shape)

    def _window_bounds(self) -> tuple:
        """(i0, j0, nx, ny) of the recompute window (see _Window): the design box's Hz
        nodes, rows i0 - 1 .. i0 + ni - 1 and columns j0 - 1 .. j0 + nj - 1, and every
        node within K + 2 of them, cut to the grid."""
        i0, j0, ni, nj = self.box
        nx, ny = self.grid.nx, self.grid.ny
        m = self.K + 2
        r0, r1 = max(0, i0 - 1 - m), min(nx, i0 + ni + m)
        c0, c1 = max(0, j0 - 1 - m), min(ny, j0 + nj + m)
        return r0, c0, r1 - r0, c1 - c0

    def _checkpoint_storage(self, solver, n_ck: int) -> dict:
        """Zeroed checkpoint arrays of the forward pass, FIELDS order: ``n_ck`` slots of
        each whole field, or with ``window`` of each field's block on the recompute
        window (all a recompute restores)."""
        real = solver.real
        if not self.window:
            return {name: zeros((n_ck, getattr(solver.state, name).size), real)
                    for name in FIELDS}
        _, _, wx, wy = self._window_bounds()
        size = {"ex": wx * (wy + 1), "ey": (wx + 1) * wy, "hz": wx * wy, "psi_ex_y": wx * (wy + 1),
                "psi_ey_x": (wx + 1) * wy, "psi_hz_x": wx * wy, "psi_hz_y": wx * wy}
        return {name: zeros((n_ck, size[name]), real) for name in FIELDS}

    def _set_forward_checkpoints(self, plan, checkpoints: dict, n_ck: int, sim: int) -> None:
        """Simulation ``sim`` of the forward plan saves into ``checkpoints``
        (_checkpoint_storage's arrays)."""
        arrays = [checkpoints[name] for name in FIELDS]
        if not self.window:
            plan.set_checkpoints(arrays, n_ck, sim=sim)
        else:
            r0, c0, wx, wy = self._window_bounds()
            plan.set_checkpoint_window(arrays, n_ck, wx, wy, r0, c0, sim=sim)

    def _window_of(self, solver) -> _Window:
        """The recompute window of ``solver``'s grid (see _Window), with its state."""
        from etalon_amd.hip.backend import to_device

        r0, c0, wx, wy = self._window_bounds()
        r1, c1 = r0 + wx, c0 + wy
        real = solver.real
        st = solver.state
        put = lambda a: to_device(np.ascontiguousarray(a, dtype=real))  # noqa: E731
        cb = self._window_cb(solver, r0, r1, c0, c1)
        host = self._profiles(solver)
        prof = {}
        for name in PROFILES:
            v = host[name]
            if "x" in name.split("_")[0]:  # bx_*, cx_*, ikx_*: along the rows
                prof[name] = v[r0:r1 + 1] if name.endswith("_e") else v[r0:r1]
            else:
                prof[name] = v[c0:c1 + 1] if name.endswith("_e") else v[c0:c1]
        # the twelve window profiles in one device array (one upload), each a view
        packed = put(np.concatenate([prof[name] for name in PROFILES]))
        at = 0
        for name in PROFILES:
            n = prof[name].size
            prof[name] = packed.rows(at, at + n)
            at += n
        shapes = {"ex": (r1 - r0, c1 - c0 + 1), "ey": (r1 - r0 + 1, c1 - c0),
                  "hz": (r1 - r0, c1 - c0), "psi_ex_y": (r1 - r0, c1 - c0 + 1),
                  "psi_ey_x": (r1 - r0 + 1, c1 - c0), "psi_hz_x": (r1 - r0, c1 - c0),
                  "psi_hz_y": (r1 - r0, c1 - c0)}
        state = TEState(**{name: zeros(shape, real) for name, shape in shapes.items()},
                        cb_ex=put(cb["cb_ex"]), cb_ey=put(cb["cb_ey"]), **prof,
                        db=st.db, inv_dx=st.inv_dx, inv_dy=st.inv_dy, nx=r1 - r0, ny=c1 - c0)
        return _Window(r0, c0, r1 - r0, c1 - c0, state)

    @staticmethod
    def _window_cb(solver, r0, r1, c0, c1) -> dict:
        """cb_ex, cb_ey of the window rows [r0, r1) x columns [c0, c1] (Ex) and rows
        [r0, r1] x columns [c0, c1) (Ey), from the solver's permittivity: the values
        (and bits) of its full arrays there."""
        cb = cb_arrays(solver.grid, {"Ex": solver.eps_r["Ex"][r0:r1, c0:c1 + 1],
                                     "Ey": solver.eps_r["Ey"][r0:r1 + 1, c0:c1]})
        return {k: np.ascontiguousarray(v, dtype=solver.real) for k, v in cb.items()}

    def _refresh_window(self, solver, win: _Window) -> None:
        """After a new design permittivity: the window's cb from the solver's."""
        cb = self._window_cb(solver, win.i0, win.i0 + win.nx, win.j0, win.j0 + win.ny)
        win.state.cb_ex.set(cb["cb_ex"])
        win.state.cb_ey.set(cb["cb_ey"])

    def _recompute_state(self, solver, rev, rtune=False) -> tuple:
        """(state, tune, window or None) of the recompute plan beside reverse plan
        ``rev``: the window's state at the reverse plan's depth, or the full state.
        ``rtune`` (default: this problem's) is the recompute plans' tune when set."""
        tune = self.rtune if rtune is False else rtune
        if rev.kernel == "tile" and (tune is None or tune.temporal == 0):
            # step blocks that pair in a co-plan (a forward plan's own default may be
            # deeper than any reverse kernel)
            tune = replace(design_tune(tune), temporal=rev.temporal())
        if not self.window:
            return solver.state, tune, None
        win = self._window_of(solver)
        return win.state, tune, win

    def _configure_recompute(self, plan, solver, sim: int, hist, win=None) -> None:
        """Sources, source table and the Hz probe on the box of simulation ``sim`` of a
        recompute plan (no monitors: the recompute only feeds the history); on a
        window, the source nodes inside it and the checkpoints' window."""
        comp, ii, jj, keep = [], [], [], []
        col = 0
        for inj in solver.injections:
            c = _COMP[inj.component]
            for i, j in zip(inj.i, inj.j):
                i, j = int(i), int(j)
                if win is not None:
                    rows = win.nx + (1 if c == 1 else 0)
                    cols = win.ny + (1 if c == 0 else 0)
                    i, j = i - win.i0, j - win.j0
                    if not (0 <= i < rows and 0 <= j < cols):
                        col += 1
                        continue  # beyond the window: cannot reach the box in a segment
                comp.append(c)
                ii.append(i)
                jj.append(j)
                keep.append(col)
                col += 1
        plan.set_sources(np.array(comp, np.int32), np.array(ii, np.int64),
                         np.array(jj, np.int64), sim=sim)
        delta, _ = solver._tables(0, self.n_steps)
        if delta is not None and keep:
            plan.source_table(0, np.ascontiguousarray(delta[:, keep]), sim=sim)
        nodes = self._box_hz_nodes()
        if win is not None:
            i, j = np.divmod(nodes, self.grid.ny)
            nodes = ((i - win.i0) * win.ny + (j - win.j0)).astype(np.int64)
        plan.add_probe("Hz", nodes, hist, sim=sim)

    def _set_recompute_checkpoints(self, plan, checkpoints: dict, n_ck: int, sim: int,
                                   win=None) -> None:
        """Simulation ``sim`` of a recompute plan restores from the forward pass's
        checkpoints: whole fields, or on a window (whose fields are the blocks the
        forward pass saved) the window's."""
        if (win is not None) != self.window:
            raise ValueError("window checkpoints go with a window recompute, whole-field "
                             "checkpoints with a whole-grid one")
        plan.set_checkpoints([checkpoints[name] for name in FIELDS], n_ck, sim=sim)

    def _reverse_segments(self, tape: Tape, rev, hist) -> None:
        """Recompute each segment from its checkpoint (with the Hz probe on the box),
        then run its reverse steps, from the last segment to the first."""
        from etalon_amd.hip.plan import FusedPlan

        solver = tape.solver
        N = self.n_steps
        state, tune, win = self._recompute_state(solver, rev)
        plan = FusedPlan(state, tune)
        try:
            self._configure_recompute(plan, solver, 0, hist, win)
            self._set_recompute_checkpoints(plan, tape.checkpoints, len(tape.starts), 0, win)
            # Schedules: "joint" — the recompute of segment k - 1 and the reverse steps
            # of segment k in one launch per step block (a co-plan); "lanes" — the
            # same two on two streams; "sequential" — one after the other.  The first
            # two alternate between two history halves.
            schedule = self.schedule if hist.shape[0] >= 2 * self.K else "sequential"
            lane = 1 if schedule == "lanes" else 0
            if lane:
                plan.set_lane(1)

            def half(k: int) -> int:
                return (k & 1) * self.K if schedule != "sequential" else 0

            def count(k: int) -> int:
                return min(self.K, N - tape.starts[k])

            def recompute(k: int) -> None:
                plan.checkpoint(k, restore=True)  # on the plan's lane
                plan.advance(tape.starts[k], count(k), half(k))

            if schedule == "joint":
                co = C.c_void_p()
                _lib.call("etl_coplan_create", plan.handle, rev.handle, C.byref(co))
                try:
                    _lib.call("etl_coplan_sweep", co, N, self.K)
                finally:
                    _lib.call("etl_coplan_destroy", co)
                return
            last = len(tape.starts) - 1
            recompute(last)
            for k in range(last, -1, -1):
                if schedule == "lanes":
                    _lib.call("etl_lane_wait", 0, 1)      # reverse k after recompute k
                    if k > 0:
                        _lib.call("etl_lane_wait", 1, 0)  # recompute k-1 after reverse k+1
                        recompute(k - 1)
                    rev.run(tape.starts[k], count(k), half(k))
                else:
                    rev.run(tape.starts[k], count(k))
                    if k > 0:
                        recompute(k - 1)
        finally:
            plan.close()

    def run(self, eps: np.ndarray) -> tuple[float, np.ndarray]:
        """(J, ∂J/∂ε per pixel)."""
        tape = self.forward(eps)
        return tape.J, self.backward(tape)


class _Reverse:
    """The reverse steps: the temporally blocked adjoint plan ("tile") or the
    one-launch-per-half-step kernels ("plain"); the same arithmetic."""

    def __init__(self, kernel: str, desc, tune) -> None:
        """``desc``: one EtlAdjointDesc, or a list of them (tile kernels: one plan over
        every simulation)."""
        self.kernel = kernel
        self.handle = C.c_void_p()
        descs = desc if isinstance(desc, (list, tuple)) else [desc]
        if kernel == "tile":
            t = None
            if tune is not None:
                # the forward band: the plan's default depth follows the forward rule
                t = _lib.EtlTune(tune.band_rows, 1, 1, 0, tune.temporal, tune.waves,
                                 tune.launch)
            arr = (_lib.EtlAdjointDesc * len(descs))(*descs)
            _lib.call("etl_adjplan_create", arr, len(descs), C.byref(t) if t else None,
                      C.byref(self.handle))
        else:
            if len(descs) != 1:
                raise ValueError("the plain reverse kernels take one simulation")
            _lib.call("etl_adjoint_create", C.byref(descs[0]), C.byref(self.handle))

    def sources(self, comp, flat, table, sim: int = 0) -> None:
        args = (int(comp.size), comp.ctypes.data_as(C.POINTER(C.c_int)),
                flat.ctypes.data_as(_lib.i64p), int(table.shape[0]), C.c_void_p(table.ctypes.data))
        if self.kernel == "tile":
            _lib.call("etl_adjplan_sources", self.handle, int(sim), *args)
        else:
            _lib.call("etl_adjoint_sources", self.handle, *args)

    def temporal(self) -> int:
        """The temporal depth of the reverse steps' blocks (1 for the plain kernels)."""
        if self.kernel != "tile":
            return 1
        t = _lib.EtlTune()
        _lib.call("etl_adjplan_tune", self.handle, C.byref(t))
        return int(t.temporal)

    def run(self, n_first: int, count: int, hist_row0: int = 0) -> None:
        if self.kernel == "tile":
            _lib.call("etl_adjplan_run_rows", self.handle, int(n_first), int(count),
                      int(hist_row0))
        else:
            if hist_row0:
                raise ValueError("the plain reverse kernels read history from row 0")
            _lib.call("etl_adjoint_run", self.handle, int(n_first), int(count))

    def close(self) -> None:
        if self.handle:
            name = "etl_adjplan_destroy" if self.kernel == "tile" else "etl_adjoint_destroy"
            _lib.call(name, self.handle)
            self.handle = C.c_void_p()


def _function():
    import torch

    class PermittivityGradient(torch.autograd.Function):
        """J(ε) with ∂J/∂ε by the adjoint; ε a (npx, npy) tensor of pixel permittivities."""

        @staticmethod
        def forward(ctx, eps, problem):
            tape = problem.forward(eps.detach().to("cpu", torch.float64).numpy())
            ctx.problem, ctx.tape = problem, tape
            return torch.tensor(tape.J, dtype=eps.dtype, device=eps.device)

        @staticmethod
        def backward(ctx, grad_output):
            g = ctx.problem.backward(ctx.tape)
            ctx.tape = None
            return grad_output * torch.as_tensor(g, dtype=grad_output.dtype,
                                                 device=grad_output.device), None

    return PermittivityGradient


def objective(eps, problem: AdjointProblem):
    """J(ε) as a differentiable torch scalar (``torch.autograd.Function``)."""
    return _function().apply(eps, problem)
