"""Fused time-stepping plan: Python side of etl_plan_* (include/etalon_hip.h).

A plan advances one simulation, or a batch of independent simulations of any
sizes, with the fused kernels.  It owns nothing but device-side bookkeeping; the
ping-pong buffer pairs are DeviceArrays held here.  After ``advance`` the state
lives in buffer ``parity`` of each pair (the same parity for every simulation).
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

from etalon_amd.hip import _lib
from etalon_amd.hip import backend as _backend
from etalon_amd.hip.backend import DeviceArray, _precision_of, zeros
from etalon_amd.state import FIELDS, PROFILES

COMPONENT_CODE = {"Ex": 0, "Ey": 1, "Hz": 2}
MAX_SOURCE_NODES = 16  # kMaxSources in csrc/fused.h, per simulation


@dataclass
class Tune:
    """Fused-kernel shape; 0 = the library default: vec 1, one wave per workgroup,
    ``launch`` 1 (one kernel per step block over every tile kind) where the shape
    allows it (vec 1 with band 8 or 16), else 3 (one kernel per tile kind on three
    streams); for large plans (at least 8 tiles of 16 x 56 per CU) temporal depth 6
    with 96-step trace chunks and 16-row tiles — depth 8 in float when the fields
    (32 bytes per cell) exceed the 256 MiB Infinity Cache and every grid has at least
    4 M cells; else 8-row tiles and depth 4
    with 64-step chunks.  Depths 6 and 8 run with launch 1 only.  ``uniform_cb`` (launch
    1; 0 = on, -1 = off): interior tiles whose window holds one cb_ex and one cb_ey value
    read those two values instead of the cb rows (plans whose cb arrays are rewritten
    every run, as in a design loop, turn it off: each write lays the tiles out again)."""

    band_rows: int = 0
    vec: int = 0
    use_graph: int = 1
    trace_chunk: int = 0
    temporal: int = 0
    waves: int = 0
    launch: int = 0
    uniform_cb: int = 0

    def as_struct(self) -> _lib.EtlTune:
        return _lib.EtlTune(self.band_rows, self.vec, self.use_graph, self.trace_chunk,
                            self.temporal, self.waves, self.launch, self.uniform_cb)


def _describe(state, pair: dict) -> _lib.EtlTeDesc:
    d = _lib.EtlTeDesc()
    d.precision = _precision_of(state.hz.dtype)
    d.nx, d.ny = int(state.nx), int(state.ny)
    d.db, d.inv_dx, d.inv_dy = float(state.db), float(state.inv_dx), float(state.inv_dy)
    for k, name in enumerate(FIELDS):
        d.field[k][0] = pair[name][0].ptr
        d.field[k][1] = pair[name][1].ptr
    d.cb_ex = state.cb_ex.ptr
    d.cb_ey = state.cb_ey.ptr
    for name in PROFILES:
        setattr(d, name, getattr(state, name).ptr)
    return d


class FusedPlan:
    """Fused kernel plan over one TEState or a sequence of them (DeviceArrays)."""

    def __init__(self, states, tune: Optional[Tune] = None):
        self.batch = isinstance(states, (list, tuple))
        self.states: Sequence = list(states) if self.batch else [states]
        self.state = self.states[0]
        self.precision = _precision_of(self.state.hz.dtype)
        self.pair_sets = [{name: [getattr(s, name), zeros(getattr(s, name).shape,
                                                          getattr(s, name).dtype)]
                           for name in FIELDS} for s in self.states]
        self.pairs = self.pair_sets[0]
        # the first simulation's seven arrays (FIELDS order) in buffer 0 and in buffer 1
        self.arrays = [tuple(self.pairs[name][b] for name in FIELDS) for b in (0, 1)]
        descs = (_lib.EtlTeDesc * len(self.states))()
        for k, (s, pair) in enumerate(zip(self.states, self.pair_sets)):
            descs[k] = _describe(s, pair)
        self._descs = descs
        handle = C.c_void_p()
        t = tune.as_struct() if tune is not None else None
        _lib.call("etl_plan_create_batch", descs, len(self.states),
                  C.byref(t) if t is not None else None, C.byref(handle))
        self.handle = handle
        self.parity = 0
        self.n_src = [0] * len(self.states)
        self._cb_marks = self._material_marks()
        self._epoch = _backend.host_write_epoch()
        # the step entry point and its parity output, looked up once (advance is
        # called once per step block by interactive loops)
        self._advance_rows = _lib.load().etl_plan_advance_rows
        self._parity_out = C.c_int()
        self._parity_ref = C.byref(self._parity_out)

    def _material_marks(self) -> tuple:
        return tuple((s.cb_ex.writes, s.cb_ey.writes) for s in self.states)

    def _check_materials(self) -> None:
        """cb_ex or cb_ey written from the host since the plan laid out its tiles: lay
        them out again (uniform-material tiles read the cb values of the layout)."""
        epoch = _backend.host_write_epoch()
        if epoch == self._epoch:  # no host write into any buffer since the last check
            return
        marks = self._material_marks()
        if marks != self._cb_marks:
            _lib.call("etl_plan_touch", self.handle)
            self._cb_marks = marks
        self._epoch = epoch

    def close(self) -> None:
        if self.handle:
            _lib.call("etl_plan_destroy", self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_sources(self, comp: np.ndarray, i: np.ndarray, j: np.ndarray, sim: int = 0) -> None:
        comp = np.ascontiguousarray(comp, dtype=np.int32)
        i = np.ascontiguousarray(i, dtype=np.int64)
        j = np.ascontiguousarray(j, dtype=np.int64)
        _lib.call("etl_plan_set_sources_sim", self.handle, int(sim), int(comp.size),
                  comp.ctypes.data_as(C.POINTER(C.c_int)), i.ctypes.data_as(_lib.i64p),
                  j.ctypes.data_as(_lib.i64p))
        self.n_src[sim] = int(comp.size)

    def source_table(self, n0: int, delta: np.ndarray, sim: int = 0) -> None:
        dtype = self.states[sim].hz.dtype
        delta = np.ascontiguousarray(delta, dtype=dtype)
        _lib.call("etl_plan_source_table_sim", self.handle, int(sim), int(n0), int(delta.shape[0]),
                  C.c_void_p(delta.ctypes.data))

    def add_monitor(self, component: str, flat_index: np.ndarray, acc: DeviceArray,
                    sim: int = 0) -> int:
        idx = np.ascontiguousarray(flat_index, dtype=np.int64)
        n_f = acc.shape[0]
        out = C.c_int()
        _lib.call("etl_plan_add_monitor_sim", self.handle, int(sim), COMPONENT_CODE[component],
                  int(idx.size), idx.ctypes.data_as(_lib.i64p), int(n_f), C.c_void_p(acc.ptr),
                  C.byref(out))
        return out.value

    def add_probe(self, component: str, flat_index: np.ndarray, out: DeviceArray,
                  sim: int = 0) -> None:
        """Time-domain probe: row r of ``out`` (rows, n_nodes) receives the nodes' values
        of step n + r during ``advance(n, count)`` (count <= rows)."""
        idx = np.ascontiguousarray(flat_index, dtype=np.int64)
        if out.ndim != 2 or out.shape[1] != idx.size:
            raise ValueError("probe output must be (rows, n_nodes)")
        _lib.call("etl_plan_add_probe_sim", self.handle, int(sim), COMPONENT_CODE[component],
                  int(idx.size), idx.ctypes.data_as(_lib.i64p), C.c_void_p(out.ptr),
                  int(out.shape[0]))

    def share_phasors(self, mon_id: int, owner: int, sim: int = 0) -> None:
        """Monitor ``mon_id`` reads the phasor columns of monitor ``owner`` (equal tables)."""
        _lib.call("etl_plan_share_phasors", self.handle, int(sim), int(mon_id), int(owner))

    def phasor_table(self, mon_id: int, n0: int, table: np.ndarray, sim: int = 0) -> None:
        table = np.ascontiguousarray(table)
        _lib.call("etl_plan_phasor_table_sim", self.handle, int(sim), int(mon_id), int(n0),
                  int(table.shape[0]), C.c_void_p(table.ctypes.data))

    def advance(self, n0: int, count: int, probe_row0: int = 0) -> int:
        """Steps [n0, n0 + count); probes write rows probe_row0 .. probe_row0 + count - 1."""
        self._check_materials()
        rc = self._advance_rows(self.handle, n0, count, probe_row0, self._parity_ref)
        if rc:
            raise _lib.HipError("etl_plan_advance_rows", rc, _lib.last_error(_lib.load()))
        self.parity = self._parity_out.value
        return self.parity

    def set_checkpoints(self, arrays: Sequence, n_seg: int, sim: int = 0) -> None:
        """Checkpoint storage: arrays[k] holds ``n_seg`` copies of field k (FIELDS order)."""
        ptrs = (C.c_void_p * 7)(*[a.ptr for a in arrays])
        _lib.call("etl_plan_set_checkpoints_sim", self.handle, int(sim), ptrs, int(n_seg))

    def set_checkpoint_window(self, arrays: Sequence, n_seg: int, win_nx: int, win_ny: int,
                              row0: int, col0: int, sim: int = 0) -> None:
        """Checkpoint storage that holds only a window of simulation ``sim``'s fields:
        arrays[k] (FIELDS order) holds ``n_seg`` slots, each the block of field k with
        Hz nodes [row0, row0 + win_nx) x [col0, col0 + win_ny) (Ex one column more, Ey
        one row more) stored row after row.  A plan over that window's grid takes the
        same arrays with ``set_checkpoints``."""
        ptrs = (C.c_void_p * 7)(*[a.ptr for a in arrays])
        _lib.call("etl_plan_set_checkpoint_window_sim", self.handle, int(sim), ptrs, int(n_seg),
                  int(win_nx), int(win_ny), int(row0), int(col0))

    def checkpoint(self, seg: int, restore: bool) -> None:
        """Save the state of every simulation into slot ``seg`` or restore it from there
        (stream-ordered, on the plan's lane), in the buffers that hold the state now."""
        _lib.call("etl_plan_checkpoint", self.handle, int(seg), 1 if restore else 0)

    def advance_checkpointed(self, n0: int, count: int, every: int) -> int:
        """Steps [n0, n0 + count), the state saved into slot c // every before each
        segment of ``every`` steps starting at step c."""
        self._check_materials()
        parity = C.c_int()
        _lib.call("etl_plan_advance_checkpointed", self.handle, int(n0), int(count), int(every),
                  C.byref(parity))
        self.parity = parity.value
        return self.parity

    def set_lane(self, lane: int) -> None:
        """Issue this plan's work on stream lane 0 or 1 (see etl_lane_wait)."""
        _lib.call("etl_plan_set_lane", self.handle, int(lane))

    def prepare(self) -> None:
        """Build the layout (tile work-list, monitor slots) now rather than at the
        first table upload or advance."""
        self._check_materials()
        _lib.call("etl_plan_prepare", self.handle)

    def resolved_tune(self) -> Tune:
        """The tuning in effect (library defaults filled in)."""
        t = _lib.EtlTune()
        _lib.call("etl_plan_tune", self.handle, C.byref(t))
        return Tune(t.band_rows, t.vec, t.use_graph, t.trace_chunk, t.temporal, t.waves, t.launch,
                    t.uniform_cb)

    def time(self, n0: int, count: int) -> dict:
        """HIP-event timing of ``count`` steps launched directly: ``launch_ms`` (average
        fused launch, all tile kinds), ``kernel_ms`` (the kernel that advances the
        interior tiles — with ``launch`` 1 the one kernel of every tile kind, with 3 the
        interior-tile kernel alone), ``total_ms``, ``kernel_cells`` (Hz cells that
        kernel advances per step), ``launches``, ``depth`` (steps per launch) and
        ``launch``."""
        self._check_materials()
        t, parity = _lib.EtlTiming(), C.c_int()
        _lib.call("etl_plan_time", self.handle, int(n0), int(count), C.byref(t), C.byref(parity))
        self.parity = parity.value
        return {name: getattr(t, name) for name, _ in _lib.EtlTiming._fields_}

    def bind(self) -> None:
        """Point each state's field attributes at the buffers holding its current state."""
        for s, pair in zip(self.states, self.pair_sets):
            for name in FIELDS:
                setattr(s, name, pair[name][self.parity])
