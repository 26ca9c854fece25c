"""Host logic of FDTDSolver.advance, without a GPU: the step tables are refilled in
growing pieces (the host builds the next piece while the GPU runs the current one),
a multi-piece advance leaves the plan the table of its whole range, and a run over
the same steps again (after reset) uses it without a refill; an advance within the
held tables takes the one-call path until a host write or new state arrays send it
through the checks again.  A recording stand-in takes the place of the fused plan."""

from types import SimpleNamespace

import numpy as np

from etalon_amd import solver as solver_mod
from etalon_amd.hip import backend as hip_backend
from etalon_amd.solver import FDTDSolver
from etalon_amd.state import FIELDS


class RecordingPlan:
    def __init__(self):
        self.calls = []
        self.binds = 0
        self.parity = 0
        self.arrays = [tuple(object() for _ in FIELDS) for _ in (0, 1)]

    def source_table(self, n0, delta):
        self.calls.append(("delta", n0, len(delta)))

    def phasor_table(self, mon_id, n0, table):
        self.calls.append(("phasor", mon_id, n0, len(table)))

    def advance(self, n0, rows):
        self.calls.append(("advance", n0, rows))
        self.parity ^= rows % 2  # one launch per step in this stand-in
        return self.parity

    def bind(self):
        self.binds += 1
        for name, array in zip(FIELDS, self.arrays[self.parity]):
            setattr(self.state, name, array)


def make_solver(plan, checked_plan=False):
    """A solver shell around ``plan``.  checked_plan: the plan is the solver's own, as
    after a first advance (the one-call path may take it); else ``_ensure_plan`` hands
    it over on every advance that goes through the checks."""
    s = object.__new__(FDTDSolver)  # no state or device: advance's host logic only
    s.n = 0
    s._plan_tables = (0, 0)
    s._plan_ids = [7]
    s._plan = None
    s._plan_epoch = -1
    s.state = SimpleNamespace(**dict(zip(FIELDS, plan.arrays[0])))
    plan.state = s.state
    s.checks = 0

    def ensure():
        s.checks += 1
        if checked_plan:
            s._plan = plan
        return plan

    s._ensure_plan = ensure
    s.fused_supported = lambda: True
    s._tables = lambda n0, rows: (np.zeros((rows, 1)), [np.zeros((rows, 2))])
    return s


def advances(plan):
    return [c[1:] for c in plan.calls if c[0] == "advance"]


def test_refills_grow_and_stage_the_whole_range():
    plan = RecordingPlan()
    s = make_solver(plan)
    s.advance(10000)
    assert advances(plan) == [(0, 192), (192, 768), (960, 3072), (4032, 5968)]
    # each piece's tables are handed over before its steps; the whole range last
    assert plan.calls[-2:] == [("delta", 0, 10000), ("phasor", 7, 0, 10000)]
    assert s.n == 10000 and s._plan_tables == (0, 10000)
    first = solver_mod._FIRST_TABLE_ROWS
    for n0, rows in advances(plan)[:-1]:
        assert rows % first == 0  # no refill splits a step block or trace chunk

    plan.calls.clear()
    s.n = 0  # reset: the same steps again
    s.advance(10000)
    assert plan.calls == [("advance", 0, 10000)]


def test_short_advances_stage_ahead():
    """A refill covers at least _FIRST_TABLE_ROWS steps, so the short advances that
    follow (a step loop of small advances, a timed region after a warm-up) upload
    nothing until they pass its end."""
    plan = RecordingPlan()
    s = make_solver(plan)
    first = solver_mod._FIRST_TABLE_ROWS
    s.advance(100)
    assert plan.calls == [("delta", 0, first), ("phasor", 7, 0, first), ("advance", 0, 100)]
    assert s._plan_tables == (0, first)
    plan.calls.clear()
    s.advance(50)  # within the held table: no refill
    assert plan.calls == [("advance", 100, 50)]
    plan.calls.clear()
    s.advance(100)  # the rest of the held table, then a new piece from its end
    assert plan.calls == [("advance", 150, first - 150), ("delta", first, first),
                          ("phasor", 7, first, first), ("advance", first, 250 - first)]


def test_warm_up_then_timed_steps_upload_nothing():
    """bench.py's pattern: W warm-up steps, then K timed ones within the first table."""
    plan = RecordingPlan()
    s = make_solver(plan)
    s.advance(5)
    plan.calls.clear()
    s.advance(20)
    assert plan.calls == [("advance", 5, 20)]


def test_long_runs_cap_the_piece_and_stage_nothing_larger():
    plan = RecordingPlan()
    s = make_solver(plan)
    cap = solver_mod._PLAN_TABLE_ROWS
    s.advance(60000)
    rows = [r for _, r in advances(plan)]
    assert rows == [192, 768, 3072, 12288, cap, 60000 - (192 + 768 + 3072 + 12288 + cap)]
    assert max(c[2] for c in plan.calls if c[0] == "delta") <= cap
    assert s.n == 60000


def test_one_call_path_and_what_sends_advance_back_through_the_checks():
    """Within the held tables an advance of a checked plan is one plan call (its state
    rebound only when the parity changed); a host write into any device buffer, new
    state arrays or steps beyond the tables go through the checks again."""
    plan = RecordingPlan()
    s = make_solver(plan, checked_plan=True)
    s.advance(5)  # checks, stages the first table
    assert s.checks == 1 and s._plan is plan
    plan.calls.clear()
    binds = plan.binds
    s.advance(20)  # even: parity unchanged, no rebind
    assert plan.binds == binds
    s.advance(3)   # odd: rebind
    assert plan.calls == [("advance", 5, 20), ("advance", 25, 3)] and plan.binds == binds + 1
    assert s.checks == 1 and s.n == 28
    assert s.state.hz is plan.arrays[plan.parity][FIELDS.index("hz")]

    hip_backend._write_epoch += 1  # what DeviceArray.set does
    plan.calls.clear()
    s.advance(2)
    assert s.checks == 2 and plan.calls[0] == ("advance", 28, 2)
    s.advance(2)
    assert s.checks == 2  # checked at the new epoch

    s.state.ex = object()  # a state array replaced by hand
    s.advance(2)
    assert s.checks == 3

    s.state.ex = plan.arrays[plan.parity][0]
    first = solver_mod._FIRST_TABLE_ROWS
    s.advance(first)  # runs past the held table: refills through the checks
    assert s.checks == 4 and s.n == 34 + first

