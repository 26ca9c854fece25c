"""Host-side checks without a GPU: host writes into device arrays are counted (a
fused plan re-checks its preconditions when the counts of its state change), a stale
design-loop tape is replayed instead of reading the overwritten session, and backend
selection never falls back to a CPU path."""

import numpy as np
import pytest

from etalon_amd.hip import backend as hb


class _FakeBuffer:
    """A stand-in device allocation (no HIP): enough for DeviceArray's bookkeeping."""

    def __init__(self, nbytes):
        self.ptr = 0x1000
        self.nbytes = nbytes
        self.writes = 0


@pytest.fixture
def no_device(monkeypatch):
    calls = []
    monkeypatch.setattr(hb._lib, "call", lambda name, *args: calls.append(name))
    return calls


def _array(shape=(4, 5), dtype=np.float32):
    buf = _FakeBuffer(int(np.prod(shape)) * np.dtype(dtype).itemsize)
    return hb.DeviceArray(shape, dtype, buffer=buf)


def test_set_and_item_assignment_count_writes(no_device, monkeypatch):
    a = _array()
    assert a.writes == 0
    a.set(np.ones((4, 5)))
    assert a.writes == 1
    monkeypatch.setattr(hb.DeviceArray, "get", lambda self: np.zeros(self.shape, self.dtype))
    a[1, 2] = -0.0  # read, modify, write back
    assert a.writes == 2
    a[...] = 3.0
    assert a.writes == 3
    row = a.rows(1, 3)  # a view shares the count of its buffer
    row.set(np.zeros((2, 5)))
    assert a.writes == 4 and row.writes == 4
    assert no_device.count("etl_memcpy_h2d") == 4


def test_solver_plan_marks_cover_both_buffers_of_every_pair(no_device):
    """The plan's marks change with a host write into either buffer of a field pair,
    not with which buffer holds the state after a step block."""
    from etalon_amd.solver import FDTDSolver
    from etalon_amd.state import FIELDS

    class Plan:
        pairs = {name: [_array(), _array()] for name in FIELDS}

    plan = Plan()
    before = FDTDSolver._pair_marks(plan)
    assert len(before) == 2 * len(FIELDS)
    plan.pairs["psi_hz_x"][1].set(np.zeros((4, 5)))
    after = FDTDSolver._pair_marks(plan)
    assert before != after
    k = 2 * FIELDS.index("psi_hz_x") + 1
    assert after[k] == before[k] + 1 and after[:k] == before[:k] and after[k + 1:] == before[k + 1:]


def _problem():
    from etalon_amd.adjoint import AdjointProblem
    from tests.adjoint_sims import design_problem

    sim, _ = design_problem("single", seed=1, steps=100)
    return AdjointProblem(sim, ["out"], checkpoint_every=64)


def test_stale_session_tape_is_replayed(monkeypatch):
    """backward(tape) of a session tape whose generation is older than the session's
    runs a new forward at the tape's own permittivity (no session) and the backward
    pass on that; a current tape goes to the session's backward."""
    from etalon_amd.adjoint import Tape, _Session

    prob = _problem()
    solver = object()
    prob._session = _Session(solver=solver, checkpoints={}, starts=[0], generation=3)
    replayed, session_calls = [], []
    fresh_solver = type("FreshSolver", (), {"real": np.float32})()

    def forward_new(eps, session):
        replayed.append((np.array(eps), session))
        return Tape(solver=fresh_solver, J=1.0, eps=np.array(eps), generation=0)

    prob._forward_new = forward_new
    prob._backward_session = lambda tape: session_calls.append(tape) or "session"
    import etalon_amd.adjoint as adjoint_mod

    def no_alloc(*args):
        raise RuntimeError("non-session backward")

    monkeypatch.setattr(adjoint_mod, "zeros", no_alloc)
    eps = np.full(prob.shape, 4.0)
    current = Tape(solver=solver, J=2.0, eps=eps, generation=3)
    assert prob.backward(current) == "session"
    stale = Tape(solver=solver, J=2.0, eps=eps, generation=2)
    with pytest.raises(RuntimeError, match="non-session backward"):
        prob.backward(stale)  # replayed, then the non-session backward on the new tape
    assert len(replayed) == 1 and replayed[0][1] is False
    assert np.array_equal(replayed[0][0], eps)
    assert len(session_calls) == 1
    prob._session = None  # nothing to release


def test_batch_stale_tape_is_replayed():
    from etalon_amd.batch_adjoint import BatchAdjointProblem, BatchTape, _BatchSession

    batch_adj = BatchAdjointProblem([_problem(), _problem()])
    batch = object()
    batch_adj._session = _BatchSession(batch=batch, checkpoints=[], histories=[], generation=2)
    replayed = []

    def forward_new(eps_list, session):
        replayed.append(session)
        raise RuntimeError("replayed")

    batch_adj._forward_new = forward_new
    batch_adj._backward_session = lambda tape: "session"
    assert batch_adj.backward(BatchTape(batch=batch, generation=2)) == "session"
    with pytest.raises(RuntimeError, match="replayed"):
        batch_adj.backward(BatchTape(batch=batch, eps=[1, 2], generation=1))
    assert replayed == [False]
    batch_adj._session = None


def test_auto_and_hip_never_fall_back_to_a_cpu_path(monkeypatch):
    """Without a device, "auto" and "hip" raise RuntimeError (no CPU compute path in
    this package); unknown names raise ValueError."""
    from etalon_amd.backends import get_backend

    def no_device(device=0):
        raise RuntimeError("no HIP device")

    monkeypatch.setattr(hb._lib, "init", no_device)
    for name in ("auto", "hip"):
        with pytest.raises(RuntimeError):
            get_backend(name, "single")
    with pytest.raises(ValueError, match="unknown backend 'numpy'"):
        get_backend("numpy", "single")
