"""Host writes into a planned state, stale design-loop tapes and backend selection
on the HIP path.

* A fused plan takes psi as +0 outside the CPML strips and fields without -0 (its
  interior tiles compute ``d`` for the reference's ``d*1 + (+0)``); values written
  from the host after an ``advance`` must make the next ``advance`` check again and
  step instead, bit for bit like the NumPy oracle given the same writes.
* A backward pass on a tape whose design-loop state a later forward overwrote
  returns the gradient of its own permittivity.
* ``"auto"`` picks the HIP backend on an MI355X (the reference's G5,
  tests/gpu/test_cupy_equality.py:183-187).
"""

import numpy as np
import pytest

from etalon_amd.solver import FDTDSolver
from oracle import te_numpy as ref
from tests.sims import bitwise_equal, full_sim, small_sim

pytestmark = pytest.mark.gpu
FIELDS = ("ex", "ey", "hz", "psi_ex_y", "psi_ey_x", "psi_hz_x", "psi_hz_y")


def _same(gpu, cpu) -> None:
    for name in FIELDS:
        assert bitwise_equal(getattr(gpu.state, name).get(), getattr(cpu.s, name)), name
    for mon, acc in zip(gpu.monitors, cpu.acc):
        assert bitwise_equal(mon.acc.get(), acc)


@pytest.mark.parametrize("precision", ["single", "double"])
@pytest.mark.parametrize("write", ["psi_interior", "negative_zero", "psi_strip"])
def test_host_write_after_advance_is_seen(hip, precision, write):
    """Steps 0-59 fused, a host write, steps 60-149 fused: equal to the oracle given the
    same write.  psi inside the interior and a -0 field value break the plan's
    preconditions (the solver then steps); psi written inside a strip does not."""
    sim = full_sim(precision, run_time=1.0)
    gpu, cpu = FDTDSolver(sim), ref.Solver(sim)
    gpu.advance(60)
    cpu.run(60)
    _same(gpu, cpu)
    nx, ny = gpu.grid.nx, gpu.grid.ny
    real = gpu.real
    if write == "psi_interior":
        name, key, value = "psi_hz_x", (nx // 2, ny // 2), real(3e-4)
    elif write == "negative_zero":
        name, key, value = "hz", (nx // 2 + 3, ny // 3), real(-0.0)
    else:
        name, key, value = "psi_ey_x", (2, ny // 2), real(-2e-3)
    plan = gpu._plan
    getattr(gpu.state, name)[key] = value
    getattr(cpu.s, name)[key] = value
    gpu.advance(90)
    cpu.run(90)
    _same(gpu, cpu)
    if write == "psi_strip":
        assert gpu._plan is not None and gpu._plan is not plan  # checked again, still fused
    else:
        assert gpu._plan is None  # the check failed: advance stepped


def test_stale_tape_backward_returns_its_own_gradient(hip):
    """J1 = forward(a), J2 = forward(b) on one design-loop problem, then backward of the
    first tape: the gradient at a (a fresh problem's, bit for bit), and the second tape
    still gives the gradient at b."""
    from etalon_amd.adjoint import AdjointProblem
    from tests.adjoint_sims import design_problem

    sim, design = design_problem("single", seed=12, steps=301)
    a = np.asarray(design.permittivity)
    b = np.random.default_rng(5).uniform(2.085, 12.11, size=a.shape)
    loop = AdjointProblem(sim, ["out"], checkpoint_every=64)
    tape_a = loop.forward(a)
    tape_b = loop.forward(b)
    g_a = loop.backward(tape_a)
    g_b = loop.backward(tape_b)
    Ja, ref_a = AdjointProblem(sim, ["out"], checkpoint_every=64, reuse=False).run(a)
    Jb, ref_b = AdjointProblem(sim, ["out"], checkpoint_every=64, reuse=False).run(b)
    assert tape_a.J == Ja and tape_b.J == Jb
    assert np.array_equal(g_a, ref_a)
    assert np.array_equal(g_b, ref_b)
    assert not np.array_equal(g_a, g_b)
    loop.close()


def test_stale_batch_tape_backward_returns_its_own_gradient(hip):
    from etalon_amd.adjoint import AdjointProblem
    from etalon_amd.batch_adjoint import BatchAdjointProblem
    from tests.adjoint_sims import design_problem

    sims = [design_problem("single", seed=s, steps=301)[0] for s in (3, 4)]
    rng = np.random.default_rng(8)
    shape = np.asarray(sims[0].structures[-1].permittivity).shape
    first = [rng.uniform(2.085, 12.11, size=shape) for _ in sims]
    second = [rng.uniform(2.085, 12.11, size=shape) for _ in sims]
    loop = BatchAdjointProblem([AdjointProblem(s, ["out"], checkpoint_every=64) for s in sims])
    tape1 = loop.forward(first)
    tape2 = loop.forward(second)
    g1 = loop.backward(tape1)
    g2 = loop.backward(tape2)
    fresh = BatchAdjointProblem([AdjointProblem(s, ["out"], checkpoint_every=64) for s in sims],
                                reuse=False)
    _, r1 = fresh.run(first)
    _, r2 = fresh.run(second)
    for got, want in zip(g1, r1):
        assert np.array_equal(got, want)
    for got, want in zip(g2, r2):
        assert np.array_equal(got, want)
    loop.close()


def test_g5_auto_picks_hip(hip):
    from etalon_amd.backends import get_backend

    assert get_backend("auto", "single").name == "hip"
    sim = full_sim("single").model_copy(update={"run_time": 5e-15, "backend": "auto"})
    result = FDTDSolver(sim).run()
    assert result.metadata["backend"] == "hip"
    assert result.metadata["device"].startswith("hip:0")


def test_plan_is_kept_across_advances_of_either_parity(hip):
    """No host write, no new plan: advances that leave the state in either buffer of
    the pairs (odd and even numbers of launches) keep the plan, its tables and graphs."""
    sim = full_sim("single")
    gpu, cpu = FDTDSolver(sim), ref.Solver(sim)
    gpu.advance(7)
    plan = gpu._plan
    for count in (5, 13, 96, 1):
        gpu.advance(count)
        assert gpu._plan is plan, count
    cpu.run(7 + 5 + 13 + 96 + 1)
    _same(gpu, cpu)
