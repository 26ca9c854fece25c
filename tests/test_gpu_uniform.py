"""Uniform-material interior tiles (EtlTune.uniform_cb) against the NumPy oracle, bit for
bit: tiles whose window holds one cb_ex and one cb_ey value read those two values
instead of the cb rows; a host write into cb_ex or cb_ey lays the tiles out again
before the next step.
"""

import numpy as np
import pytest

from oracle import te_numpy as ref
from tests.sims import bitwise_equal

pytestmark = pytest.mark.gpu

PRECISIONS = [("single", np.float32), ("double", np.float64)]
FIELDS = ("ex", "ey", "hz", "psi_ex_y", "psi_ey_x", "psi_hz_x", "psi_hz_y")


def _compare(gpu, cpu):
    assert np.abs(cpu.s.hz).max() > 0
    for name in FIELDS:
        assert bitwise_equal(getattr(gpu.state, name).get(), getattr(cpu.s, name)), name
    for mg, acc in zip(gpu.monitors, cpu.acc):  # the pulse reaches the monitor later
        assert bitwise_equal(mg.acc.get(), acc), mg.sampling.component


def _uniform_share(solver):
    """(uniform cells, all cells) of the solver's plan, as etl_plan_time reports them for
    a block of the next 24 steps (which the solver then counts as run)."""
    plan = solver._plan
    n, count = solver.n, 24
    delta, phasors = solver._tables(n, count)
    if delta is not None:
        plan.source_table(n, delta)
    for mon_id, table in zip(solver._plan_ids, phasors):
        if table is not None:
            plan.phasor_table(mon_id, n, table)
    tm = plan.time(n, count)
    solver._plan_tables = (n, n + count)
    plan.bind()
    solver.n += count
    return tm["uniform_cells"], tm["kernel_cells"]


@pytest.mark.parametrize(("precision", "dtype"), PRECISIONS)
def test_uniform_material_tiles_bitwise(hip, precision, dtype):
    """The throughput workload at 400 x 400 interior cells (oxide cladding, one silicon
    strip): most interior tiles are of one material, the strip's are not; 240 steps in
    uneven advances plus a timed block, fields and the line monitor bit for bit."""
    from etalon_amd.solver import FDTDSolver
    from etalon_amd.workloads import throughput_sim

    sim = throughput_sim(precision, 400)
    gpu, cpu = FDTDSolver(sim), ref.Solver(sim)
    for count in (37, 96, 83):
        gpu.advance(count)
    uniform, cells = _uniform_share(gpu)
    print(f"{precision}: uniform-material cells {uniform} of {cells}", flush=True)
    assert 0 < uniform < cells
    cpu.run(gpu.n)
    _compare(gpu, cpu)


@pytest.mark.parametrize(("precision", "dtype"), PRECISIONS)
def test_material_write_lays_tiles_out_again(hip, precision, dtype):
    """cb_ex and cb_ey rewritten from the host halfway through a run (a silicon band
    across uniform cladding): the next advance classifies the tiles again, and the run
    ends bit for bit where the oracle with the same write ends — stale constants in the
    tiles the band crosses would not."""
    from etalon_amd.solver import FDTDSolver
    from etalon_amd.workloads import throughput_sim

    sim = throughput_sim(precision, 400)
    gpu, cpu = FDTDSolver(sim), ref.Solver(sim)
    gpu.advance(120)
    cpu.run(120)
    before, cells = _uniform_share(gpu)
    cpu.run(24)
    rows = slice(200, 232)
    for name in ("cb_ex", "cb_ey"):
        host = getattr(gpu.state, name).get()
        host[rows] = host[rows] * dtype(0.5)
        getattr(gpu.state, name).set(host)
        getattr(cpu.s, name)[rows] = getattr(cpu.s, name)[rows] * dtype(0.5)
    gpu.advance(100)
    after, _ = _uniform_share(gpu)
    cpu.run(124)
    print(f"{precision}: uniform-material cells {before} -> {after} of {cells}", flush=True)
    assert after < before
    _compare(gpu, cpu)
