"""HIP kernels vs the NumPy oracle, bit for bit (the reference's G1–G4 equality suite,
tests/gpu/test_cupy_equality.py:36-180, tightened from rtol to bitwise equality).

Every comparison is on the bit pattern (±0 distinguished) in float32 and float64,
through the C ABI of libetalon_hip.so.
"""

import numpy as np
import pytest

from oracle import te_numpy as ref
from tests.sims import bitwise_equal, full_sim, small_sim, vacuum_pulse

pytestmark = pytest.mark.gpu

PRECISIONS = [("single", np.float32), ("double", np.float64)]
FIELDS = ("ex", "ey", "hz", "psi_ex_y", "psi_ey_x", "psi_hz_x", "psi_hz_y")


def random_state(dtype, seed=1, nx=67, ny=53):
    """Random fields, ψ and CPML profiles on every node, Δx ≠ Δy (G1/G2 of the reference)."""
    rng = np.random.default_rng(seed)
    dx, dy = 0.05, 0.04
    g = ref.Grid(nx, ny, dx, dy, ref.time_step(0.99, dx, dy), 0, 0.0, 0.0, 0.0)
    eps = {c: rng.uniform(1.0, 12.0, g.shape(c)) for c in ("Ex", "Ey")}

    def prof(n):
        return (rng.uniform(0.2, 1.0, n + 1), -rng.uniform(0, 0.8, n + 1), rng.uniform(0.3, 1.0, n + 1),
                rng.uniform(0.2, 1.0, n), -rng.uniform(0, 0.8, n), rng.uniform(0.3, 1.0, n))

    s = ref.State(g, eps, prof(nx), prof(ny), dtype)
    for name in FIELDS:
        v = rng.standard_normal(getattr(s, name).shape).astype(dtype)
        if name == "hz":
            v *= dtype(1 / 377.0)
        setattr(s, name, v)
    return s


def to_hip_state(s):
    from etalon_amd.hip.backend import to_device
    from etalon_amd.state import TEState

    arrays = {k: to_device(getattr(s, k)) for k in (*FIELDS, "cb_ex", "cb_ey",
              "bx_e", "cx_e", "ikx_e", "bx_h", "cx_h", "ikx_h",
              "by_e", "cy_e", "iky_e", "by_h", "cy_h", "iky_h")}
    return TEState(**arrays, db=s.db, inv_dx=s.inv_dx, inv_dy=s.inv_dy, nx=s.nx, ny=s.ny)


@pytest.mark.parametrize(("precision", "dtype"), PRECISIONS)
@pytest.mark.parametrize("update", ["h", "e"])
def test_g1_g2_update_kernels_bitwise(hip, precision, dtype, update):
    cpu = random_state(dtype)
    gpu = to_hip_state(cpu)
    (ref.update_h if update == "h" else ref.update_e)(cpu)
    (hip.update_h_te if update == "h" else hip.update_e_te)(gpu)
    for name in FIELDS:
        assert bitwise_equal(getattr(gpu, name).get(), getattr(cpu, name)), (update, name)


@pytest.mark.parametrize(("precision", "dtype"), PRECISIONS)
def test_g3_dft_kernel_bitwise(hip, precision, dtype):
    rng = np.random.default_rng(2)
    cdt = np.complex64 if dtype == np.float32 else np.complex128
    field = rng.standard_normal((67, 54)).astype(dtype)
    index = rng.choice(field.size, size=301, replace=False).astype(np.int64)
    ph = (np.exp(1j * rng.uniform(0, 2 * np.pi, 13)) * 1e-16).astype(cdt)
    acc = ((rng.standard_normal((13, 301)) + 1j * rng.standard_normal((13, 301))) * 1e-15).astype(cdt)
    acc_gpu = hip.to_device(acc)
    ref.dft(acc, field, index, ph)
    hip.accumulate_dft(acc_gpu, hip.to_device(field), hip.to_device(index), hip.to_device(ph))
    assert bitwise_equal(acc_gpu.get(), acc)


def _compare_solver(gpu_solver, cpu_solver, nonzero_monitors=True):
    for name, value in cpu_solver.fields().items():
        assert np.abs(value).max() > 0, name
        assert bitwise_equal(gpu_solver.fields()[name], value), name
    for mg, acc in zip(gpu_solver.monitors, cpu_solver.acc):
        label = f"{mg.sampling.monitor}/{mg.sampling.component}"
        if nonzero_monitors:
            assert np.abs(acc).max() > 0, label
        assert bitwise_equal(mg.acc.get(), acc), label


@pytest.mark.parametrize(("precision", "dtype"), PRECISIONS)
def test_g4_full_run_step_path(hip, precision, dtype):
    from etalon_amd.solver import FDTDSolver

    sim = full_sim(precision)
    gpu, cpu = FDTDSolver(sim), ref.Solver(sim)
    for _ in range(300):
        gpu.step()
    cpu.run(300)
    _compare_solver(gpu, cpu)


@pytest.mark.parametrize(("precision", "dtype"), PRECISIONS)
def test_g4_full_run_fused_path(hip, precision, dtype):
    """600 steps through the fused plan in uneven chunks (odd counts flip the buffers)."""
    from etalon_amd.solver import FDTDSolver

    sim = full_sim(precision)
    gpu, cpu = FDTDSolver(sim), ref.Solver(sim)
    for count in (1, 63, 64, 129, 343):
        gpu.advance(count)
    cpu.run(600)
    assert gpu.n == cpu.n == 600
    _compare_solver(gpu, cpu)


@pytest.mark.parametrize(("precision", "dtype"), PRECISIONS)
def test_fused_and_step_paths_interleave(hip, precision, dtype):
    from etalon_amd.solver import FDTDSolver

    sim = small_sim(precision=precision)
    gpu, cpu = FDTDSolver(sim), ref.Solver(sim)
    gpu.advance(37)
    for _ in range(5):
        gpu.step()
    gpu.advance(58)
    cpu.run(100)
    _compare_solver(gpu, cpu)


@pytest.mark.parametrize(("precision", "dtype"), PRECISIONS)
def test_negative_zero_state_steps_instead(hip, precision, dtype):
    """The tile kernels take d*1 + (+0) as d, which needs fields without -0: a state
    holding -0 makes plan creation return ETL_ESTATE, and advance() steps with the
    per-half-step kernels instead — bit for bit the oracle's result, -0 included."""
    from etalon_amd.hip._lib import HipError
    from etalon_amd.hip.plan import FusedPlan
    from etalon_amd.solver import FDTDSolver

    sim = small_sim(precision=precision)
    gpu, cpu = FDTDSolver(sim), ref.Solver(sim)
    for name in ("hz", "ey"):  # -0 on interior nodes away from the source
        host = np.zeros(getattr(cpu.s, name).shape, dtype)
        host[5:9, 20:26] = -0.0
        setattr(cpu.s, name, host.copy())
        getattr(gpu.state, name).set(host)
    with pytest.raises(HipError) as err:
        FusedPlan(gpu.state)
    assert err.value.code == -3 and "-0" in str(err.value)
    gpu.advance(40)
    cpu.run(40)
    _compare_solver(gpu, cpu)
    zeros = gpu.state.hz.get()[5:9, 20:26]
    print(f"{precision}: -0 kept at {int(np.signbit(zeros).sum())} Hz nodes", flush=True)


@pytest.mark.parametrize(("precision", "dtype"), PRECISIONS)
def test_vacuum_pulse_2000_steps_bitwise(hip, precision, dtype):
    """184² grid (8 µm + 12-layer CPML), courant 0.5, 4 point monitors × 11 λ, 2000 steps."""
    from etalon_amd.solver import FDTDSolver

    sim = vacuum_pulse(precision)
    gpu = FDTDSolver(sim)
    assert (gpu.grid.nx, gpu.grid.ny, gpu.n_steps) == (184, 184, 2000)
    result = gpu.run()
    cpu = ref.Solver(sim).run()
    # Ex on the source's axis (y = 0) is exactly zero by mirror symmetry.
    _compare_solver(gpu, cpu, nonzero_monitors=False)
    assert sum(np.abs(acc).max() > 0 for acc in cpu.acc) >= 10
    values = cpu.monitor_values()
    for mon in sim.monitors:
        for comp in ("Ex", "Ey", "Hz"):
            assert bitwise_equal(result[mon.name][comp].values, values[(mon.name, comp)])


@pytest.mark.parametrize(("precision", "dtype"), PRECISIONS)
def test_largest_grid_bitwise(hip, precision, dtype):
    """8024 x 8024 cells (64.4 M; the fields exceed the GPU's caches): the throughput
    workload at 8000 x 8000 interior cells from random Ex, Ey, Hz (zero psi), 18 steps of
    the default plan (depth 8 in float, 6 in double) in three advances (6 + 1, 2 + 1,
    then one block of 8, or 6 + 2), every field and the monitor against the oracle, bit
    for bit."""
    from etalon_amd.solver import FDTDSolver
    from etalon_amd.workloads import throughput_sim

    sim = throughput_sim(precision, 8000)
    gpu, cpu = FDTDSolver(sim), ref.Solver(sim)
    assert (gpu.grid.nx, gpu.grid.ny) == (8024, 8024)
    rng = np.random.default_rng(5)
    for name in ("ex", "ey", "hz"):
        host = rng.standard_normal(getattr(cpu.s, name).shape, dtype=dtype)
        if name == "hz":
            host *= dtype(1 / 377.0)
        setattr(cpu.s, name, host)
        getattr(gpu.state, name).set(host)
    gpu.advance(7)
    gpu.advance(3)
    gpu.advance(8)
    cpu.run(18)
    assert gpu.n == cpu.n == 18
    tune = gpu._plan.resolved_tune()
    print(f"{precision} 8024^2: tune {tune.__dict__}", flush=True)
    assert tune.temporal == (8 if precision == "single" else 6)
    for name in FIELDS:
        value = getattr(cpu.s, name)
        assert np.abs(value).max() > 0, name
        assert bitwise_equal(getattr(gpu.state, name).get(), value), name
    assert len(gpu.monitors) == len(cpu.acc) == 1
    for mg, acc in zip(gpu.monitors, cpu.acc):
        assert np.abs(acc).max() > 0
        assert bitwise_equal(mg.acc.get(), acc), mg.sampling.component


_ORACLE_CACHE = {}


def _oracle_full(precision, steps):
    key = (precision, steps)
    if key not in _ORACLE_CACHE:
        _ORACLE_CACHE[key] = ref.Solver(full_sim(precision)).run(steps)
    return _ORACLE_CACHE[key]


@pytest.mark.parametrize(("precision", "dtype"), PRECISIONS)
def test_reset_and_run_again_bitwise(hip, precision, dtype):
    """A design loop's pattern: advance, reset, advance again.  The first advance
    refills the step tables in pieces and leaves the plan the table of its whole range;
    the second run uses that table, a third one other pieces; all three end bit for bit
    at the oracle's state."""
    from etalon_amd.solver import FDTDSolver

    sim = full_sim(precision)
    gpu = FDTDSolver(sim)
    cpu = _oracle_full(precision, 600)
    gpu.advance(600)
    assert gpu._plan_tables == (0, 600)
    _compare_solver(gpu, cpu)
    gpu.reset()
    gpu.advance(600)
    assert gpu._plan_tables == (0, 600)
    _compare_solver(gpu, cpu)
    gpu.reset()
    for count in (100, 37, 463):
        gpu.advance(count)
    _compare_solver(gpu, cpu)


# (vec, temporal depth, band rows): every compiled tile shape is its own code path
TUNES_F32 = [(1, 1, 8), (1, 2, 8), (1, 4, 8), (1, 1, 16), (1, 2, 16), (1, 4, 16),
             (2, 1, 8), (2, 2, 8), (2, 4, 8)]
TUNES_F64 = [(1, 1, 8), (1, 2, 8), (1, 4, 8), (1, 1, 16), (1, 2, 16), (1, 4, 16)]
# launch 3: one kernel per tile kind on three streams; 1: one kernel per step block
@pytest.mark.parametrize(("precision", "dtype"), PRECISIONS)
def test_step_loop_of_short_advances_bitwise(hip, precision, dtype):
    """An interactive loop of short advances (1 to 20 steps, the one-call path of a
    checked plan, whose step counter each advance finds where the previous one left
    it), broken by a timed block and a reset; every run ends bit for bit at the
    oracle's state."""
    from etalon_amd.solver import FDTDSolver

    sim = full_sim(precision)
    gpu = FDTDSolver(sim)
    cpu = _oracle_full(precision, 600)
    counts = [1, 2, 3, 5, 20, 7, 1, 11, 13, 17] * 4 + [18]  # 378 steps
    for count in counts:
        gpu.advance(count)
    lo, hi = gpu._plan_tables
    assert lo <= gpu.n and gpu.n + 24 <= hi
    gpu._plan.time(gpu.n, 24)  # a timed block of direct launches on the same tables
    gpu._plan.bind()
    gpu.n += 24
    gpu.advance(600 - gpu.n)
    _compare_solver(gpu, cpu)
    gpu.reset()
    for count in counts + [600 - sum(counts)]:
        gpu.advance(count)
    _compare_solver(gpu, cpu)


SHAPES = [("single", np.float32, *s, 3) for s in TUNES_F32] + \
    [("double", np.float64, *s, 3) for s in TUNES_F64] + \
    [(p, d, 1, t, 8, 1) for p, d in (("single", np.float32), ("double", np.float64))
     for t in (1, 2, 4, 6)] + [(p, d, 1, t, 16, 1) for p, d in (("single", np.float32),
                                                                 ("double", np.float64))
                               for t in (1, 2, 4, 6)] + \
    [("single", np.float32, 1, 8, 16, 1)]


@pytest.mark.parametrize(("precision", "dtype", "vec", "depth", "band", "launch"), SHAPES)
def test_fused_tile_shapes_bitwise(hip, precision, dtype, vec, depth, band, launch):
    """Temporal blocking depth, lane width, tile height and launch mode never change a
    bit."""
    from etalon_amd.hip.plan import Tune
    from etalon_amd.solver import FDTDSolver

    sim = full_sim(precision)
    gpu = FDTDSolver(sim, tune=Tune(band_rows=band, vec=vec, temporal=depth,
                                    trace_chunk=8 * depth, use_graph=1, launch=launch))
    for count in (3, 64, 129, 204):  # 400 steps: graph chunks, depth-1 tails, odd launches
        gpu.advance(count)
    _compare_solver(gpu, _oracle_full(precision, 400))
