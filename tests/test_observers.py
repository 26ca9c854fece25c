"""Run observers (field energy, block-mean snapshot), the result-file layout and the
throughput calibration.

CPU: the oracle against the reference tests' known answers — the energy of constant
fields from the permittivity map (tests/studio/test_jobs.py:277-304), the block
shape and first mean of an arange field (:307-315) — and against exact sums; the
result layout of point, line and plane monitors (etalon/studio/runner.py:314-356).
GPU: etl_field_energy against the oracle's sums (float64 rounding) and
etl_downsample against the oracle's block means bit for bit, on states of real runs.
"""

import json

import numpy as np
import pytest

import etalon_amd as et
from oracle import observe_numpy as obs
from oracle import te_numpy as ref
from tests.sims import full_sim, small_sim


def _energy_sim(precision="single"):
    """The reference test's simulation: a 1.0 x 0.6 µm domain with an n = 2 box."""
    return et.Simulation(
        size=(1.0, 0.6, 0.0),
        structures=[et.Box(size=(0.4, 0.2, 0.0), medium=et.Medium(n=2.0))],
        grid=et.UniformGrid(dl=0.05),
        sources=[et.PointSource(center=(0, 0, 0), component="Hz",
                                pulse=et.GaussianPulse(wavelength=1.55, bandwidth=0.3))],
        run_time=1e-15, precision=precision, backend="hip")


def _constant_fields_expected(solver):
    """½ ΔxΔy (ε0 · 4 · Σ εr(Ex nodes) + μ0 · n_Hz) for Ex = 2 V/m, Hz = 1 A/m."""
    eps = solver.eps["Ex"]
    area = (0.05e-6) ** 2
    return 0.5 * area * (ref.EPS0 * 4.0 * eps.sum() + ref.MU0 * solver.s.hz.size)


def test_oracle_energy_of_constant_fields():
    solver = ref.Solver(_energy_sim())
    solver.s.ex[...] = 2.0
    solver.s.hz[...] = 1.0
    g = solver.g
    W = obs.field_energy(obs.energy_sums(solver.s), g.dt, solver.s.db, g.dx, g.dy)
    assert W == pytest.approx(_constant_fields_expected(solver), rel=1e-5)


def test_oracle_energy_sums_are_the_exact_sums_to_rounding():
    solver = ref.Solver(full_sim("double")).run(120)
    got = obs.energy_sums(solver.s)
    exact = obs.exact_energy_sums(solver.s)
    assert all(e > 0 for e in exact)
    for a, b in zip(got, exact):
        assert a == pytest.approx(b, rel=1e-13, abs=0)


def test_oracle_downsample_known_answer():
    field = np.arange(600 * 10, dtype=np.float32).reshape(600, 10)
    sx, sy = obs.downsample_steps(*field.shape)
    assert (sx, sy) == (3, 1)
    blocks = obs.block_means(field, sx, sy)
    assert blocks.shape == (200, 10) and blocks.dtype == np.float32
    assert blocks[0, 0] == pytest.approx(field[0:3, 0].mean())
    np.testing.assert_allclose(blocks, field.reshape(200, 3, 10, 1).mean(axis=(1, 3)), rtol=1e-6)
    assert obs.downsample_steps(256, 256) == (1, 1)
    assert obs.downsample_steps(257, 513) == (2, 3)


def test_oracle_block_means_of_a_ragged_field():
    """Rows and columns beyond whole blocks are dropped; means to float32 rounding."""
    rng = np.random.default_rng(3)
    f = rng.standard_normal((1031, 777)).astype(np.float32)
    sx, sy = obs.downsample_steps(*f.shape)
    assert (sx, sy) == (5, 4)
    got = obs.block_means(f, sx, sy)
    want = f[:1030, :776].reshape(206, 5, 194, 4).astype(np.float64).mean(axis=(1, 3))
    np.testing.assert_allclose(got, want, rtol=2e-7, atol=1e-7)


def _oracle_result(sim):
    """A SimulationResult assembled from an oracle run (the layout needs no GPU)."""
    from etalon_amd.grid import build_yee_grid
    from etalon_amd.result import SimulationResult, monitor_data
    from etalon_amd.sampling import monitor_samplings

    solver = ref.Solver(sim).run()
    grid = build_yee_grid(sim)
    samplings = [s for m in sim.monitors for s in monitor_samplings(m, grid)]
    data = {m.name: {} for m in sim.monitors}
    for smp, acc in zip(samplings, solver.acc):
        data[smp.monitor][smp.component] = monitor_data(smp, acc, grid.z)
    meta = {"backend": "hip", "device": "hip:0", "precision": sim.precision,
            "n_steps": solver.n, "wall_time_s": float("nan"), "cell_updates_per_s": None}
    return SimulationResult(sim, grid, data, meta)


def test_result_layout_per_monitor_type(tmp_path):
    from etalon_amd.result import monitor_layout, save_result

    wls = [1.5, 1.55, 1.6]
    monitors = [
        et.DFTMonitor(name="point", center=(0.31, 0.12, 0.0), wavelengths=wls),
        et.DFTMonitor(name="line", center=(0.0, 0.1, 0.0), size=(0.6, 0.0, 0.0), wavelengths=wls),
        et.DFTMonitor(name="plane", size=(2.0, 1.5, 0.0), wavelengths=wls, components=("Hz",)),
    ]
    sim = small_sim(monitors=monitors)
    result = _oracle_result(sim)
    arrays, listing = monitor_layout(sim, result)
    assert arrays["m0.Ey"].shape == (1, 1, 3) and arrays["m0.Ey"].dtype == np.complex64
    assert arrays["m0.Ey.x"].tolist() == [0.31] and arrays["m0.Ey.y"].tolist() == [0.12]
    n_line = result["line"]["Hz"].values.shape[0]
    assert arrays["m1.Hz"].shape == (n_line, 1, 3)
    assert np.all(np.abs(arrays["m1.Hz.x"]) <= 0.3 + 1e-9) and arrays["m1.Hz.y"].tolist() == [0.1]
    nx_p, ny_p = result["plane"]["Hz"].values.shape[:2]
    assert arrays["m2.Hz"].shape == (nx_p, ny_p, 3)
    assert "m2.Ex" not in arrays
    np.testing.assert_array_equal(arrays["m2.wavelengths"], wls)
    assert [m["geometry"] for m in listing] == ["point", "line", "plane"]
    assert listing[0]["shape"] is None and listing[0]["extent"] is None
    assert listing[0]["components"] == ["Ex", "Ey", "Hz"]
    assert listing[2]["shape"] == [nx_p, ny_p]
    x = arrays["m2.Hz.x"]
    y = arrays["m2.Hz.y"]
    assert listing[2]["extent"] == [float(x[0]), float(x[-1]), float(y[0]), float(y[-1])]
    assert listing[1]["wavelengths_um"] == wls

    meta = save_result(tmp_path / "run", sim, result)
    assert {p.name for p in (tmp_path / "run").iterdir()} == {"result.npz", "meta.json"}
    with np.load(tmp_path / "run" / "result.npz") as saved:
        assert set(saved.files) == set(arrays)
        np.testing.assert_array_equal(saved["m2.Hz"], arrays["m2.Hz"])
    on_disk = json.loads((tmp_path / "run" / "meta.json").read_text())
    assert on_disk["monitors"] == listing and on_disk["steps"] == result.metadata["n_steps"]
    assert meta["wall_time_s"] is None  # NaN is not JSON


# ------------------------------------------------------------------ GPU
@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["single", "double"])
def test_field_energy_matches_the_oracle(hip, precision):
    from etalon_amd.observe import energy_sums, field_energy
    from etalon_amd.solver import FDTDSolver

    sim = full_sim(precision)
    gpu, cpu = FDTDSolver(sim), ref.Solver(sim)
    gpu.advance(150)
    cpu.run(150)
    got = energy_sums(gpu.state)
    exact = obs.exact_energy_sums(cpu.s)  # the same fields (bitwise parity)
    numpy_sums = obs.energy_sums(cpu.s)
    for a, b, c in zip(got, exact, numpy_sums):
        assert b > 0
        assert a == pytest.approx(b, rel=1e-13, abs=0)
        assert a == pytest.approx(c, rel=1e-13, abs=0)
    assert np.array_equal(energy_sums(gpu.state), got)  # a fixed order: the same bits again
    g = cpu.g
    W = obs.field_energy(exact, g.dt, cpu.s.db, g.dx, g.dy)
    assert field_energy(gpu) == pytest.approx(W, rel=1e-13, abs=0)


@pytest.mark.gpu
def test_field_energy_of_constant_fields(hip):
    from etalon_amd.observe import field_energy
    from etalon_amd.solver import FDTDSolver

    sim = _energy_sim()
    gpu, cpu = FDTDSolver(sim), ref.Solver(sim)
    gpu.state.ex.set(np.full(gpu.state.ex.shape, 2.0))
    gpu.state.hz.set(np.full(gpu.state.hz.shape, 1.0))
    assert field_energy(gpu) == pytest.approx(_constant_fields_expected(cpu), rel=1e-5)


@pytest.mark.gpu
def test_reference_style_host_energy_runs_on_device_arrays(hip):
    """The reference's observer does host arithmetic through ``astype`` and ``xp`` =
    NumPy on the backend's arrays; that works on DeviceArrays (host copies) and agrees
    with the device reduction."""
    from etalon_amd.observe import energy_sums
    from etalon_amd.solver import FDTDSolver

    gpu = FDTDSolver(full_sim("single"))
    gpu.advance(90)
    s = gpu.state
    hz = s.hz.astype(np.float64)
    assert isinstance(hz, np.ndarray) and hz.dtype == np.float64
    host = obs.energy_sums(s)  # np.asarray / astype on DeviceArrays
    for a, b in zip(energy_sums(s), host):
        assert a == pytest.approx(b, rel=1e-13, abs=0)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["single", "double"])
@pytest.mark.parametrize("shape", [(600, 10), (1031, 777), (256, 256), (2024, 2024)])
def test_downsample_bitwise_against_the_oracle(hip, precision, shape):
    from etalon_amd.observe import downsample

    dtype = np.float32 if precision == "single" else np.float64
    rng = np.random.default_rng(sum(shape))
    field = rng.standard_normal(shape).astype(dtype)
    blocks, sx, sy = downsample(hip.to_device(field))
    assert (sx, sy) == obs.downsample_steps(*shape)
    got = blocks.get()
    want = obs.block_means(field, sx, sy)
    assert got.dtype == np.float32 and got.shape == want.shape
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


@pytest.mark.gpu
def test_snapshot_of_a_run(hip):
    from etalon_amd.observe import snapshot
    from etalon_amd.solver import FDTDSolver

    sim = et.Simulation(
        size=(12.48, 1.28, 0.0), grid=et.UniformGrid(dl=0.02), boundary=et.PML(layers=12),
        sources=[et.PointSource(center=(0, 0, 0), component="Hz",
                                pulse=et.GaussianPulse(wavelength=1.55, bandwidth=0.3))],
        run_time=1.0, backend="hip")
    gpu = FDTDSolver(sim)
    gpu.advance(200)
    snap = snapshot(gpu)
    assert (gpu.grid.nx, gpu.grid.ny) == (648, 88)
    assert snap["data"].shape == (216, 88) and (snap["sx"], snap["sy"]) == (3, 1)
    assert np.abs(snap["data"]).max() > 0
    x0, x1, y0, y1 = snap["extent"]
    assert -6.24 - 0.24 < x0 < x1 < 6.24 + 0.24 and -0.64 - 0.24 < y0 < y1 < 0.64 + 0.24
    want = obs.block_means(gpu.state.hz.get(), 3, 1)
    assert np.array_equal(snap["data"], want)


@pytest.mark.gpu
def test_save_result_of_a_hip_run(hip, tmp_path):
    from etalon_amd.result import save_result
    from etalon_amd.solver import FDTDSolver

    monitors = [et.DFTMonitor(name="point", center=(0.31, 0.12, 0.0), wavelengths=[1.55]),
                et.DFTMonitor(name="plane", size=(2.0, 1.5, 0.0), wavelengths=[1.55],
                              components=("Hz",))]
    sim = small_sim(monitors=monitors)
    result = FDTDSolver(sim).run()
    meta = save_result(tmp_path, sim, result)
    assert meta["backend"] == "hip" and meta["cells_per_s"] > 0
    with np.load(tmp_path / "result.npz") as saved:
        assert saved["m1.Hz"].shape[2] == 1 and saved["m0.Hz"].shape == (1, 1, 1)


@pytest.mark.gpu
@pytest.mark.parametrize("fused", [True, False])
def test_measure_throughput(hip, fused):
    from etalon_amd.throughput import measure_throughput

    r = measure_throughput(region_cells=500, steps=60, warmup=6, fused=fused)
    assert r.cells == (524, 524) and r.steps == 60 and r.backend == "hip"
    assert r.cell_updates_per_s == pytest.approx(524 * 524 * 60 / r.wall_time_s)
    assert r.cell_updates_per_s > 1e8
