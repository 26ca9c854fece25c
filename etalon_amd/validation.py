"""M0 physics validation on the HIP path (etalon/validation/m0.py:1-257) and the
point2d convention check against Meep (bench/etalon/run_device.py:120-149).

Each function runs small simulations through this package's solver and returns
plain NumPy data; the thresholds are in tests/test_gpu_m0.py (the reference's
tests/validation/test_m0.py:17-72 and bench/metrics.py:20).
"""

from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
from scipy.special import hankel1

from etalon_amd import models as et
from etalon_amd.analysis import measured_phase_velocity, numerical_phase_velocity
from etalon_amd.constants import C_0, EPSILON_0, MU_0, UM
from etalon_amd.solver import FDTDSolver
from etalon_amd.sources import pulse_signal

# A broad pulse covering 9-36 cells per wavelength on 0.1 / 0.05 µm grids.
BROADBAND = et.GaussianPulse(wavelength=1.35, bandwidth=0.9)
BROADBAND_WAVELENGTHS = np.linspace(0.9, 1.8, 19)


def centred_size(half_width: float, dl: float) -> float:
    """A region size with an odd number of cells: the origin is then an Hz node."""
    return (2 * math.ceil(half_width / dl) + 1) * dl


@dataclass
class DispersionCase:
    dl: float
    angle: float
    n: float
    wavelengths: np.ndarray
    cells_per_wavelength: np.ndarray  # λ / (n Δ)
    v_measured: np.ndarray            # ṽp / (c/n), from the probes
    v_theory: np.ndarray              # ṽp / (c/n), Yee relation


def dispersion_cases(dl: float, n: float = 1.0, run_time: float = 200e-15, r_near: float = 1.5,
                     r_far: float = 4.5, pulse=BROADBAND, wavelengths=BROADBAND_WAVELENGTHS,
                     precision: str = "single") -> list[DispersionCase]:
    """One run: an Hz point source on an Hz node, Hz probes on Hz nodes at two radii
    along 0° and along 45° (lattice symmetry directions, where rays follow k̃)."""
    m_axis = (round(r_near / dl), round(r_far / dl))
    m_diag = (round(r_near / (math.sqrt(2) * dl)), round(r_far / (math.sqrt(2) * dl)))
    probes = {
        0.0: [(m * dl, 0.0, m * dl) for m in m_axis],
        45.0: [(m * dl, m * dl, math.sqrt(2) * m * dl) for m in m_diag],
    }
    monitors = [et.DFTMonitor(name=f"{angle:g}-{k}", center=(x, y, 0.0), wavelengths=wavelengths,
                              components=("Hz",))
                for angle, points in probes.items() for k, (x, y, _) in enumerate(points)]
    size = centred_size(r_far + 1.0, dl)
    sim = et.Simulation(size=(size, size, 0.0), medium=et.Medium(n=n), grid=et.UniformGrid(dl=dl),
                        boundary=et.PML(layers=12),
                        sources=[et.PointSource(center=(0.0, 0.0, 0.0), component="Hz",
                                                pulse=pulse)],
                        monitors=monitors, run_time=run_time, precision=precision, backend="hip")
    result = sim.run()
    cases = []
    for angle, points in probes.items():
        (_, _, r1), (_, _, r2) = points
        v_meas = measured_phase_velocity(result[f"{angle:g}-0"]["Hz"], result[f"{angle:g}-1"]["Hz"],
                                         r1, r2, wave="cylindrical", n=n)
        v_yee = numerical_phase_velocity(wavelengths, result.grid, angle=angle, n=n)
        lam = np.asarray(wavelengths, dtype=np.float64)
        cases.append(DispersionCase(dl, angle, n, lam, lam / (n * dl), v_meas, v_yee))
    return cases


@dataclass
class ReflectionCase:
    t: np.ndarray
    probes: list[str]
    hz_small: np.ndarray      # (n_t, n_probes)
    hz_reference: np.ndarray  # same probes in a domain too large to reflect in time
    error: np.ndarray         # max_t |Hz − Hz_ref| / max_t |Hz_ref| per probe


def cpml_reflection_case(layers: int = 12, run_time: float = 100e-15,
                         precision: str = "single") -> ReflectionCase:
    """CPML reflection (Taflove & Hagness, ch. 7): Hz(t) two cells inside the region,
    beside a face and at a corner, against a domain large enough that no reflection
    returns within the run."""
    dl = 0.05
    size = centred_size(1.5, dl)
    edge = size / 2.0
    probe_xy = {"side": (edge - 2.5 * dl, 0.0), "corner": (edge - 2.5 * dl, edge - 2.5 * dl)}
    source = et.PointSource(center=(0.0, 0.0, 0.0), component="Hz",
                            pulse=et.GaussianPulse(wavelength=1.55, bandwidth=0.5))
    margin = C_0 * run_time / 2.0  # µm

    def record(region: float):
        sim = et.Simulation(size=(region, region, 0.0), grid=et.UniformGrid(dl=dl),
                            boundary=et.PML(layers=layers), sources=[source], run_time=run_time,
                            precision=precision, backend="hip")
        solver = FDTDSolver(sim)
        x, y = solver.grid.coords("Hz")
        nodes = [(int(np.argmin(np.abs(x - px))), int(np.argmin(np.abs(y - py))))
                 for px, py in probe_xy.values()]
        trace = np.empty((solver.n_steps, len(nodes)))
        for step in range(solver.n_steps):
            solver.step()
            for k, (i, j) in enumerate(nodes):
                trace[step, k] = float(solver.state.hz[i, j])
        t = (np.arange(solver.n_steps) + 0.5) * solver.grid.dt
        return t, trace

    t, small = record(size)
    _, reference = record(centred_size(1.5 + margin, dl))
    error = np.abs(small - reference).max(0) / np.abs(reference).max(0)
    return ReflectionCase(t, list(probe_xy), small, reference, error)


@dataclass
class LateTimeCase:
    steps: np.ndarray
    energy: np.ndarray  # total field energy per unit length (J/m), PML included


def late_time_case(n_steps: int = 30_000, precision: str = "single") -> LateTimeCase:
    """Run long after the pulse has left the domain: the CPML must keep absorbing.

    The energy is sampled after steps 1, 51, 101, … (the reference samples after
    every step n with n % 50 == 0, counting from 0); the fused path advances
    between samples."""
    dl = 0.05
    size = centred_size(1.0, dl)
    sim = et.Simulation(size=(size, size, 0.0), grid=et.UniformGrid(dl=dl),
                        boundary=et.PML(layers=12),
                        sources=[et.PointSource(center=(0.0, 0.0, 0.0), component="Hz",
                                                pulse=et.GaussianPulse(wavelength=1.55,
                                                                       bandwidth=0.5))],
                        run_time=1e-15, precision=precision, backend="hip")
    solver = FDTDSolver(sim)
    area = (solver.grid.dx * UM) * (solver.grid.dy * UM)
    steps, energy = [], []
    for n in range(0, n_steps, 50):
        solver.advance(n + 1 - solver.n)
        f = solver.fields()
        w = 0.5 * EPSILON_0 * (np.sum(f["Ex"].astype(float) ** 2) + np.sum(f["Ey"].astype(float) ** 2))
        w += 0.5 * MU_0 * np.sum(f["Hz"].astype(float) ** 2)
        steps.append(n)
        energy.append(w * area)
    return LateTimeCase(np.array(steps), np.array(energy))


@dataclass
class GreenCase:
    wavelengths: np.ndarray
    radius: float
    hz_measured: np.ndarray
    hz_analytic: np.ndarray


def green_function_case(dl: float = 0.05, radius: float = 3.0,
                        precision: str = "single") -> GreenCase:
    """Hz(r, ω) = −(ωε0/4) K̂(ω) H0⁽¹⁾(k̃ r) for a magnetic line current M_z = K(t) δ²(r)
    (time convention e^(−iωt)); K̂ is the DFT of the injected samples K(nΔt) with the
    monitor's Δt, k̃ the Yee wavenumber along x.  Checks the source normalisation
    1/(ΔxΔy), the DFT's Δt and every sign together."""
    wavelengths = np.linspace(1.3, 1.8, 11)
    pulse = et.GaussianPulse(wavelength=1.55, bandwidth=0.5, amplitude=1.0)
    size = centred_size(radius + 1.5, dl)
    m = round(radius / dl)
    sim = et.Simulation(size=(size, size, 0.0), grid=et.UniformGrid(dl=dl),
                        boundary=et.PML(layers=12),
                        sources=[et.PointSource(center=(0.0, 0.0, 0.0), component="Hz",
                                                pulse=pulse)],
                        monitors=[et.DFTMonitor(name="probe", center=(m * dl, 0.0, 0.0),
                                                wavelengths=wavelengths, components=("Hz",))],
                        run_time=250e-15, precision=precision, backend="hip")
    result = sim.run()
    grid = result.grid
    t = np.arange(result.metadata["n_steps"]) * grid.dt  # M is sampled at nΔt
    omega = 2 * np.pi * C_0 / wavelengths
    k_hat = (pulse_signal(pulse, t)[None, :] * np.exp(1j * omega[:, None] * t[None, :])).sum(1)
    k_hat *= grid.dt
    k_yee = 2 * np.pi / (wavelengths * UM) / numerical_phase_velocity(wavelengths, grid, angle=0.0)
    analytic = -(omega * EPSILON_0 / 4.0) * k_hat * hankel1(0, k_yee * m * dl * UM)
    measured = np.asarray(result["probe"]["Hz"].values, dtype=np.complex128)
    return GreenCase(wavelengths, m * dl, measured, analytic)


# point2d (bench/config.yaml:64-70, bench/config.py:75-100,200-215,390-401): SiO2,
# 8 x 4 µm region at resolution 50, 1 µm PML = 50 layers, Meep S = 0.5 → courant
# 0.5·√2, Hz source at (−2.99, 0.01), probes at (−1.99, 0.01) and (2.01, 0.01),
# band 1.50–1.60 µm (51 points), 300 fs, single precision.
POINT2D_BAND = (1.50, 1.60, 51)


def point2d_simulation(precision: str = "single"):
    lo, hi, points = POINT2D_BAND
    return et.Simulation(
        center=(0.0, 0.0, 0.0), size=(8.0, 4.0, 0.0), medium=et.Medium(n=1.444),
        grid=et.UniformGrid(dl=1.0 / 50), boundary=et.PML(layers=50),
        sources=[et.PointSource(center=(-2.99, 0.01, 0.0), component="Hz",
                                pulse=et.GaussianPulse(wavelength=0.5 * (lo + hi),
                                                       bandwidth=hi - lo))],
        monitors=[et.DFTMonitor(name=f"probe{k}", center=(x, y, 0.0),
                                wavelengths=np.linspace(lo, hi, points), components=("Hz",))
                  for k, (x, y) in enumerate([(-1.99, 0.01), (2.01, 0.01)], start=1)],
        run_time=300e-15, courant=0.5 * math.sqrt(2), precision=precision, backend="hip")


def point2d_t21(precision: str = "single") -> tuple[np.ndarray, np.ndarray, int]:
    """(wavelengths, T21 = Hz(probe 2) / Hz(probe 1), steps) of the point2d device."""
    result = point2d_simulation(precision).run()
    h1 = np.asarray(result["probe1"]["Hz"].values)
    h2 = np.asarray(result["probe2"]["Hz"].values)
    lo, hi, points = POINT2D_BAND
    return np.linspace(lo, hi, points), h2 / h1, int(result.metadata["n_steps"])
