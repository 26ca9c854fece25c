"""Simulations shared by the tests (the reference's tests/sims.py:12-24 and the GPU
equality suite's full_sim, tests/gpu/test_cupy_equality.py:113-160)."""

import numpy as np

import etalon_amd as et


def pulse(wavelength=1.55, bandwidth=0.5, **kwargs):
    return et.GaussianPulse(wavelength=wavelength, bandwidth=bandwidth, **kwargs)


def small_sim(**overrides):
    """2 µm × 1.5 µm vacuum, 0.05 µm cells, 8-layer CPML, one Hz source, one point probe."""
    kwargs = {
        "size": (2.0, 1.5, 0.0),
        "grid": et.UniformGrid(dl=0.05),
        "boundary": et.PML(layers=8),
        "sources": [et.PointSource(center=(-0.5, 0.1, 0.0), component="Hz", pulse=pulse())],
        "monitors": [et.DFTMonitor(name="probe", center=(0.5, 0.0, 0.0), wavelengths=[1.5, 1.6])],
        "run_time": 20e-15,
        "backend": "hip",
    }
    kwargs.update(overrides)
    return et.Simulation(**kwargs)


def full_sim(precision="single", run_time=1.0):
    """Dielectric box, off-grid Ey and Hz sources, point and plane monitors, κ_max = 2."""
    si = et.Medium(n=3.476)
    return et.Simulation(
        size=(131 * 0.03, 97 * 0.03, 0.0),
        grid=et.UniformGrid(dl=0.03),
        medium=et.Medium(n=1.444),
        structures=[et.Box(center=(0.3, -0.2, 0.0), size=(1.5, 0.5, 0.22), medium=si)],
        boundary=et.PML(layers=10, kappa_max=2.0),
        sources=[
            et.PointSource(center=(-1.234, 0.071, 0.0), component="Ey",
                           pulse=et.GaussianPulse(wavelength=1.55, bandwidth=0.3)),
            et.PointSource(center=(0.517, 0.633, 0.0), component="Hz",
                           pulse=et.GaussianPulse(wavelength=1.31, bandwidth=0.2, amplitude=0.5)),
        ],
        monitors=[
            et.DFTMonitor(name="point", center=(1.11, -0.37, 0.0),
                          wavelengths=np.linspace(1.3, 1.7, 9)),
            et.DFTMonitor(name="plane", size=(3.0, 2.0, 0.0), wavelengths=[1.31, 1.55]),
        ],
        run_time=run_time,
        precision=precision,
        backend="hip",
    )


def vacuum_pulse(precision="single"):
    """8 × 8 µm vacuum at 0.05 µm (160² + 12-layer CPML = 184²), courant 0.5, Hz source
    at (−3, 0), pulse 1.55 ± 0.25 µm, 4 point monitors at 11 wavelengths, 2000 steps."""
    wl = np.linspace(1.45, 1.65, 11)
    probes = [(-1.0, 0.0), (3.0, 0.0), (0.0, 2.0), (2.0, -2.0)]
    sim = et.Simulation(
        size=(8.0, 8.0, 0.0),
        grid=et.UniformGrid(dl=0.05),
        boundary=et.PML(layers=12),
        sources=[et.PointSource(center=(-3.0, 0.0, 0.0), component="Hz",
                                pulse=et.GaussianPulse(wavelength=1.55, bandwidth=0.5))],
        monitors=[et.DFTMonitor(name=f"p{k}", center=(x, y, 0.0), wavelengths=wl)
                  for k, (x, y) in enumerate(probes)],
        run_time=1.0,
        courant=0.5,
        precision=precision,
        backend="hip",
    )
    return sim.model_copy(update={"run_time": 2000 * _dt(sim)})


def _dt(sim):
    from etalon_amd.grid import build_yee_grid

    return build_yee_grid(sim).dt


def bitwise_equal(a, b) -> bool:
    """Equal bit patterns (distinguishes ±0), for real or complex arrays."""
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    if a.shape != b.shape or a.dtype != b.dtype:
        return False
    view = {4: np.uint32, 8: np.uint64, 16: np.uint64}[a.dtype.itemsize]
    return bool(np.array_equal(a.view(view), b.view(view)))
