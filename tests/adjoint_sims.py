"""Small design problems for the adjoint tests: a strip waveguide into a square
design region of random permittivity, an Hz source upstream, an Hz line monitor
downstream; J = Σ_f Σ |Hz(f, y)|² on the monitor."""

import numpy as np

import etalon_amd as et

OXIDE, SILICON = 2.085, 12.11


def design_problem(precision="double", seed=0, pixels=(6, 6), steps=240, size=(3.0, 2.0),
                   dl=0.05, pml=8, region=0.6):
    rng = np.random.default_rng(seed)
    eps = rng.uniform(OXIDE, SILICON, size=pixels)
    oxide = et.Medium(n=float(np.sqrt(OXIDE)))
    si = et.Medium(n=float(np.sqrt(SILICON)))
    sx, sy = size
    design = et.PixelGrid(center=(0.0, 0.0, 0.0), size=(region, region, 0.0), permittivity=eps)
    from etalon_amd.grid import courant_time_step

    dt = courant_time_step(0.99, dl, dl)
    sim = et.Simulation(
        size=(sx, sy, 0.0), grid=et.UniformGrid(dl=dl), medium=oxide,
        structures=[et.Box(center=(-sx / 2, 0.0, 0.0), size=(sx, 0.5, 0.0), medium=si),
                    et.Box(center=(sx / 2, 0.0, 0.0), size=(sx, 0.5, 0.0), medium=si), design],
        boundary=et.PML(layers=pml),
        sources=[et.PointSource(center=(-sx / 2 + 0.4, 0.0, 0.0), component="Hz",
                                pulse=et.GaussianPulse(wavelength=1.55, bandwidth=0.3))],
        monitors=[et.DFTMonitor(name="out", center=(sx / 2 - 0.5, 0.0, 0.0), size=(0.0, 1.0, 0.0),
                                wavelengths=np.linspace(1.50, 1.60, 5), components=("Hz",))],
        run_time=(steps - 0.5) * dt, precision=precision, backend="hip")
    return sim, design


def with_design(sim, design, eps):
    """The same simulation with the design region's permittivity replaced."""
    new = design.model_copy(update={"permittivity": np.asarray(eps)})
    structures = tuple(new if s is design else s for s in sim.structures)
    return sim.model_copy(update={"structures": structures}), new
