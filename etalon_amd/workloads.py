"""Synthetic workloads of the benchmarks: the throughput grid and the batch of
uneven simulations.  Random geometry is drawn from ``numpy.random.default_rng(seed)``.
"""

from __future__ import annotations

import numpy as np

from etalon_amd import models as et

SILICON = 12.11
OXIDE = 2.085
WIDTH = 0.5   # strip width, µm
GAP = 0.2     # ring-to-bus gap, µm
DL = 0.05     # grid step, µm


def throughput_sim(precision: str = "single", interior: int = 2000) -> et.Simulation:
    """``interior`` x ``interior`` cells at dl = 0.05 µm + 12-layer CPML (2000: 2024 x
    2024 cells; 8000: 8024 x 8024, whose fields no longer fit the GPU's caches): a
    silicon strip (ε 12.11, 0.5 µm) along x through y = 0 into the PML, oxide cladding
    (ε 2.085), an Hz Gaussian source (1.55 µm, bandwidth 0.1 µm) 2 µm from the left
    edge in the strip, a 4 µm Hz line monitor across the guide 5 µm from the right edge
    (x = +45 µm at 2000) at 21 wavelengths 1.50–1.60 µm, courant 0.99."""
    oxide = et.Medium(n=float(np.sqrt(OXIDE)))
    silicon = et.Medium(n=float(np.sqrt(SILICON)))
    size = round(interior * DL, 9)
    return et.Simulation(
        size=(size, size, 0.0),
        grid=et.UniformGrid(dl=DL),
        medium=oxide,
        structures=[et.Box(center=(0.0, 0.0, 0.0), size=(2 * size, WIDTH, 0.0), medium=silicon)],
        boundary=et.PML(layers=12),
        sources=[et.PointSource(center=(-size / 2 + 2.0, 0.0, 0.0), component="Hz",
                                pulse=et.GaussianPulse(wavelength=1.55, bandwidth=0.1))],
        monitors=[et.DFTMonitor(name="line", center=(size / 2 - 5.0, 0.0, 0.0),
                                size=(0.0, 4.0, 0.0),
                                wavelengths=np.linspace(1.50, 1.60, 21), components=("Hz",))],
        run_time=1.0,
        courant=0.99,
        precision=precision,
        backend="hip",
    )


def _run_time_for(steps: int, courant: float = 0.99) -> float:
    """A run_time that gives exactly ``steps`` steps on a DL grid."""
    from etalon_amd.grid import courant_time_step

    dt = courant_time_step(courant, DL, DL)
    return (steps - 0.5) * dt


def uneven_batch(n: int = 64, seed: int = 0, lo: int = 200, hi: int = 1500, steps: int = 4000,
                 precision: str = "single") -> list:
    """``n`` independent simulations with interior sizes drawn uniformly from
    [lo, hi] cells per side (independently for x and y) at dl = 0.05 µm, 12-layer
    CPML, each holding seeded geometry in oxide (ε 2.085): a silicon (ε 12.11) strip
    bend of radius 1–5 µm or a ring resonator of radius 2–6 µm with a 0.2 µm gap to
    its bus (radii clipped to fit the grid, which needs lo >= 80), an Hz Gaussian
    source at the input end
    of the guide, and 1–8 point or line DFT monitors at 11 wavelengths 1.50–1.60 µm
    (all three components), ``steps`` steps each."""
    rng = np.random.default_rng(seed)
    oxide = et.Medium(n=float(np.sqrt(OXIDE)))
    silicon = et.Medium(n=float(np.sqrt(SILICON)))
    wavelengths = np.linspace(1.50, 1.60, 11)
    run_time = _run_time_for(steps)
    sims = []
    for k in range(n):
        nxi, nyi = (int(v) for v in rng.integers(lo, hi + 1, size=2))
        sx, sy = nxi * DL, nyi * DL
        half = 0.5 * min(sx, sy)
        structures: list = []
        if rng.random() < 0.5:
            # bend: a strip from the left edge along y = -R, a quarter arc about the
            # origin, a strip up to the top edge along x = +R
            r = float(np.clip(rng.uniform(1.0, 5.0), 0.4, max(0.4, half - 0.6)))
            structures += [
                et.Box(center=(-0.5 * sx, -r, 0.0), size=(sx, WIDTH, 0.0), medium=silicon),
                et.Ring(center=(0.0, 0.0, 0.0), inner_radius=r - 0.5 * WIDTH,
                        outer_radius=r + 0.5 * WIDTH, angle_start=270.0, angle_stop=360.0,
                        medium=silicon),
                et.Box(center=(r, 0.5 * sy, 0.0), size=(WIDTH, sy, 0.0), medium=silicon),
            ]
            guide_y = -r
        else:
            # ring resonator above a bus along y = -(R + w + gap)
            r = float(np.clip(rng.uniform(2.0, 6.0), 0.4, max(0.4, half - 1.2)))
            bus = -(r + WIDTH + GAP)
            structures += [
                et.Ring(center=(0.0, 0.0, 0.0), inner_radius=r - 0.5 * WIDTH,
                        outer_radius=r + 0.5 * WIDTH, medium=silicon),
                et.Box(center=(0.0, bus, 0.0), size=(2.0 * sx, WIDTH, 0.0), medium=silicon),
            ]
            guide_y = bus
        source = et.PointSource(center=(-0.5 * sx + 1.0, guide_y, 0.0), component="Hz",
                                pulse=et.GaussianPulse(wavelength=1.55, bandwidth=0.1))
        monitors = []
        for m in range(int(rng.integers(1, 9))):
            x = float(rng.uniform(-0.45, 0.45) * sx)
            if rng.random() < 0.5:
                y = float(rng.uniform(-0.45, 0.45) * sy)
                monitors.append(et.DFTMonitor(name=f"p{m}", center=(x, y, 0.0),
                                              wavelengths=wavelengths))
            else:
                length = float(min(rng.uniform(1.0, 3.0), 0.5 * sy))
                reach = 0.5 * sy - 0.5 * length - 0.1  # the line stays inside the region
                y = float(rng.uniform(-reach, reach))
                monitors.append(et.DFTMonitor(name=f"l{m}", center=(x, y, 0.0),
                                              size=(0.0, length, 0.0), wavelengths=wavelengths))
        sims.append(et.Simulation(size=(sx, sy, 0.0), grid=et.UniformGrid(dl=DL), medium=oxide,
                                  structures=structures, boundary=et.PML(layers=12),
                                  sources=[source], monitors=monitors, run_time=run_time,
                                  courant=0.99, precision=precision, backend="hip"))
    return sims


def adjoint_design_sim(precision: str = "single", seed: int = 0, steps: int = 4000,
                       interior: int = 400, region: float = 2.0, pixels: int = 40):
    """``interior`` x ``interior`` cells at dl = 0.05 µm + 12-layer CPML, an input
    silicon strip (0.5 µm) from the left face and an output strip to the right face,
    a ``region`` µm square design region of ``pixels`` x ``pixels`` pixels between
    them with seeded permittivity in [2.085, 12.11], an Hz source in the input strip
    and a 2 µm Hz line monitor across the output strip at 5 wavelengths 1.50–1.60
    µm; ``steps`` steps.  Returns (simulation, design PixelGrid, objective names)."""
    rng = np.random.default_rng(seed)
    oxide = et.Medium(n=float(np.sqrt(OXIDE)))
    silicon = et.Medium(n=float(np.sqrt(SILICON)))
    size = interior * DL
    half = size / 2.0
    eps = rng.uniform(OXIDE, SILICON, size=(pixels, pixels))
    design = et.PixelGrid(center=(0.0, 0.0, 0.0), size=(region, region, 0.0), permittivity=eps)
    strip_len = half - region / 2.0
    sim = et.Simulation(
        size=(size, size, 0.0), grid=et.UniformGrid(dl=DL), medium=oxide,
        structures=[
            et.Box(center=(-half + strip_len / 2.0, 0.0, 0.0), size=(strip_len + 1.0, WIDTH, 0.0),
                   medium=silicon),
            et.Box(center=(half - strip_len / 2.0, 0.0, 0.0), size=(strip_len + 1.0, WIDTH, 0.0),
                   medium=silicon),
            design,
        ],
        boundary=et.PML(layers=12),
        sources=[et.PointSource(center=(-half + 1.0, 0.0, 0.0), component="Hz",
                                pulse=et.GaussianPulse(wavelength=1.55, bandwidth=0.1))],
        monitors=[et.DFTMonitor(name="out", center=(half - 2.0, 0.0, 0.0), size=(0.0, 2.0, 0.0),
                                wavelengths=np.linspace(1.50, 1.60, 5), components=("Hz",))],
        run_time=_run_time_for(steps), courant=0.99, precision=precision, backend="hip")
    return sim, design, ("out",)


LATTICE = 0.42       # photonic-crystal lattice constant, µm
HOLE_RADIUS = 0.126  # air holes, µm


def phc_holes(half: float, lattice: float = LATTICE, radius: float = HOLE_RADIUS,
              keep_out: float = 0.0) -> list:
    """Air holes of a triangular lattice (rows along x, spacing ``lattice``, every
    other row shifted by half a period, rows ``lattice`` √3/2 apart) covering
    |x|, |y| <= half, without the row at y = 0 (a W1 line-defect waveguide) and
    without holes whose centre lies in the square |x|, |y| < keep_out."""
    air = et.Medium(n=1.0)
    dy = lattice * np.sqrt(3.0) / 2.0
    n_rows = int(np.ceil(half / dy))
    n_cols = int(np.ceil(half / lattice)) + 1
    holes = []
    for row in range(-n_rows, n_rows + 1):
        if row == 0:
            continue
        y = row * dy
        shift = 0.5 * lattice if row % 2 else 0.0
        for col in range(-n_cols, n_cols + 1):
            x = col * lattice + shift
            if abs(x) > half or abs(y) > half:
                continue
            if abs(x) < keep_out and abs(y) < keep_out:
                continue
            holes.append(et.Cylinder(center=(float(x), float(y), 0.0), radius=radius,
                                     medium=air))
    return holes


def phc_adjoint_candidates(n: int = 32, seed: int = 0, steps: int = 8000, interior: int = 1000,
                           region: float = 4.0, pixels: int = 80,
                           precision: str = "single") -> tuple[list, list, tuple]:
    """``n`` design candidates on ``interior`` x ``interior`` cells at dl = 0.05 µm +
    12-layer CPML: a silicon slab (ε 12.11) with a triangular lattice of air holes
    (a = 0.42 µm, r = 0.126 µm) and a W1 line defect along y = 0, a ``region`` µm
    square design region (``pixels`` x ``pixels``, seeded ε in [2.085, 12.11], one
    draw per candidate) at the centre, an Hz source in the defect 4 µm before the
    region, a 3 µm Hz line monitor across the defect 4 µm after it at 9 wavelengths
    1.50–1.60 µm; ``steps`` steps.  Returns (simulations, design
    permittivities, objective names); the candidates share every structure but the
    design region."""
    rng = np.random.default_rng(seed)
    size = interior * DL
    half = size / 2.0
    silicon = et.Medium(n=float(np.sqrt(SILICON)))
    base_design = et.PixelGrid(center=(0.0, 0.0, 0.0), size=(region, region, 0.0),
                               permittivity=np.full((pixels, pixels), SILICON))
    base = et.Simulation(
        size=(size, size, 0.0), grid=et.UniformGrid(dl=DL), medium=silicon,
        structures=[*phc_holes(half + 12 * DL, keep_out=region / 2.0), base_design],
        boundary=et.PML(layers=12),
        sources=[et.PointSource(center=(-region / 2.0 - 4.0, 0.0, 0.0), component="Hz",
                                pulse=et.GaussianPulse(wavelength=1.55, bandwidth=0.15))],
        monitors=[et.DFTMonitor(name="out", center=(region / 2.0 + 4.0, 0.0, 0.0),
                                size=(0.0, 3.0, 0.0),
                                wavelengths=np.linspace(1.50, 1.60, 9), components=("Hz",))],
        run_time=_run_time_for(steps), courant=0.99, precision=precision, backend="hip")
    sims, eps = [], []
    for _ in range(n):
        e = rng.uniform(OXIDE, SILICON, size=(pixels, pixels))
        design = base_design.model_copy(update={"permittivity": e})
        sims.append(base.model_copy(update={"structures": (*base.structures[:-1], design)}))
        eps.append(e)
    return sims, eps, ("out",)
