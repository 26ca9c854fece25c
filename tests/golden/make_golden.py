"""Write the golden vectors in tests/golden/*.npz from the reference's own functions.

Run in a container where the reference tree is mounted (it never travels with the
repository; the vectors do):

    python tests/golden/make_golden.py /root/reference

Only the reference's NumPy-only modules are imported (grid, materials, cpml, dft,
sources, state, the NumPy backend, the models); its solver module needs xarray,
which is absent, so the time loop below is the documented step order of
etalon/fdtd/solver.py:120-128 written out over the reference's own functions.
What is captured (float64 inputs/outputs, plus float32 runs):

* cpml_*.npz      — cpml_axis profiles for the 184-cell (courant 0.5) and the
                    2024-cell (courant 0.99, ε_bg 2.085) axes and a κ_max = 2 case;
* inject.npz      — point_injection nodes/weights/coefficients and 40 steps of
                    delta(n) for an on-node Hz source, a mid-cell Hz source and an
                    off-grid Ey source in a dielectric;
* sampling.npz    — monitor_samplings flat indices/weights for point, line and
                    plane monitors, and dft_phasors for a set of times;
* run_small_*.npz — 50 steps of the reference small simulation (2 × 1.5 µm, 8-layer
                    CPML, Hz source, point probe) in float64 and float32: fields,
                    ψ and DFT accumulators after step 50.
"""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

OUT = Path(__file__).resolve().parent


def main(ref_root: str) -> None:
    sys.path.insert(0, ref_root)
    import etalon as et  # the reference package (lazy top-level names)
    from etalon.fdtd.backends.numpy_backend import NumpyBackend
    from etalon.fdtd.cpml import cpml_axis
    from etalon.fdtd.dft import dft_phasors, monitor_samplings
    from etalon.fdtd.sources import point_injection
    from etalon.fdtd.state import build_te_state
    from etalon.grid.materials import permittivity_on_nodes
    from etalon.grid.yee import build_yee_grid, courant_time_step

    # ---------------------------------------------------------------- CPML
    dt1 = courant_time_step(0.5, 0.05, 0.05)
    dt2 = courant_time_step(0.99, 0.05, 0.05)
    cases = {
        "cpml_184": cpml_axis(184, 0.05, dt1, et.PML(), 1.0),
        "cpml_2024": cpml_axis(2024, 0.05, dt2, et.PML(), 2.085),
        "cpml_k2": cpml_axis(151, 0.03, courant_time_step(0.99, 0.03, 0.03),
                             et.PML(layers=10, kappa_max=2.0), 1.444**2),
    }
    for name, ax in cases.items():
        np.savez(OUT / f"{name}.npz", b_e=ax.b_e, c_e=ax.c_e, inv_kappa_e=ax.inv_kappa_e,
                 b_h=ax.b_h, c_h=ax.c_h, inv_kappa_h=ax.inv_kappa_h,
                 dt=np.float64(dt1 if name == "cpml_184" else dt2 if name == "cpml_2024"
                               else courant_time_step(0.99, 0.03, 0.03)))

    # ---------------------------------------------------------------- sims
    def small_sim(**overrides):
        kwargs = {
            "size": (2.0, 1.5, 0.0),
            "grid": et.UniformGrid(dl=0.05),
            "boundary": et.PML(layers=8),
            "sources": [et.PointSource(center=(-0.5, 0.1, 0.0), component="Hz",
                                       pulse=et.GaussianPulse(wavelength=1.55, bandwidth=0.5))],
            "monitors": [et.DFTMonitor(name="probe", center=(0.5, 0.0, 0.0),
                                       wavelengths=[1.5, 1.6])],
            "run_time": 20e-15,
            "backend": "numpy",
        }
        kwargs.update(overrides)
        return et.Simulation(**kwargs)

    # ---------------------------------------------------------------- injection
    vac = et.Simulation(size=(8.0, 8.0, 0.0), grid=et.UniformGrid(dl=0.05),
                        boundary=et.PML(layers=12), courant=0.5, run_time=1e-13,
                        sources=[et.PointSource(center=(-3.0, 0.0, 0.0), component="Hz",
                                                pulse=et.GaussianPulse(wavelength=1.55,
                                                                       bandwidth=0.5))],
                        backend="numpy")
    diel = small_sim(medium=et.Medium(n=1.5),
                     structures=[et.Box(center=(0.0, 0.0, 0.0), size=(0.4, 0.3, 0.0),
                                        medium=et.Medium(n=3.476))],
                     sources=[et.PointSource(center=(0.013, -0.031, 0.0), component="Ey",
                                             pulse=et.GaussianPulse(wavelength=1.31,
                                                                    bandwidth=0.2,
                                                                    amplitude=0.5))])
    inj_data = {}
    for tag, sim in (("vac", vac), ("diel", diel), ("small", small_sim())):
        grid = build_yee_grid(sim)
        eps = permittivity_on_nodes(sim, grid)
        inj = point_injection(sim.sources[0], grid, eps)
        inj_data[f"{tag}_i"] = inj.i
        inj_data[f"{tag}_j"] = inj.j
        inj_data[f"{tag}_weights"] = inj.weights
        inj_data[f"{tag}_coefficient"] = inj.coefficient
        inj_data[f"{tag}_delta"] = np.array([inj.delta(n, grid.dt) for n in range(40)])
        if tag == "diel":
            inj_data[f"{tag}_eps_ex"] = eps["Ex"]
            inj_data[f"{tag}_eps_ey"] = eps["Ey"]
    np.savez(OUT / "inject.npz", **inj_data)

    # ---------------------------------------------------------------- sampling
    mons = [
        et.DFTMonitor(name="point", center=(0.31, 0.12, 0.0), wavelengths=[1.5, 1.55, 1.6]),
        et.DFTMonitor(name="line", center=(0.0, 0.1, 0.0), size=(0.6, 0.0, 0.0),
                      wavelengths=[1.55]),
        et.DFTMonitor(name="plane", size=(1.0, 0.6, 0.0), wavelengths=[1.31, 1.55],
                      components=("Hz",)),
    ]
    sim = small_sim(monitors=mons)
    grid = build_yee_grid(sim)
    samp = {}
    for mon in mons:
        for s in monitor_samplings(mon, grid):
            key = f"{mon.name}_{s.component}"
            samp[f"{key}_flat"] = s.flat_index
            samp[f"{key}_xw"] = s.x.weights if s.x.weights is not None else np.zeros(0)
            samp[f"{key}_yw"] = s.y.weights if s.y.weights is not None else np.zeros(0)
    f = et.C_0 / np.linspace(1.3, 1.8, 7)
    times = (np.arange(0, 300, 7, dtype=np.float64) + 0.5) * grid.dt
    samp["phasor_freqs"] = f
    samp["phasor_times"] = times
    samp["phasors"] = dft_phasors(f, times, grid.dt)
    np.savez(OUT / "sampling.npz", **samp)

    # ---------------------------------------------------------------- 50-step runs
    for precision, real, cplx in (("double", np.float64, np.complex128),
                                  ("single", np.float32, np.complex64)):
        sim = small_sim(precision=precision)
        grid = build_yee_grid(sim)
        eps = permittivity_on_nodes(sim, grid)
        bg = sim.medium.permittivity
        cx = cpml_axis(grid.nx, grid.dx, grid.dt, sim.boundary, bg)
        cy = cpml_axis(grid.ny, grid.dy, grid.dt, sim.boundary, bg)
        s = build_te_state(grid, eps, cx, cy, real, lambda a: a)
        backend = NumpyBackend()
        inj = [point_injection(src, grid, eps) for src in sim.sources]
        samplings = [x for mon in sim.monitors for x in monitor_samplings(mon, grid)]
        accs = [np.zeros((x.wavelengths.size, x.flat_index.size), cplx) for x in samplings]
        for n in range(50):
            # etalon/fdtd/solver.py:120-128 with _inject (:83-89) and _phasor_rows (:91-110)
            backend.update_h_te(s)
            for p in inj:
                if p.is_magnetic:
                    getattr(s, p.component.lower())[p.i, p.j] -= p.delta(n, grid.dt).astype(real)
            steps = np.arange(n, n + 1, dtype=np.float64)
            for x, acc in zip(samplings, accs):
                if x.is_magnetic:
                    ph = dft_phasors(x.frequencies, (steps + 0.5) * grid.dt, grid.dt).astype(cplx)
                    backend.accumulate_dft(acc, getattr(s, x.component.lower()), x.flat_index,
                                           ph[0])
            backend.update_e_te(s)
            for p in inj:
                if not p.is_magnetic:
                    getattr(s, p.component.lower())[p.i, p.j] -= p.delta(n, grid.dt).astype(real)
            for x, acc in zip(samplings, accs):
                if not x.is_magnetic:
                    ph = dft_phasors(x.frequencies, (steps + 1.0) * grid.dt, grid.dt).astype(cplx)
                    backend.accumulate_dft(acc, getattr(s, x.component.lower()), x.flat_index,
                                           ph[0])
        out = {name: getattr(s, name) for name in ("ex", "ey", "hz", "psi_ex_y", "psi_ey_x",
                                                   "psi_hz_x", "psi_hz_y", "cb_ex", "cb_ey")}
        for x, acc in zip(samplings, accs):
            out[f"acc_{x.monitor}_{x.component}"] = acc
        np.savez(OUT / f"run_small_{precision}.npz", **out)
    print("wrote", sorted(p.name for p in OUT.glob("*.npz")))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
