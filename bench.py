"""Throughput benchmark of the fused 2D TE FDTD path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision single|double]
                    [--no-cpu-baseline] [--hbm-steps S] [--interior N]
    python bench.py --workload batched [--batch N] [--batch-steps S] [--seed K]
    python bench.py --workload adjoint [--seed K] [--fd-pixels P]
    python bench.py --workload batched-adjoint [--candidates N] [--phc-steps S] [--fd-pixels P]
    python bench.py --workload adjoint|batched-adjoint --one-pass   (one gradient, for profilers)

Workload (the throughput grid): 2000 × 2000 interior cells at dl = 0.05 µm plus a
12-layer CPML on every face (2024 × 2024 cells), a silicon strip waveguide
(ε = 12.11, 0.5 µm wide, along x through y = 0, into the PML) in oxide cladding
(ε = 2.085, set as n = sqrt(2.085)), an Hz Gaussian source (1.55 µm, bandwidth
0.1 µm) 2 µm from the left edge inside the strip, and a line DFT monitor across
the guide at x = +45 µm (4 µm long, Hz, 21 wavelengths 1.50–1.60 µm), courant
0.99.  Fields start at zero on the device; the timed region is K full steps
(fused kernel + monitor DFT), bracketed by device synchronisation.

One JSON line on stdout: value = cell-updates/s over the whole grid (PML
included) ÷ 1e9, with ``timed_region.host_share`` the part of the timed wall that
the GPU's own step time does not cover; a roofline object for the dominant kernel —
the one kernel per step block (6 steps on this grid) that advances every tile
(joint_kernel) — whose frac is the HBM bytes one launch moves (rocprofv3 PMC counts
of this build of the kernel; the compulsory bytes, each array read or written once,
when the table holds none) over its average duration from HIP events on the stream
it runs on, against 8 TB/s, with the compulsory and the "effective" figure (32 B per
cell-update in float32, 64 B in float64, times the cells and the steps a launch
advances; above the peak under temporal blocking) beside it; `hbm_resident`, the
same workload at 8000 x 8000 interior cells, whose fields do not fit the GPU's
caches; summaries of the other precision, the batched workload and the adjoint
workload (``--no-extras`` leaves them out); and the NumPy oracle on the host (one
core) on a bounded sample of the same grid.

``--workload batched``: N (64) simulations of uneven sizes (interior 200–1500
cells per side, seeded bends and rings, 1–8 monitors each, 4000 steps) advanced
together by one autotuned fused plan.  value = aggregate cell-updates/s; also
sims/s, the same simulations run one after another (fused, and with the per-step
primitives on a bounded sample of steps), a bitwise check of every simulation's
fields against its sequential run, and the float32 error of every monitor against
the float64 run of the same batch (float64 on this path is bit-identical to the
NumPy oracle); cpu_baseline: the NumPy oracle over every simulation, one process
per core (at most 16), ``--cpu-batch-steps`` steps each.

``--workload adjoint``: a 424 x 424 grid (400 x 400 interior + 12 CPML) with a
2 x 2 µm design region of 40 x 40 pixels between an input and an output strip,
J = Σ over 5 wavelengths of |DFT Hz|² on the output line monitor, 4000 steps,
float32, checkpoints every 64 steps.  value = wall time of J and ∂J/∂ε over the
wall time of one plain forward run; the gradient is checked against the float64
gradient and float64 central differences on random pixels; the NumPy float64
adjoint is timed on a bounded sample of steps of the same grid.

``--workload batched-adjoint``: N (32) design candidates, each 1000 x 1000
interior + 12 CPML, a triangular-lattice photonic crystal (a = 0.42 µm, air holes
of r = 0.126 µm in ε 12.11) with a W1 line defect and a 4 x 4 µm design region of
80 x 80 pixels, J = Σ over 9 wavelengths of |DFT Hz|² on the output line monitor,
8000 steps, checkpoints every ~√N steps.  One fused plan runs every candidate's
forward pass; the backward sweep runs the recompute and the reverse steps of every
candidate in one joint launch per step block.  value = wall time of every J and
∂J/∂ε over the wall time of one batched plain forward run of the candidates;
candidate 0's gradient is checked against its float64 gradient and float64
central differences.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL = {"single": 32, "double": 64}
# Reference-published throughput on this grid (2024^2 cells incl. CPML), RTX 4060 Ti,
# CuPy kernels: README.md:57, bench/results/b2ff355.json:76-87, docs/plan-m0.md:232.
REFERENCE_GCELLS = {"single": 2.944929933, "double": 1.5}
# Bytes the timed kernel moves per launch, from rocprofv3 PMC runs of this script
# (profiles/fused_2024_pmc_traffic.md, profiles/joint_kernel_8024_hbm.md): FETCH_SIZE and
# WRITE_SIZE in KiB per dispatch, keyed by (precision, vec, temporal, band_rows, launch)
# for the 2024 x 2024 grid and by (precision, temporal, band_rows) at 8024 x 8024, and
# the factors that a 256 MiB copy gives them (FETCH_SIZE counts half the bytes read).
# The counts belong to one build of the kernels: PMC_BUILD is the fingerprint of the
# sources they were measured with (kernel_build_id), and a build with another
# fingerprint reports no measured traffic until the counters are taken again.
PMC_BUILD = "8349b2dbe7a09e00"
PMC_KIB = {("single", 1, 6, 16, 1): (39201.9, 49890.3),
           ("double", 1, 6, 16, 1): (93115.3, 106755.8)}
FETCH_SCALE, WRITE_SCALE = 1.99985, 1.0
PMC_KIB_8024 = {("single", 8, 16): (440256.2, 795735.5),
                ("double", 6, 16): (902089.5, 1563518.7)}
# VALU wave-instructions one launch of the same build issues (rocprofv3 SQ_INSTS_VALU,
# profiles/joint_kernel_sq_counters.md), keyed by (precision, temporal, band_rows, grid
# rows).  A CU issues at most one wave64 float VALU instruction per cycle (four SIMDs of
# 16 lanes, four cycles each, float32 and float64 alike), at up to 2.4 GHz: the count
# over CUs x clock is the launch's VALU issue time, the time it could not go below with
# the instructions it has.
SQ_VALU = {("single", 6, 16, 2024): 12986303, ("double", 6, 16, 2024): 15597430,
           ("single", 8, 16, 8024): 280114133}
VALU_CLOCK_HZ = 2.4e9
# Infinity Cache (MALL) of the MI355X: a launch whose working set fits it has part of
# its L2 misses served there, not by HBM.
MALL_BYTES = 256 * 2**20
# The sources that define the forward joint kernel and its tile work-list.
KERNEL_SOURCES = ("common.h", "fused.h", "te_tile.h", "adj_tile.h", "joint.h", "coplan.h",
                  "joint_fwd.hip", "plan.hip")


def kernel_build_id() -> str:
    """sha256 (16 hex digits) over the sources of the forward joint kernel and its
    tile layout (KERNEL_SOURCES): equal ids, equal kernels and work-lists."""
    import hashlib

    h = hashlib.sha256()
    for name in KERNEL_SOURCES:
        h.update(name.encode())
        h.update((ROOT / "etalon_amd" / "csrc" / name).read_bytes())
    return h.hexdigest()[:16]


def measured_traffic(table: dict, key: tuple) -> "int | None":
    """HBM bytes per launch from the PMC table, or None when the table is of another
    build of the kernels or holds no entry for this shape."""
    pmc = table.get(key)
    if pmc is None or kernel_build_id() != PMC_BUILD:
        return None
    return round((pmc[0] * FETCH_SCALE + pmc[1] * WRITE_SCALE) * 1024)


from etalon_amd.workloads import adjoint_design_sim, throughput_sim, uneven_batch  # noqa: E402


def cpu_baseline(precision: str, sample_steps: int) -> dict:
    """The NumPy oracle on the host, one core, 20 warm-up + ``sample_steps`` timed steps."""
    from oracle import te_numpy as ref

    solver = ref.Solver(throughput_sim(precision))
    for _ in range(20):
        solver.step()
    t0 = time.perf_counter()
    for _ in range(sample_steps):
        solver.step()
    wall = time.perf_counter() - t0
    cells = solver.g.nx * solver.g.ny
    return {
        "value": cells * sample_steps / wall / 1e9,
        "unit": "Gcell-updates/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{sample_steps} steps (after 20 warm-up) of the same {solver.g.nx}x{solver.g.ny} "
                  f"grid on the NumPy oracle, {wall:.1f} s; host {os.cpu_count()} CPUs, "
                  f"{_cpu_model()}",
    }


def _cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def run(precision: str, steps: int, warmup: int, interior: int = 2000) -> dict:
    from etalon_amd.solver import FDTDSolver

    sim = throughput_sim(precision, interior)
    solver = FDTDSolver(sim)
    nx, ny = solver.grid.nx, solver.grid.ny
    solver.advance(warmup)
    solver.backend.synchronize()
    t0 = time.perf_counter()
    solver.advance(steps)
    solver.backend.synchronize()
    wall = time.perf_counter() - t0
    cells = nx * ny
    # Kernel-level timing with HIP events: the same plan, launched directly.
    plan = solver._plan
    n_time = 480  # 5 trace chunks of 96, a multiple of depths 6 and 8: every launch is full
    delta, phasors = solver._tables(solver.n, n_time)
    if delta is not None:
        plan.source_table(solver.n, delta)
    for mon_id, table in zip(solver._plan_ids, phasors):
        if table is not None:
            plan.phasor_table(mon_id, solver.n, table)
    tm = plan.time(solver.n, n_time)
    solver._plan_tables = (solver.n, solver.n + n_time)
    plan.bind()
    solver.n += n_time
    depth = tm["depth"]
    per_cell = BYTES_PER_CELL[precision]
    kernel_bytes = per_cell * tm["kernel_cells"] * depth  # algorithmic, per launch
    gpu_ms_per_step = tm["total_ms"] / (tm["launches"] * depth)
    return {
        "cells": (nx, ny),
        "wall": wall,
        "value": cells * steps / wall / 1e9,
        "ms_per_step": wall / steps * 1e3,
        "timing": tm,
        "kernel_bytes": kernel_bytes,
        "gpu_ms_per_step": gpu_ms_per_step,
        # the share of the timed wall that the GPU's own step time does not cover
        # (launch and table work on the host, the syncs around the region)
        "host_share": max(0.0, 1.0 - steps * gpu_ms_per_step / (wall * 1e3)),
        "launch_gbs": per_cell * cells * depth / (tm["launch_ms"] * 1e-3) / 1e9,
        "tune": plan.resolved_tune().__dict__,
        "device": solver.backend.device,
    }


def roofline(precision: str, r: dict, traffic: "int | None") -> dict:
    """The roofline object of the dominant kernel (the one joint kernel per step block).

    frac = HBM bytes one launch moves over its event-timed duration, against 8 TB/s.
    The bytes are the measured ones (rocprofv3 FETCH_SIZE and WRITE_SIZE of this build
    of the kernels) when the PMC table holds them, else the compulsory ones — every
    array the launch needs read once and every array it updates written once, 32 B per
    cell in float32 and 64 in float64 (24 and 48 in the uniform-material interior tiles,
    which read no cb array), a lower bound on its traffic.  The effective
    figure (the algorithmic bytes of the T unfused steps a launch advances) is kept
    beside it: temporal blocking puts it above the peak."""
    tm = r["timing"]
    kernel_s = tm["kernel_ms"] * 1e-3
    # cells of uniform-material tiles read no cb array: two REALs less per cell
    cb_bytes = 2 * BYTES_PER_CELL[precision] // 8
    compulsory = (BYTES_PER_CELL[precision] * tm["kernel_cells"]
                  - cb_bytes * tm.get("uniform_cells", 0))
    moved = traffic if traffic is not None else compulsory
    gbs = moved / kernel_s / 1e9
    roof = {
        "bound": "hbm",
        "achieved": round(gbs, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(gbs / HBM_PEAK_GBS, 4),
        "traffic": traffic,
        "basis": ("measured: bytes per launch between L2 and memory from rocprofv3 FETCH_SIZE "
                  "and WRITE_SIZE of this build of the kernel (scaled by a copy kernel's "
                  "factors), over the launch's event-timed duration" if traffic is not None else
                  "compulsory: no PMC counts for this build of the kernel, so the bytes are "
                  "the compulsory ones (each array read or written once per launch), a lower "
                  "bound on its traffic"),
        "kernel_build": kernel_build_id(),
        "compulsory_bytes": compulsory,
        "uniform_cells": tm.get("uniform_cells", 0),
        "compulsory_gbs": round(compulsory / kernel_s / 1e9, 1),
        "compulsory_frac": round(compulsory / kernel_s / 1e9 / HBM_PEAK_GBS, 4),
        "effective_gbs": round(r["kernel_bytes"] / kernel_s / 1e9, 1),
        "effective_frac": round(r["kernel_bytes"] / kernel_s / 1e9 / HBM_PEAK_GBS, 4),
        "effective_note": f"{BYTES_PER_CELL[precision]} B per cell-update (Ex, Ey, Hz read and "
                          "written, cb_ex, cb_ey read) x cells x the steps one launch advances, "
                          "over its duration: above the peak because temporal blocking moves "
                          "each array once per launch, not once per step",
    }
    if traffic is not None:
        roof["traffic_over_compulsory"] = round(traffic / compulsory, 3)
    valu = _valu_issue(precision, r, kernel_s)
    if valu is not None:
        roof["valu_issue"] = valu
    # both buffers of Ex, Ey and Hz, cb_ex and cb_ey (the CPML strips' psi arrays aside)
    working_set = tm["kernel_cells"] * BYTES_PER_CELL[precision]
    if working_set < MALL_BYTES:
        roof["cache_note"] = (f"the launch's working set ({working_set / 1e6:.0f} MB: both "
                              "buffers of Ex, Ey, Hz and cb_ex, cb_ey) fits the 256 MiB "
                              "Infinity Cache, so FETCH_SIZE and WRITE_SIZE count the L2 "
                              "misses to the memory fabric, part of which that cache serves "
                              "instead of HBM; hbm_resident is the grid whose launches stream "
                              "from HBM")
    return roof


def _valu_issue(precision: str, r: dict, kernel_s: float) -> "dict | None":
    """The launch's VALU issue time (SQ_INSTS_VALU of this build over CUs x clock)
    against its duration, when the SQ table holds this shape."""
    import re

    tune = r["tune"]
    insts = SQ_VALU.get((precision, tune["temporal"], tune["band_rows"], r["cells"][0]))
    cus = re.search(r"CUs=(\d+)", r.get("device", ""))
    if insts is None or cus is None or kernel_build_id() != PMC_BUILD:
        return None
    issue_s = insts / (int(cus.group(1)) * VALU_CLOCK_HZ)
    return {
        "insts_per_launch": insts,
        "issue_us": round(issue_s * 1e6, 2),
        "frac": round(issue_s / kernel_s, 4),
        "note": "VALU wave-instructions of one launch (rocprofv3 SQ_INSTS_VALU of this build) "
                "at one per CU per cycle and 2.4 GHz, over the launch's event-timed "
                "duration: how close the launch runs to the issue limit of its own "
                "instructions",
    }


def hbm_resident(precision: str, steps: int) -> dict:
    """The same workload at 8000 x 8000 interior cells (8024 x 8024): its fields (1.3 GB
    in float32) exceed the 256 MB Infinity Cache, so every launch streams them from
    HBM.  96 warm-up steps, then ``steps`` timed ones."""
    r = run(precision, steps, 96, interior=8000)
    tm = r["timing"]
    nx, ny = r["cells"]
    traffic = None
    if (nx, ny) == (8024, 8024):
        traffic = measured_traffic(PMC_KIB_8024, (precision, r["tune"]["temporal"],
                                                  r["tune"]["band_rows"]))
    return {
        "grid": [nx, ny],
        "value": round(r["value"], 3),
        "unit": "Gcell-updates/s",
        "steps": steps,
        "ms_per_step": round(r["ms_per_step"], 5),
        "avg_ms_per_launch": round(tm["kernel_ms"], 5),
        "steps_per_launch": tm["depth"],
        "tune": r["tune"],
        "roofline": roofline(precision, r, traffic),
    }


def _fingerprint(solver) -> str:
    import hashlib

    h = hashlib.sha256()
    for name in ("ex", "ey", "hz", "psi_ex_y", "psi_ey_x", "psi_hz_x", "psi_hz_y"):
        h.update(np.ascontiguousarray(getattr(solver.state, name).get()).tobytes())
    for mon in solver.monitors:
        h.update(np.ascontiguousarray(mon.acc.get()).tobytes())
    return h.hexdigest()


def _monitor_values(solver) -> list:
    return [np.asarray(mon.acc.get()).astype(np.complex128) for mon in solver.monitors]


def run_batched(n: int, steps: int, seed: int, plain_steps: int) -> dict:
    from etalon_amd.batch import BatchSolver
    from etalon_amd.solver import FDTDSolver

    sims = uneven_batch(n, seed=seed, steps=steps)
    t0 = time.perf_counter()
    batch = BatchSolver(sims, autotune=True)
    setup = time.perf_counter() - t0
    cells = batch.cells
    batch.backend.synchronize()
    t0 = time.perf_counter()
    batch.advance(steps)
    batch.backend.synchronize()
    wall = time.perf_counter() - t0
    prints = [_fingerprint(s) for s in batch.solvers]
    f32_values = [_monitor_values(s) for s in batch.solvers]
    tune = batch._plan.resolved_tune().__dict__
    report = batch.tune_report
    del batch

    # the same simulations one after another on the fused path
    seq_wall, same = 0.0, 0
    for sim, fp in zip(sims, prints):
        solver = FDTDSolver(sim)
        solver.backend.synchronize()
        t0 = time.perf_counter()
        solver.advance(steps)
        solver.backend.synchronize()
        seq_wall += time.perf_counter() - t0
        same += _fingerprint(solver) == fp
        del solver

    # ... and with the per-step primitives (one launch per half-step, the reference's
    # loop), on the first `plain_steps` steps of each, scaled to the full step count
    plain_wall = 0.0
    for sim in sims:
        solver = FDTDSolver(sim)
        for _ in range(3):
            solver.step()
        solver.backend.synchronize()
        t0 = time.perf_counter()
        for _ in range(plain_steps):
            solver.step()
        solver.backend.synchronize()
        plain_wall += (time.perf_counter() - t0) * steps / plain_steps
        del solver

    # float32 error of every monitor against the float64 run of the same batch
    sims64 = uneven_batch(n, seed=seed, steps=steps, precision="double")
    batch64 = BatchSolver(sims64, autotune=True)
    batch64.advance(steps)
    batch64.backend.synchronize()
    errors = []
    for s32, s64 in zip(f32_values, (_monitor_values(s) for s in batch64.solvers)):
        for a, b in zip(s32, s64):
            scale = float(np.max(np.abs(b)))
            errors.append(float(np.max(np.abs(a - b)) / scale) if scale > 0 else 0.0)
    del batch64
    return {"n": n, "steps": steps, "cells": cells, "setup_s": setup, "wall": wall,
            "seq_wall": seq_wall, "plain_wall": plain_wall, "bitwise_same": same,
            "errors": errors, "tune": tune, "tune_report": report}


def _oracle_sample(task) -> tuple:
    """One simulation on the NumPy oracle: 2 warm-up steps, then ``steps`` timed ones;
    returns (cell-updates, seconds)."""
    from oracle import te_numpy as ref

    sim, steps = task
    solver = ref.Solver(sim)
    for _ in range(2):
        solver.step()
    t0 = time.perf_counter()
    for _ in range(steps):
        solver.step()
    return solver.g.nx * solver.g.ny * steps, time.perf_counter() - t0


def batched_cpu_baseline(sims: list, steps: int) -> dict:
    """The NumPy oracle over every simulation of the batch, one process per core (at
    most 16), a bounded number of steps each.  Runs before anything touches the GPU,
    so the worker processes are forked from a process without GPU state."""
    import multiprocessing

    workers = max(1, min(16, os.cpu_count() or 1))
    t0 = time.perf_counter()
    with multiprocessing.get_context("fork").Pool(workers) as pool:
        out = pool.map(_oracle_sample, [(sim, steps) for sim in sims], chunksize=1)
    wall = time.perf_counter() - t0
    updates = sum(u for u, _ in out)
    busy = sum(t for _, t in out)
    return {
        "value": updates / (busy / workers) / 1e9,
        "unit": "Gcell-updates/s",
        "cores": workers,
        "kind": "port",
        "single_core": updates / busy / 1e9,
        "sample": f"{steps} steps (after 2 warm-up) of each of the {len(sims)} simulations on "
                  f"the NumPy oracle, one process per core over {workers} cores; value = "
                  f"cell-updates / (summed step time / cores), single_core = cell-updates / "
                  f"summed step time; {wall:.1f} s wall with set-up; host {os.cpu_count()} "
                  f"CPUs, {_cpu_model()}",
    }


def main_batched(args) -> int:
    cpu = None
    if not args.no_cpu_baseline:
        cpu = batched_cpu_baseline(uneven_batch(n=args.batch, seed=args.seed,
                                                steps=args.batch_steps), args.cpu_batch_steps)
    r = run_batched(args.batch, args.batch_steps, args.seed, args.plain_steps)
    updates = r["cells"] * r["steps"]
    errors = np.array(r["errors"])
    line = {
        "metric": "batched sims/s (aggregate Gcell-updates/s)",
        "value": round(updates / r["wall"] / 1e9, 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": r["steps"],
        "warmup": 0,
        "ms_per_step": round(r["wall"] / r["steps"] * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": f"synthetic: {r['n']} seeded simulations (seed {args.seed})",
        "config": {"workload": f"batched uneven runs: {r['n']} simulations, interior 200-1500 "
                               f"cells per side at dl 0.05 um + 12 CPML, bends/rings eps 12.11 in "
                               f"2.085, Hz source, 1-8 monitors at 11 wavelengths, "
                               f"{r['steps']} steps each", "cells_total": r["cells"]},
        "sims_per_s": round(r["n"] / r["wall"], 2),
        "wall_s": round(r["wall"], 4),
        "sequential_fused": {"wall_s": round(r["seq_wall"], 4),
                             "gcells": round(updates / r["seq_wall"] / 1e9, 3),
                             "speedup": round(r["seq_wall"] / r["wall"], 2)},
        "sequential_plain": {"wall_s": round(r["plain_wall"], 3),
                             "gcells": round(updates / r["plain_wall"] / 1e9, 3),
                             "speedup": round(r["plain_wall"] / r["wall"], 2),
                             "sample": f"{args.plain_steps} step() calls per simulation, "
                                       "scaled to the full step count"},
        "bitwise_equal_to_sequential": f"{r['bitwise_same']}/{r['n']}",
        "f32_vs_f64": {"monitors": int(errors.size), "max": float(f"{errors.max():.3g}"),
                       "median": float(f"{np.median(errors):.3g}"),
                       "per_monitor": [float(f"{e:.3g}") for e in errors]},
        "tune": r["tune"],
        "autotune": [{**{k: t[k] for k in ("vec", "temporal", "band_rows", "waves", "launch")},
                      "gcells": round(t["gcells"], 1)}
                     for t in (r["tune_report"] or {}).get("timings", [])],
        "setup_s": round(r["setup_s"], 2),
    }
    if cpu is not None:
        line["cpu_baseline"] = cpu
    print(json.dumps(line), flush=True)
    return 0


def run_adjoint(seed: int, fd_pixels: int, cpu_steps: int, full: bool = True) -> dict:
    """The adjoint workload's figures; ``full`` False skips the sequential-schedule and
    store-mode gradients and the NumPy sample (the summary in the throughput line)."""
    from etalon_amd.adjoint import AdjointProblem
    from etalon_amd.solver import FDTDSolver

    sim, design, names = adjoint_design_sim("single", seed=seed)
    eps = np.asarray(design.permittivity)
    reps = 3  # every figure is the best of 3

    def grad_time(prob) -> tuple:
        best, g = None, None
        for _ in range(reps):
            tape = prob.forward(eps)
            t0 = time.perf_counter()
            g = prob.backward(tape)
            t = tape.forward_s + time.perf_counter() - t0
            best = t if best is None else min(best, t)
        return best, g

    FDTDSolver(sim).run()  # warm: first launches of the kernels
    # one plain forward run: a fresh solver (cold), and the same solver run again
    # from step 0 (warm: its plan, graphs and tables kept, as in a design loop)
    t_fwd_cold = min(FDTDSolver(sim).run().metadata["wall_time_s"] for _ in range(reps))
    loop_solver = FDTDSolver(sim)
    loop_solver.run()
    t_fwd = None
    for _ in range(reps):
        loop_solver.reset()
        t = loop_solver.run().metadata["wall_time_s"]
        t_fwd = t if t_fwd is None else min(t_fwd, t)
    # gradients: a design loop's problem (session kept), and fresh problems (cold)
    prob = AdjointProblem(sim, names)
    prob.run(eps)
    t_grad, grad = grad_time(prob)
    cold = [AdjointProblem(sim, names, reuse=False) for _ in range(reps)]
    t_cold = None
    for c in cold:
        tape = c.forward(eps)
        t0 = time.perf_counter()
        c.backward(tape)
        t = tape.forward_s + time.perf_counter() - t0
        t_cold = t if t_cold is None else min(t_cold, t)
    t_seq = t_store = grad_q = grad_s = None
    if full:
        seq = AdjointProblem(sim, names, schedule="sequential")
        seq.run(eps)
        t_seq, grad_q = grad_time(seq)
        store = AdjointProblem(sim, names, mode="store")
        store.run(eps)
        t_store, grad_s = grad_time(store)

    sim64, _, _ = adjoint_design_sim("double", seed=seed)
    prob64 = AdjointProblem(sim64, names)
    J64, grad64 = prob64.run(eps)
    scale = float(np.abs(grad64).max())
    rng = np.random.default_rng(seed + 1)
    fd_rel = []
    for flat in rng.choice(eps.size, size=fd_pixels, replace=False):
        p, q = divmod(int(flat), eps.shape[1])
        h = 1e-4 * eps[p, q]
        plus, minus = eps.copy(), eps.copy()
        plus[p, q] += h
        minus[p, q] -= h
        fd = (prob64.forward(plus).J - prob64.forward(minus).J) / (2 * h)
        fd_rel.append(abs(grad[p, q] - fd) / scale)

    # NumPy float64 adjoint (the oracle) on the first `cpu_steps` steps of the grid
    cpu_wall = None
    if full:
        from oracle import adjoint_numpy as adj
        from etalon_amd.workloads import adjoint_design_sim as make

        short, short_design, _ = make("double", seed=seed, steps=cpu_steps)
        t0 = time.perf_counter()
        adj.gradient(short, set(names), short_design)
        cpu_wall = time.perf_counter() - t0
    return {"t_fwd": t_fwd, "t_grad": t_grad, "t_store": t_store, "t_seq": t_seq,
            "t_fwd_cold": t_fwd_cold, "t_grad_cold": t_cold,
            "seq_equal": full and bool(np.array_equal(grad, grad_q)), "J": tape.J, "J64": J64,
            "grad_rel_f64": float(np.abs(grad - grad64).max() / scale),
            "store_equal": full and bool(np.array_equal(grad, grad_s)), "fd_rel": fd_rel,
            "cpu_wall": cpu_wall, "cpu_steps": cpu_steps, "K": prob.K,
            "cells": prob.grid.nx * prob.grid.ny, "steps": prob.n_steps, "pixels": eps.shape}


def main_adjoint(args) -> int:
    r = run_adjoint(args.seed, args.fd_pixels, args.cpu_adjoint_steps)
    fd = np.array(r["fd_rel"])
    cpu_grad = r["cpu_wall"] * r["steps"] / r["cpu_steps"]
    line = {
        "metric": "adjoint grad time vs 1 fwd",
        "value": round(r["t_grad"] / r["t_fwd"], 3),
        "unit": "x (wall of J and dJ/deps / wall of one plain forward run, both in a design "
                "loop: solver, plans and graphs kept between runs; best of 3 each)",
        "n_gpus": 1,
        "steps": r["steps"],
        "warmup": 1,
        "ms_per_step": round(r["t_grad"] / r["steps"] * 1e3, 5),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": f"synthetic: seeded design permittivity (seed {args.seed})",
        "config": {"workload": "adjoint gradient: 400x400 interior + 12 CPML at dl 0.05 um, "
                               "2x2 um design region of 40x40 pixels (eps 2.085-12.11) between "
                               "input/output strips, J = sum over 5 wavelengths of |DFT Hz|^2 on "
                               "the output line monitor, 4000 steps, checkpoints every "
                               f"{r['K']} steps", "cells": r["cells"],
                   "pixels": list(r["pixels"])},
        "t_fwd_s": round(r["t_fwd"], 4),
        "t_grad_s": round(r["t_grad"], 4),
        "cold": {"t_fwd_s": round(r["t_fwd_cold"], 4), "t_grad_s": round(r["t_grad_cold"], 4),
                 "ratio": round(r["t_grad_cold"] / r["t_fwd_cold"], 3),
                 "note": "fresh solver / fresh AdjointProblem per run (plans, graphs and "
                         "buffers built inside the timed region)"},
        "sequential_segments": {"t_grad_s": round(r["t_seq"], 4),
                                "ratio_vs_cold_fwd": round(r["t_seq"] / r["t_fwd_cold"], 3),
                                "gradient_bitwise_equal": r["seq_equal"]},
        "store_history": {"t_grad_s": round(r["t_store"], 4),
                          "ratio_vs_cold_fwd": round(r["t_store"] / r["t_fwd_cold"], 3),
                          "gradient_bitwise_equal": r["store_equal"]},
        "gradient_check": {"vs_f64_adjoint_max_rel": float(f"{r['grad_rel_f64']:.3g}"),
                           "fd_pixels": int(fd.size),
                           "vs_f64_central_fd_max_rel": float(f"{fd.max():.3g}"),
                           "J_f32": r["J"], "J_f64": r["J64"]},
        "cpu_baseline": {"value": round(cpu_grad, 2), "unit": "s per gradient (extrapolated)",
                         "cores": 1, "kind": "port",
                         "sample": f"NumPy float64 adjoint of the same grid over {r['cpu_steps']} "
                                   f"steps ({r['cpu_wall']:.1f} s), scaled to {r['steps']}; "
                                   f"host {os.cpu_count()} CPUs, {_cpu_model()}",
                         "vs_gpu_grad": round(cpu_grad / r["t_grad"], 1)},
    }
    print(json.dumps(line), flush=True)
    return 0


def run_batched_adjoint(args) -> dict:
    from etalon_amd.adjoint import AdjointProblem
    from etalon_amd.batch import BatchSolver
    from etalon_amd.batch_adjoint import BatchAdjointProblem
    from etalon_amd.workloads import phc_adjoint_candidates

    t_setup = time.perf_counter()
    sims, eps, names = phc_adjoint_candidates(n=args.candidates, seed=args.seed,
                                              steps=args.phc_steps)
    problems = [AdjointProblem(sim, names) for sim in sims]
    batch = BatchAdjointProblem(problems)
    print(f"batched adjoint: {len(sims)} candidates set up in "
          f"{time.perf_counter() - t_setup:.1f} s", file=sys.stderr, flush=True)
    # the first gradient of a fresh problem: the batch's states and forward plan are
    # built outside the timed region (as for the forward run below), its checkpoint
    # storage, reverse plans, co-plan and graphs inside
    tape = batch.forward(eps)
    t0 = time.perf_counter()
    batch.backward(tape)
    t_cold = tape.forward_s + time.perf_counter() - t0
    del tape
    # a design loop: the session of the first gradient kept (best of 3)
    t_grad, grads, J32 = float("inf"), None, None
    for _ in range(3):
        tape = batch.forward(eps)
        t0 = time.perf_counter()
        g = batch.backward(tape)
        t = tape.forward_s + time.perf_counter() - t0
        if t < t_grad:
            t_grad, grads, J32 = t, g, [float(j) for j in tape.J]
        del tape
    batch.close()
    # store mode: the forward pass keeps every design box's Hz of every step in HBM and
    # the reverse steps read it, no recompute (first gradient warms, best of 3 after)
    stored = BatchAdjointProblem([AdjointProblem(sim, names, mode="store") for sim in sims])
    t_store, g_store = float("inf"), None
    for rep in range(4):
        tape = stored.forward(eps)
        t0 = time.perf_counter()
        g = stored.backward(tape)
        t = tape.forward_s + time.perf_counter() - t0
        hist_bytes = sum(h.nbytes for h in tape.histories)
        if rep > 0 and t < t_store:
            t_store, g_store = t, g
        del tape
    stored.close()
    # the store sweep groups the reverse steps into other temporal blocks than the
    # checkpointed one: the float64 gradients agree to rounding, not bit for bit
    store_rel = max(float(np.abs(a - b).max() / np.abs(b).max()) for a, b in zip(g_store, grads))
    design_sims = [p.with_permittivity(e) for p, e in zip(problems, eps)]
    BatchSolver(design_sims).run()  # warm
    fwd = BatchSolver(design_sims)
    t_fwd = fwd.run()[0].metadata["wall_time_s"]
    del fwd
    print(f"batched adjoint: forward {t_fwd:.3f} s, J and gradients {t_grad:.3f} s",
          file=sys.stderr, flush=True)

    # candidate 0 in float64: the adjoint gradient, and central differences on random
    # pixels (the 2 P perturbed runs advanced as one float64 batch)
    sims64, _, _ = phc_adjoint_candidates(n=1, seed=args.seed, steps=args.phc_steps,
                                          precision="double")
    p64 = AdjointProblem(sims64[0], names)
    J64, g64 = p64.run(eps[0])
    scale = float(np.abs(g64).max())
    rng = np.random.default_rng(args.seed + 1)
    picks = rng.choice(eps[0].size, size=args.fd_pixels, replace=False)
    runs = []
    for flat in picks:
        p, q = divmod(int(flat), eps[0].shape[1])
        h = 1e-4 * eps[0][p, q]
        for sign in (1.0, -1.0):
            e = eps[0].copy()
            e[p, q] += sign * h
            runs.append(p64.with_permittivity(e))
    fd_batch = BatchSolver(runs)
    fd_batch.run()
    Jfd = [p64._objective(s) for s in fd_batch.solvers]
    del fd_batch
    fd_rel = []
    for k, flat in enumerate(picks):
        p, q = divmod(int(flat), eps[0].shape[1])
        h = 1e-4 * eps[0][p, q]
        fd_rel.append(abs(grads[0][p, q] - (Jfd[2 * k] - Jfd[2 * k + 1]) / (2 * h)) / scale)

    # the NumPy float64 adjoint on a bounded sample: a grid of the same size (the cost
    # per step does not depend on the permittivity map)
    from oracle import adjoint_numpy as adj

    short, short_design, _ = adjoint_design_sim("double", seed=args.seed,
                                                steps=args.cpu_phc_steps, interior=1000,
                                                region=4.0, pixels=80)
    t0 = time.perf_counter()
    adj.gradient(short, set(names), short_design)
    cpu_wall = time.perf_counter() - t0
    n_cells = sum(p.grid.nx * p.grid.ny for p in problems)
    return {"t_grad": t_grad, "t_cold": t_cold, "t_fwd": t_fwd, "J": J32, "t_store": t_store,
            "store_rel": store_rel, "hist_bytes": hist_bytes,
            "J64_0": J64, "grad_rel_f64": float(np.abs(grads[0] - g64).max() / scale),
            "fd_rel": fd_rel, "cpu_wall": cpu_wall, "cells": n_cells, "n": len(problems),
            "steps": problems[0].n_steps, "K": problems[0].K, "pixels": eps[0].shape,
            "grid": (problems[0].grid.nx, problems[0].grid.ny), "holes": len(sims[0].structures) - 1}


def main_batched_adjoint(args) -> int:
    r = run_batched_adjoint(args)
    fd = np.array(r["fd_rel"])
    cpu_total = r["cpu_wall"] / args.cpu_phc_steps * r["steps"] * r["n"]
    line = {
        "metric": "batched adjoint grad time vs batched fwd",
        "value": round(r["t_grad"] / r["t_fwd"], 3),
        "unit": "x (wall of J and dJ/deps of every candidate in a design loop: batch, plans "
                "and graphs kept between gradients, best of 3 / wall of one batched plain "
                "forward run of them)",
        "n_gpus": 1,
        "steps": r["steps"],
        "warmup": 1,
        "ms_per_step": round(r["t_grad"] / r["steps"] * 1e3, 5),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": f"synthetic: {r['n']} seeded design candidates (seed {args.seed})",
        "config": {"workload": f"batched adjoint: {r['n']} candidates, 1000x1000 interior + 12 "
                               "CPML at dl 0.05 um, triangular-lattice photonic crystal (a 0.42 "
                               "um, air holes r 0.126 um in eps 12.11) with a W1 line defect, "
                               "4x4 um design region of 80x80 pixels (eps 2.085-12.11), J = "
                               "sum over 9 wavelengths 1.50-1.60 um of |DFT Hz|^2 on the output "
                               f"line monitor, {r['steps']} steps, checkpoints every {r['K']} "
                               "steps; forward recompute and reverse steps of every candidate "
                               "in one launch per step block",
                   "grid": list(r["grid"]), "holes": r["holes"], "pixels": list(r["pixels"]),
                   "cells_total": r["cells"]},
        "t_grad_s": round(r["t_grad"], 3),
        "t_fwd_s": round(r["t_fwd"], 3),
        "cold": {"t_grad_s": round(r["t_cold"], 3), "ratio": round(r["t_cold"] / r["t_fwd"], 3),
                 "note": "first gradient of a fresh problem (checkpoint storage, reverse "
                         "plans, co-plan and graphs built inside the timed region)"},
        "store_history": {"t_grad_s": round(r["t_store"], 3),
                          "ratio": round(r["t_store"] / r["t_fwd"], 3),
                          "candidates_per_s": round(r["n"] / r["t_store"], 2),
                          "history_bytes": int(r["hist_bytes"]),
                          "vs_checkpoint_max_rel": float(f"{r['store_rel']:.3g}"),
                          "note": "design loop, best of 3: the forward pass keeps each design "
                                  "box's Hz of every step in HBM, the reverse steps read it "
                                  "(no recompute)"},
        "candidates_per_s": round(r["n"] / r["t_grad"], 2),
        "gcells_grad": round(3 * r["cells"] * r["steps"] / r["t_grad"] / 1e9, 1),
        "gradient_check": {"candidate": 0,
                           "vs_f64_adjoint_max_rel": float(f"{r['grad_rel_f64']:.3g}"),
                           "fd_pixels": int(fd.size),
                           "vs_f64_central_fd_max_rel": float(f"{fd.max():.3g}"),
                           "J_f32": r["J"][0], "J_f64": r["J64_0"]},
        "cpu_baseline": {"value": round(cpu_total, 1), "unit": "s per batch of gradients "
                         "(extrapolated)", "cores": 1, "kind": "port",
                         "sample": f"NumPy float64 adjoint of a 1024x1024 grid over "
                                   f"{args.cpu_phc_steps} steps ({r['cpu_wall']:.1f} s), "
                                   f"scaled to {r['steps']} steps x {r['n']} candidates; host "
                                   f"{os.cpu_count()} CPUs, {_cpu_model()}",
                         "vs_gpu_grad": round(cpu_total / r["t_grad"], 1)},
    }
    print(json.dumps(line), flush=True)
    return 0


def one_pass(args) -> int:
    """One gradient of the adjoint workload (or of every candidate of the batched one)
    after one warm-up, for profilers: prints the forward and backward wall times."""
    from etalon_amd.adjoint import AdjointProblem

    if args.workload == "adjoint":
        sim, design, names = adjoint_design_sim("single", seed=args.seed)
        prob, eps = AdjointProblem(sim, names), np.asarray(design.permittivity)
    else:
        from etalon_amd.batch_adjoint import BatchAdjointProblem
        from etalon_amd.workloads import phc_adjoint_candidates

        sims, eps, names = phc_adjoint_candidates(n=args.candidates, seed=args.seed,
                                                  steps=args.phc_steps)
        prob = BatchAdjointProblem([AdjointProblem(s, names) for s in sims])
    prob.run(eps)
    tape = prob.forward(eps)
    t0 = time.perf_counter()
    prob.backward(tape)
    print(json.dumps({"workload": args.workload, "forward_s": round(tape.forward_s, 5),
                      "backward_s": round(time.perf_counter() - t0, 5)}), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--precision", choices=("single", "double"), default="single")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="throughput line without the other precision, batched and adjoint "
                         "summaries (profiling runs)")
    ap.add_argument("--cpu-steps", type=int, default=400)
    ap.add_argument("--copy-mb", type=int, default=1024,
                    help="MiB of the device copy timed for the achievable HBM rate (0: skip)")
    ap.add_argument("--interior", type=int, default=2000,
                    help="interior cells per side of the throughput grid (profiling runs)")
    ap.add_argument("--hbm-steps", type=int, default=1920,
                    help="steps of the 8024x8024 companion run (0: skip it)")
    ap.add_argument("--workload", choices=("throughput", "batched", "adjoint", "batched-adjoint"),
                    default="throughput")
    ap.add_argument("--fd-pixels", type=int, default=16)
    ap.add_argument("--cpu-adjoint-steps", type=int, default=300)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--batch-steps", type=int, default=4000)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--plain-steps", type=int, default=100)
    ap.add_argument("--cpu-batch-steps", type=int, default=40,
                    help="batched workload: oracle steps per simulation for cpu_baseline")
    ap.add_argument("--candidates", type=int, default=32)
    ap.add_argument("--phc-steps", type=int, default=8000)
    ap.add_argument("--cpu-phc-steps", type=int, default=20)
    ap.add_argument("--one-pass", action="store_true",
                    help="adjoint workloads: one warm-up and one timed gradient only (profiling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the FDTD "
              "path runs on a single MI355X; multi-GPU sharding is a design note only)",
              file=sys.stderr)
        return 2
    if args.workload == "batched":
        return main_batched(args)
    if args.one_pass and args.workload in ("adjoint", "batched-adjoint"):
        return one_pass(args)
    if args.workload == "adjoint":
        return main_adjoint(args)
    if args.workload == "batched-adjoint":
        return main_batched_adjoint(args)

    main_run = run(args.precision, args.steps, args.warmup, args.interior)
    nx, ny = main_run["cells"]
    tune = main_run["tune"]
    traffic = None
    if (nx, ny) == (2024, 2024):
        traffic = measured_traffic(PMC_KIB, (args.precision, tune["vec"], tune["temporal"],
                                             tune["band_rows"], tune["launch"]))
    roof = roofline(args.precision, main_run, traffic)
    if args.copy_mb > 0:
        from etalon_amd.hip.calibrate import copy_rate

        # an achievable HBM rate: a device copy of buffers well beyond the Infinity Cache
        copy = copy_rate(args.copy_mb, 10)
        roof["copy_gbs"] = max(copy["gbs_read_plus_write"], copy["stream_copy_gbs"])
        roof["copy_note"] = (f"the faster of two copies of {args.copy_mb} MiB into another "
                             f"buffer, bytes read + written per second: the HIP runtime's "
                             f"{copy['gbs_read_plus_write']}, a 16-byte streaming kernel "
                             f"{copy['stream_copy_gbs']}")
        roof["achieved_over_copy"] = round(roof["achieved"] / roof["copy_gbs"], 4)
    line = {
        "metric": "Gcell-updates/s & %HBM roofline",
        "value": round(main_run["value"], 3),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(main_run["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": (round(main_run["value"] / REFERENCE_GCELLS[args.precision], 2)
                        if args.interior == 2000 else None),
        "dtype": "f32" if args.precision == "single" else "f64",
        "data": "synthetic: strip waveguide geometry, zero initial fields, seeded-free",
        "config": {
            "workload": f"throughput grid: {args.interior}x{args.interior} interior + 12-layer "
                        f"CPML = {nx}x{ny} cells, "
                        "Si strip eps 12.11 w 0.5 um in oxide eps 2.085, Hz Gaussian source "
                        "1.55 um, line DFT monitor Hz at 21 wavelengths, courant 0.99",
            "grid": [nx, ny],
            "precision": args.precision,
        },
        "roofline": roof,
        "timed_region": _timed_region(main_run),
        "kernel": _kernel_summary(main_run, args.precision),
        "baseline_note": "vs_baseline divides by the reference's published CuPy throughput on "
                         "this grid (RTX 4060 Ti: 2.94 Gcell/s f32, 1.5 f64)",
        "device": main_run["device"],
    }
    if args.hbm_steps > 0:
        line["hbm_resident"] = hbm_resident(args.precision, args.hbm_steps)
    if not args.no_extras:
        other = "double" if args.precision == "single" else "single"
        summary = precision_summary(other, args.steps, args.warmup, args.interior)
        if args.hbm_steps > 0:
            summary["hbm_resident"] = hbm_resident(other, args.hbm_steps)
        line[f"{'f64' if other == 'double' else 'f32'}"] = summary
        line["batched"] = batched_summary(args)
        line["adjoint"] = adjoint_summary(args)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(args.precision, args.cpu_steps)
    print(json.dumps(line), flush=True)
    return 0


def _timed_region(r: dict) -> dict:
    return {"wall_ms": round(r["wall"] * 1e3, 4),
            "gpu_ms_per_step": round(r["gpu_ms_per_step"], 5),
            "host_share": round(r["host_share"], 4),
            "note": "host_share = 1 - steps x the GPU's own time per step (HIP events over "
                    "direct launches of the same plan) / the timed wall: what the host's "
                    "launches, table work and the syncs around the region add"}


def _kernel_summary(r: dict, precision: str) -> dict:
    tm = r["timing"]
    return {
        "name": ("joint_kernel<R, T, 0, 0, BR> (every tile kind, one launch per step block)"
                 if tm["launch"] == 1 else "tile_kernel (interior tiles)"),
        "avg_ms_per_launch": round(tm["kernel_ms"], 5),
        "steps_per_launch": tm["depth"],
        "cells_per_step": tm["kernel_cells"],
        "algorithmic_bytes_per_launch": r["kernel_bytes"],
        "bytes_per_cell_update": BYTES_PER_CELL[precision],
        "tune": r["tune"],
        "fused_launch_ms": round(tm["launch_ms"], 5),
        "events_ms_per_step": round(r["gpu_ms_per_step"], 5),
    }


def precision_summary(precision: str, steps: int, warmup: int, interior: int) -> dict:
    """The throughput grid in the other precision: the same steps and warm-up."""
    r = run(precision, steps, warmup, interior)
    nx, ny = r["cells"]
    tune = r["tune"]
    traffic = None
    if (nx, ny) == (2024, 2024):
        traffic = measured_traffic(PMC_KIB, (precision, tune["vec"], tune["temporal"],
                                             tune["band_rows"], tune["launch"]))
    return {"value": round(r["value"], 3), "unit": "Gcell-updates/s",
            "dtype": "f64" if precision == "double" else "f32",
            "steps": steps, "warmup": warmup, "ms_per_step": round(r["ms_per_step"], 5),
            "vs_baseline": (round(r["value"] / REFERENCE_GCELLS[precision], 2)
                            if interior == 2000 else None),
            "roofline": roofline(precision, r, traffic), "timed_region": _timed_region(r),
            "kernel": _kernel_summary(r, precision)}


def batched_summary(args) -> dict:
    """The batched workload (64 uneven simulations, 4000 steps) in the throughput line."""
    r = run_batched(args.batch, args.batch_steps, args.seed, args.plain_steps)
    updates = r["cells"] * r["steps"]
    errors = np.array(r["errors"])
    return {"metric": "batched sims/s (aggregate Gcell-updates/s)",
            "value": round(updates / r["wall"] / 1e9, 3), "unit": "Gcell-updates/s",
            "sims_per_s": round(r["n"] / r["wall"], 2), "n_sims": r["n"], "steps": r["steps"],
            "cells_total": r["cells"], "wall_s": round(r["wall"], 4), "dtype": "f32",
            "speedup_vs_sequential_plain": round(r["plain_wall"] / r["wall"], 2),
            "speedup_vs_sequential_fused": round(r["seq_wall"] / r["wall"], 2),
            "bitwise_equal_to_sequential": f"{r['bitwise_same']}/{r['n']}",
            "f32_vs_f64_max_rel": float(f"{errors.max():.3g}"),
            "f32_vs_f64_monitors": int(errors.size), "tune": r["tune"],
            "note": f"seed {args.seed}; sequential plain = {args.plain_steps} step() calls per "
                    "simulation scaled to the full step count; full line: --workload batched"}


def adjoint_summary(args) -> dict:
    """The adjoint workload (424 x 424, 40 x 40 pixels, 4000 steps) in the throughput line."""
    r = run_adjoint(args.seed, args.fd_pixels, args.cpu_adjoint_steps, full=False)
    fd = np.array(r["fd_rel"])
    return {"metric": "adjoint grad time vs 1 fwd",
            "value": round(r["t_grad"] / r["t_fwd"], 3), "unit": "x",
            "t_grad_s": round(r["t_grad"], 4), "t_fwd_s": round(r["t_fwd"], 4),
            "cold_ratio": round(r["t_grad_cold"] / r["t_fwd_cold"], 3),
            "cold_t_grad_s": round(r["t_grad_cold"], 4), "cold_t_fwd_s": round(r["t_fwd_cold"], 4),
            "steps": r["steps"], "K": r["K"], "pixels": list(r["pixels"]), "dtype": "f32",
            "fd_pixels": int(fd.size), "vs_f64_central_fd_max_rel": float(f"{fd.max():.3g}"),
            "vs_f64_adjoint_max_rel": float(f"{r['grad_rel_f64']:.3g}"),
            "note": "value: design loop (solver, plans and graphs kept), best of 3; cold: fresh "
                    "solver / problem per run; full line: --workload adjoint"}


if __name__ == "__main__":
    sys.exit(main())
