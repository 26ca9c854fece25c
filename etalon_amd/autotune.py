"""Autotuner of the fused kernels' shape: tile rows, vector width, temporal depth,
waves per workgroup and launch mode (one kernel per step block over every tile
kind, or one per kind on three streams), measured on the simulations it will run.

The candidates are the register-tile shapes compiled into libetalon_hip.so (see
``CANDIDATES``).  Each one advances the real simulations by ``steps`` steps from
their current (zero) state through a fused plan replayed from its graph — the
timed region is the production path — and the fastest per cell-update wins.  The
simulations are then reset to step 0, so tuning never changes a result.  Choices
are cached per device, precision and workload signature in a JSON file
(``$ETALON_AMD_TUNE_CACHE``, default ``~/.cache/etalon_amd/tune.json``).
"""

from __future__ import annotations

import json
import os
import time
from pathlib import Path
from typing import Optional, Sequence

from etalon_amd.hip.plan import FusedPlan, Tune

# (vec, temporal, band_rows, waves, launch) per precision, all compiled shapes
CANDIDATES = {
    "single": [(1, 4, 8, 1, 3), (1, 4, 8, 2, 3), (1, 4, 8, 4, 3), (1, 4, 16, 1, 3),
               (1, 4, 16, 2, 3), (2, 4, 8, 1, 3), (1, 2, 8, 1, 3), (1, 4, 8, 1, 1),
               (1, 2, 8, 1, 1), (1, 4, 16, 1, 1), (1, 6, 8, 1, 1), (1, 6, 16, 1, 1),
               (1, 8, 16, 1, 1)],
    "double": [(1, 4, 8, 1, 3), (1, 4, 8, 2, 3), (1, 4, 8, 4, 3), (1, 2, 16, 1, 3),
               (1, 4, 16, 1, 3), (1, 2, 8, 1, 3), (1, 4, 8, 1, 1), (1, 2, 8, 1, 1),
               (1, 4, 16, 1, 1), (1, 6, 8, 1, 1), (1, 6, 16, 1, 1)],
}
_KEYS = ("vec", "temporal", "band_rows", "waves", "launch")


def cache_path() -> Path:
    env = os.environ.get("ETALON_AMD_TUNE_CACHE")
    return Path(env) if env else Path.home() / ".cache" / "etalon_amd" / "tune.json"


def signature(solvers: Sequence, device: str) -> str:
    """Workload key: device, precision, simulation count and grid sizes (to 2 digits)."""
    cells = sorted(s.grid.nx * s.grid.ny for s in solvers)
    total = sum(cells)
    digits = f"{total:.2g}" if total else "0"
    return (f"{device}|{solvers[0].sim.precision}|n={len(solvers)}|cells={digits}"
            f"|max={cells[-1]:.2g}|min={cells[0]:.2g}")


def _load_cache() -> dict:
    try:
        return json.loads(cache_path().read_text())
    except (OSError, ValueError):
        return {}


def _save_cache(cache: dict) -> None:
    path = cache_path()
    try:
        path.parent.mkdir(parents=True, exist_ok=True)
        tmp = path.with_suffix(".tmp")
        tmp.write_text(json.dumps(cache, indent=1, sort_keys=True))
        tmp.replace(path)
    except OSError:
        pass  # a read-only home only costs a re-tune


def _time_candidate(solvers: Sequence, tune: Tune, steps: int) -> float:
    plan = FusedPlan([s.state for s in solvers], tune)
    try:
        ids = [s.register(plan, k) for k, s in enumerate(solvers)]
        for k, (s, mon_ids) in enumerate(zip(solvers, ids)):
            delta, phasors = s._tables(0, 2 * steps)
            if delta is not None:
                plan.source_table(0, delta, sim=k)
            for mid, table in zip(mon_ids, phasors):
                if table is not None:
                    plan.phasor_table(mid, 0, table, sim=k)
        backend = solvers[0].backend
        plan.advance(0, steps)  # builds the graph, warms the caches
        backend.synchronize()
        t0 = time.perf_counter()
        plan.advance(steps, steps)
        backend.synchronize()
        return time.perf_counter() - t0
    finally:
        plan.close()


def autotune(solvers: Sequence, steps: int = 128, candidates: Optional[list] = None,
             use_cache: bool = True) -> tuple[Tune, dict]:
    """The fastest compiled shape for advancing ``solvers`` together; every solver is
    reset to step 0 afterwards.  Returns (tune, report)."""
    backend = solvers[0].backend
    precision = solvers[0].sim.precision
    key = signature(solvers, backend.device)
    cache = _load_cache() if use_cache else {}
    if use_cache and key in cache:
        v = cache[key]
        return Tune(band_rows=v["band_rows"], vec=v["vec"], temporal=v["temporal"],
                    waves=v["waves"], launch=v.get("launch", 0)), \
            {"key": key, "cached": True, "choice": v}
    cands = candidates or CANDIDATES[precision]
    cells = sum(s.grid.nx * s.grid.ny for s in solvers)
    timings = []
    for vec, depth, band, waves, launch in cands:
        tune = Tune(band_rows=band, vec=vec, temporal=depth, waves=waves, launch=launch)
        try:
            wall = _time_candidate(solvers, tune, steps)
        finally:
            for s in solvers:
                s.reset()
            backend.synchronize()
        timings.append({"vec": vec, "temporal": depth, "band_rows": band, "waves": waves,
                        "launch": launch, "gcells": cells * steps / wall / 1e9})
    best = max(timings, key=lambda t: t["gcells"])
    choice = {k: best[k] for k in _KEYS}
    if use_cache:
        cache[key] = {**choice, "gcells": round(best["gcells"], 2)}
        _save_cache(cache)
    return (Tune(band_rows=choice["band_rows"], vec=choice["vec"], temporal=choice["temporal"],
                 waves=choice["waves"], launch=choice["launch"]),
            {"key": key, "cached": False, "timings": timings,
                                          "choice": choice})
