"""Write tests/golden/adjoint_cfg4.npz and adjoint_cfg5.npz: the float64 NumPy
adjoint gradients (oracle/adjoint_numpy.py) of the two adjoint workloads at their
own sizes, which the oracle needs minutes for (about 30 ms and 170 ms per forward
and reverse step on one core), so the GPU tests compare against these instead of
running it:

* cfg4 — workloads.adjoint_design_sim("double", seed=0): 424 x 424 cells, a 40 x 40
  pixel design region, 4000 steps;
* cfg5 — candidate 0 of workloads.phc_adjoint_candidates(n=1, seed=0,
  steps=CFG5_STEPS, precision="double"): 1024 x 1024 cells of photonic crystal, an
  80 x 80 pixel design region, CFG5_STEPS steps (enough for the pulse to cross the
  design region and reach the output monitor).

Each file holds J, the gradient and the step count.

    python tests/golden/make_adjoint_golden.py [cfg4] [cfg5]
"""

import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from etalon_amd.workloads import adjoint_design_sim, phc_adjoint_candidates  # noqa: E402
from oracle import adjoint_numpy as adj  # noqa: E402

CFG5_STEPS = 2048
OUT = Path(__file__).resolve().parent


def cfg4() -> None:
    sim, design, names = adjoint_design_sim("double", seed=0)
    t0 = time.perf_counter()
    J, g = adj.gradient(sim, set(names), design)
    np.savez(OUT / "adjoint_cfg4.npz", J=J, gradient=g, steps=4000, seed=0)
    print(f"cfg4: J = {J:.6e}, max |g| = {np.abs(g).max():.3e}, {time.perf_counter() - t0:.0f} s")


def cfg5() -> None:
    sims, eps, names = phc_adjoint_candidates(n=1, seed=0, steps=CFG5_STEPS, precision="double")
    t0 = time.perf_counter()
    J, g = adj.gradient(sims[0], set(names), sims[0].structures[-1])
    np.savez(OUT / "adjoint_cfg5.npz", J=J, gradient=g, steps=CFG5_STEPS, seed=0)
    print(f"cfg5: J = {J:.6e}, max |g| = {np.abs(g).max():.3e}, {time.perf_counter() - t0:.0f} s")


if __name__ == "__main__":
    which = sys.argv[1:] or ["cfg4", "cfg5"]
    for name in which:
        {"cfg4": cfg4, "cfg5": cfg5}[name]()
