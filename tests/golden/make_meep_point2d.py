"""Write tests/golden/meep_point2d.npz from the reference's committed Meep result.

    python tests/golden/make_meep_point2d.py /root/reference

Source: bench/meep/out/2d/point2d.json of the reference tree (Meep 1.35.0,
resolution 50, 8994 steps, double precision): the complex T21 = Hz(probe 2) /
Hz(probe 1) at 51 wavelengths.  Data only; the gate is 1e-3 on max |ΔT21|
(bench/metrics.py:20).
"""

from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np


def main(ref_root: str) -> None:
    path = Path(ref_root) / "bench/meep/out/2d/point2d.json"
    data = json.loads(path.read_text())
    run = data["runs"][0]
    t21 = np.array([complex(re, im) for re, im in run["s"]["T21"]])
    np.savez(Path(__file__).resolve().parent / "meep_point2d.npz",
             wavelengths=np.array(data["wavelengths_um"], dtype=np.float64), t21=t21,
             steps=np.int64(run["grid"]["steps"]), resolution=np.int64(run["resolution"]),
             meep_version=np.str_(run["versions"]["meep"]))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
