"""Write tests/golden/m0_record.json: the M0 physics values the reference measured
and recorded (bench/results/b2ff355.json:8-75, CUDA float32 on an RTX 4060 Ti) —
dispersion residuals against the Yee relation and minimum v/c per (dl, angle), the
12-layer CPML reflection at the side and corner probes, the late-time energy after
30 000 steps, and the Green's-function error.  Data only: the reference's JSON file
is read as text.

    python tests/golden/make_m0_record.py [/root/reference]
"""

import json
import sys
from pathlib import Path

SOURCE = "bench/results/b2ff355.json"


def main(reference: Path) -> None:
    record = json.loads((reference / SOURCE).read_text(encoding="utf-8"))
    out = {
        "source": f"{SOURCE} (git_sha {record['git_sha']}, backend {record['backend']}, "
                  f"{record['device']})",
        "dispersion": [{"dl_um": d["dl_um"], "angle_deg": d["angle_deg"],
                        "max_abs_residual_vs_yee": d["max_abs_residual_vs_yee"],
                        "min_v_over_c": d["min_v_over_c"]} for d in record["dispersion"]],
        "cpml_reflection": record["cpml_reflection"],
        "late_time_final_energy_rel": record["late_time_final_energy_rel"],
        "green_function_max_rel_error": record["green_function_max_rel_error"],
    }
    path = Path(__file__).resolve().parent / "m0_record.json"
    path.write_text(json.dumps(out, indent=1) + "\n", encoding="utf-8")
    print(path)


if __name__ == "__main__":
    main(Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference"))
