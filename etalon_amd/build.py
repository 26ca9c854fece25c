"""Build libetalon_hip.so (gfx950) in-tree with hipcc.

    python -m etalon_amd.build          # compile every csrc/*.hip into etalon_amd/libetalon_hip.so

Floating-point flags are part of the contract with the NumPy reference: no FP
contraction (every multiply and add rounds separately, as NumPy does), no
flush-to-zero of f32 denormals, no fast-math.  Packed-FP32 instructions are
disabled (`-packed-fp32-ops`), so the compiler emits scalar v_add/v_mul/v_fma-free
code for float arithmetic.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
INCLUDE = PKG.parent / "include"
LIB_NAME = "libetalon_hip.so"
ARCH = "gfx950"

DEVICE_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++20",
    "-ffp-contract=off",
    "-fno-gpu-flush-denormals-to-zero",
    "-fno-fast-math",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
]


def hipcc() -> str:
    for candidate in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if candidate and Path(candidate).exists():
            return candidate
    raise RuntimeError("hipcc not found (ROCm is needed to build libetalon_hip.so)")


def sources() -> list[Path]:
    return sorted(CSRC.glob("*.hip"))


def build(verbose: bool = False, jobs: int = 8) -> Path:
    """Compile all sources (always; no staleness check) and link the shared library."""
    out_dir = PKG / "_build"
    out_dir.mkdir(exist_ok=True)
    cc = hipcc()
    objects = []
    procs = []
    for src in sources():
        obj = out_dir / (src.stem + ".o")
        cmd = [cc, *DEVICE_FLAGS, f"-I{INCLUDE}", "-c", str(src), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objects.append(obj)
        if len(procs) >= jobs:
            _wait(procs.pop(0))
    for p in procs:
        _wait(p)
    lib = PKG / LIB_NAME
    tmp = PKG / (LIB_NAME + ".tmp")
    cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objects), "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, lib)
    return lib


def _wait(item: tuple[Path, subprocess.Popen]) -> None:
    src, proc = item
    out, _ = proc.communicate()
    text = out.decode(errors="replace")
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src.name}:\n{text}")
    # The host pass does not know the packed-fp32 target feature; that warning is expected.
    lines = [ln for ln in text.splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
    if lines:
        print(f"[{src.name}]\n" + "\n".join(lines), file=sys.stderr)


if __name__ == "__main__":
    path = build(verbose="-v" in sys.argv)
    print(path)
