"""Build libetalon_hip's host code against the HIP runtime stub (hip_stub.cpp), with
AddressSanitizer, for lifecycle checks on a machine without a GPU.

    python tests/hoststub/build.py [OUT_DIR]    # prints the path of the library

Every csrc/*.hip is compiled for the host only (no device code); the fat-binary
symbols that host-only objects still reference are defined empty; the library links
the stub and the sanitizer runtime (shared, so that a Python process can preload it).
"""

from __future__ import annotations

import os
import re
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
CSRC = ROOT / "etalon_amd" / "csrc"
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
HIPCC = ROCM / "bin" / "hipcc"
CLANG = ROCM / "lib" / "llvm" / "bin" / "clang++"
LIB_NAME = "libetalon_hip_hoststub.so"

COMMON = ["-O1", "-g", "-fno-omit-frame-pointer", "-fPIC", "-std=c++20", "-ffp-contract=off",
          "-Wno-unused-function", "-Wno-unused-variable", f"-I{ROOT / 'include'}"]


def asan_runtime() -> Path:
    """The shared AddressSanitizer runtime of ROCm's clang (what the library links)."""
    found = sorted((ROCM / "lib" / "llvm" / "lib" / "clang").glob(
        "*/lib/linux/libclang_rt.asan-x86_64.so"))
    if not found:
        raise FileNotFoundError("no libclang_rt.asan-x86_64.so under ROCm's clang")
    return found[-1]


def available() -> bool:
    try:
        asan_runtime()
    except FileNotFoundError:
        return False
    return HIPCC.exists() and CLANG.exists()


def _run(cmd: list) -> None:
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{' '.join(map(str, cmd))}\n{r.stdout}")


def build(out_dir: Path) -> Path:
    out_dir.mkdir(parents=True, exist_ok=True)
    jobs = []
    for src in sorted(CSRC.glob("*.hip")):
        obj = out_dir / (src.stem + ".o")
        jobs.append([str(HIPCC), "--offload-arch=gfx950", "--cuda-host-only", *COMMON,
                     "-Xarch_host", "-fsanitize=address", "-c", str(src), "-o", str(obj)])
    stub_obj = out_dir / "hip_stub.o"
    jobs.append([str(CLANG), "-D__HIP_PLATFORM_AMD__", f"-I{ROCM / 'include'}", *COMMON,
                 "-fsanitize=address", "-c", str(HERE / "hip_stub.cpp"), "-o", str(stub_obj)])
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as pool:
        list(pool.map(_run, jobs))
    objs = [out_dir / (src.stem + ".o") for src in sorted(CSRC.glob("*.hip"))]
    # the fat-binary symbols each host-only object's registration still names
    names = set()
    for obj in objs:
        nm = subprocess.run(["nm", "-u", str(obj)], stdout=subprocess.PIPE, text=True,
                            check=True).stdout
        names.update(re.findall(r"\b(__hip_fatbin_[0-9a-f]+)\b", nm))
    fat_src = out_dir / "fatbin_stub.cpp"
    fat_src.write_text("".join(f'extern "C" const char {n}[16] = {{0}};\n' for n in sorted(names)))
    fat_obj = out_dir / "fatbin_stub.o"
    _run([str(CLANG), *COMMON, "-c", str(fat_src), "-o", str(fat_obj)])
    lib = out_dir / LIB_NAME
    _run([str(CLANG), "-shared", "-fsanitize=address", "-shared-libsan", "-o", str(lib),
          *map(str, objs), str(stub_obj), str(fat_obj)])
    return lib


if __name__ == "__main__":
    out = Path(sys.argv[1]) if len(sys.argv) > 1 else ROOT / "etalon_amd" / "_build" / "hoststub"
    print(build(out))
