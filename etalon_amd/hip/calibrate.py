"""Copy of a known size, for calibrating hardware byte counters.

    python -m etalon_amd.hip.calibrate [--mb 256] [--reps 5]

Copies an ``--mb`` MiB device buffer into another ``--reps`` times with
etl_memcpy_d2d (the HIP runtime's copy kernel) and prints one JSON line with the
bytes each copy must read and write and the measured rate.  Run under
``rocprofv3 --pmc FETCH_SIZE`` (then ``WRITE_SIZE``): the counter's value per
copy dispatch divided by the known bytes is the scale factor for the same counter
on the stencil kernels.
"""

from __future__ import annotations

import argparse
import ctypes as C
import json
import time

import numpy as np

from etalon_amd.hip import _lib
from etalon_amd.hip.backend import HipBackend, zeros


def copy_rate(mb: int, reps: int) -> dict:
    """Copy an ``mb`` MiB device buffer into another, once untimed and ``reps`` times
    timed: with the HIP runtime's copy (wall clock between device synchronisations)
    and with the library's 16-byte streaming copy kernel (HIP events)."""
    backend = HipBackend("single")
    n = mb << 20
    a = zeros((n // 4,), np.float32)
    b = zeros((n // 4,), np.float32)
    _lib.call("etl_memcpy_d2d", C.c_void_p(b.ptr), C.c_void_p(a.ptr), n)
    backend.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        _lib.call("etl_memcpy_d2d", C.c_void_p(b.ptr), C.c_void_p(a.ptr), n)
    backend.synchronize()
    dt = (time.perf_counter() - t0) / reps
    del a, b
    ms = C.c_float()
    _lib.call("etl_copy_rate", n, reps, C.byref(ms))
    return {"copy_bytes_read": n, "copy_bytes_written": n, "dispatches": reps + 1,
            "ms_per_copy": round(dt * 1e3, 4), "gbs_read_plus_write": round(2 * n / dt / 1e9, 1),
            "stream_copy_ms": round(ms.value, 4),
            "stream_copy_gbs": round(2 * n / (ms.value * 1e-3) / 1e9, 1)}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--mb", type=int, default=256)
    ap.add_argument("--reps", type=int, default=5)
    args = ap.parse_args()
    print(json.dumps(copy_rate(args.mb, args.reps)), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
