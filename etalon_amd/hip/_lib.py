"""ctypes binding of libetalon_hip.so (declarations in include/etalon_hip.h).

The library is built in-tree by ``python -m etalon_amd.build``.  Loading it never
falls back to anything: a missing library or device raises RuntimeError.
"""

from __future__ import annotations

import ctypes as C
import os
import threading
from pathlib import Path

# The in-tree build (python -m etalon_amd.build); $ETALON_AMD_LIB names another build.
LIB_PATH = Path(os.environ.get("ETALON_AMD_LIB") or
                Path(__file__).resolve().parent.parent / "libetalon_hip.so")

F32, F64 = 0, 1
_lock = threading.Lock()
_lib = None
_device = None

vp = C.c_void_p
i64p = C.POINTER(C.c_int64)


class EtlTeDesc(C.Structure):
    _fields_ = [
        ("precision", C.c_int),
        ("nx", C.c_int),
        ("ny", C.c_int),
        ("db", C.c_double),
        ("inv_dx", C.c_double),
        ("inv_dy", C.c_double),
        ("field", (vp * 2) * 7),
        ("cb_ex", vp),
        ("cb_ey", vp),
        ("bx_e", vp), ("cx_e", vp), ("ikx_e", vp),
        ("bx_h", vp), ("cx_h", vp), ("ikx_h", vp),
        ("by_e", vp), ("cy_e", vp), ("iky_e", vp),
        ("by_h", vp), ("cy_h", vp), ("iky_h", vp),
    ]


class EtlTune(C.Structure):
    _fields_ = [
        ("band_rows", C.c_int),
        ("vec", C.c_int),
        ("use_graph", C.c_int),
        ("trace_chunk", C.c_int),
        ("temporal", C.c_int),
        ("waves", C.c_int),
        ("launch", C.c_int),
        ("uniform_cb", C.c_int),
    ]


class EtlAdjointDesc(C.Structure):
    _fields_ = [
        ("precision", C.c_int),
        ("nx", C.c_int),
        ("ny", C.c_int),
        ("db", C.c_double),
        ("inv_dx", C.c_double),
        ("inv_dy", C.c_double),
        ("lam", vp * 7),
        ("cb_ex", vp),
        ("cb_ey", vp),
        ("bx_e", vp), ("cx_e", vp), ("ikx_e", vp),
        ("bx_h", vp), ("cx_h", vp), ("ikx_h", vp),
        ("by_e", vp), ("cy_e", vp), ("iky_e", vp),
        ("by_h", vp), ("cy_h", vp), ("iky_h", vp),
        ("box_i0", C.c_int), ("box_j0", C.c_int), ("box_ni", C.c_int), ("box_nj", C.c_int),
        ("grad_ex", vp),
        ("grad_ey", vp),
        ("hist", vp),
        ("hist_rows", C.c_int),
    ]


class EtlTiming(C.Structure):
    _fields_ = [
        ("launch_ms", C.c_double),
        ("kernel_ms", C.c_double),
        ("total_ms", C.c_double),
        ("kernel_cells", C.c_int64),
        ("launches", C.c_int),
        ("depth", C.c_int),
        ("launch", C.c_int),
        ("uniform_cells", C.c_int64),
    ]


# name -> argtypes; every function returns int
_SIGNATURES = {
    "etl_last_error": [C.c_char_p, C.c_size_t],
    "etl_init": [C.c_int],
    "etl_device_info": [C.c_char_p, C.c_size_t],
    "etl_device_count": [C.POINTER(C.c_int)],
    "etl_malloc": [C.POINTER(vp), C.c_size_t],
    "etl_free": [vp],
    "etl_memcpy_h2d": [vp, vp, C.c_size_t],
    "etl_memcpy_d2h": [vp, vp, C.c_size_t],
    "etl_memcpy_d2d": [vp, vp, C.c_size_t],
    "etl_memcpy_d2d_lane": [vp, vp, C.c_size_t, C.c_int],
    "etl_lane_wait": [C.c_int, C.c_int],
    "etl_zero": [vp, C.c_size_t],
    "etl_copy_rate": [C.c_size_t, C.c_int, C.POINTER(C.c_float)],
    "etl_synchronize": [],
    "etl_update_h_te": [C.c_int, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                        C.c_double, C.c_double, C.c_double, C.c_int, C.c_int],
    "etl_update_e_te": [C.c_int, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                        C.c_double, C.c_double, C.c_int, C.c_int],
    "etl_accumulate_dft": [C.c_int, vp, vp, vp, vp, C.c_int, C.c_int],
    "etl_scatter_sub": [C.c_int, vp, C.c_int64, vp, vp, vp, C.c_int],
    "etl_gather": [C.c_int, vp, vp, C.c_int64, vp, vp, C.c_int],
    "etl_field_energy": [C.c_int, vp, vp, vp, vp, vp, C.c_int, C.c_int, vp],
    "etl_field_energy_noalloc": [C.c_int, vp, vp, vp, vp, vp, C.c_int, C.c_int, vp],
    "etl_downsample": [C.c_int, vp, C.c_int, C.c_int, C.c_int, C.c_int, vp],
    "etl_plan_create": [C.POINTER(EtlTeDesc), C.POINTER(EtlTune), C.POINTER(vp)],
    "etl_plan_destroy": [vp],
    "etl_plan_touch": [vp],
    "etl_plan_set_sources": [vp, C.c_int, C.POINTER(C.c_int), i64p, i64p],
    "etl_plan_source_table": [vp, C.c_int64, C.c_int, vp],
    "etl_plan_add_monitor": [vp, C.c_int, C.c_int, i64p, C.c_int, vp, C.POINTER(C.c_int)],
    "etl_plan_phasor_table": [vp, C.c_int, C.c_int64, C.c_int, vp],
    "etl_plan_create_batch": [C.POINTER(EtlTeDesc), C.c_int, C.POINTER(EtlTune), C.POINTER(vp)],
    "etl_plan_set_sources_sim": [vp, C.c_int, C.c_int, C.POINTER(C.c_int), i64p, i64p],
    "etl_plan_source_table_sim": [vp, C.c_int, C.c_int64, C.c_int, vp],
    "etl_plan_add_monitor_sim": [vp, C.c_int, C.c_int, C.c_int, i64p, C.c_int, vp,
                                 C.POINTER(C.c_int)],
    "etl_plan_phasor_table_sim": [vp, C.c_int, C.c_int, C.c_int64, C.c_int, vp],
    "etl_plan_share_phasors": [vp, C.c_int, C.c_int, C.c_int],
    "etl_plan_add_probe_sim": [vp, C.c_int, C.c_int, C.c_int, i64p, vp, C.c_int],
    "etl_adjoint_create": [C.POINTER(EtlAdjointDesc), C.POINTER(vp)],
    "etl_adjoint_destroy": [vp],
    "etl_adjoint_sources": [vp, C.c_int, C.POINTER(C.c_int), i64p, C.c_int, vp],
    "etl_adjoint_run": [vp, C.c_int64, C.c_int],
    "etl_adjplan_create": [C.POINTER(EtlAdjointDesc), C.c_int, C.POINTER(EtlTune), C.POINTER(vp)],
    "etl_adjplan_tune": [vp, C.POINTER(EtlTune)],
    "etl_adjplan_destroy": [vp],
    "etl_adjplan_sources": [vp, C.c_int, C.c_int, C.POINTER(C.c_int), i64p, C.c_int, vp],
    "etl_adjplan_run": [vp, C.c_int64, C.c_int],
    "etl_adjplan_run_rows": [vp, C.c_int64, C.c_int, C.c_int],
    "etl_adjplan_set_lane": [vp, C.c_int],
    "etl_adjplan_reset": [vp],
    "etl_coplan_create": [vp, vp, C.POINTER(vp)],
    "etl_coplan_destroy": [vp],
    "etl_coplan_run": [vp, C.c_int64, C.c_int, C.c_int, C.c_int64, C.c_int, C.c_int],
    "etl_coplan_sweep": [vp, C.c_int64, C.c_int],
    "etl_plan_set_checkpoints_sim": [vp, C.c_int, C.POINTER(vp), C.c_int],
    "etl_plan_set_checkpoint_window_sim": [vp, C.c_int, C.POINTER(vp), C.c_int, C.c_int,
                                           C.c_int, C.c_int, C.c_int],
    "etl_plan_checkpoint": [vp, C.c_int, C.c_int],
    "etl_plan_advance_checkpointed": [vp, C.c_int64, C.c_int64, C.c_int, C.POINTER(C.c_int)],
    "etl_adjplan_lambda": [vp, C.c_int, C.c_int, C.POINTER(vp)],
    "etl_plan_tune": [vp, C.POINTER(EtlTune)],
    "etl_plan_prepare": [vp],
    "etl_plan_advance": [vp, C.c_int64, C.c_int64, C.POINTER(C.c_int)],
    "etl_plan_advance_rows": [vp, C.c_int64, C.c_int64, C.c_int, C.POINTER(C.c_int)],
    "etl_plan_set_lane": [vp, C.c_int],
    "etl_plan_time": [vp, C.c_int64, C.c_int64, C.POINTER(EtlTiming), C.POINTER(C.c_int)],
}


def exported_symbols() -> list[str]:
    """The function names the C ABI declares (for the load/export test)."""
    return list(_SIGNATURES)


def _check_single_hip_runtime() -> None:
    """torch ships its own libamdhip64.so.7; the soname is shared with ROCm's, so only
    one copy must be mapped (whichever was loaded first serves both)."""
    try:
        maps = Path("/proc/self/maps").read_text()
    except OSError:
        return
    paths = {line.split()[-1] for line in maps.splitlines()
             if "libamdhip64" in line and line.split()[-1].startswith("/")}
    real = {os.path.realpath(p) for p in paths}
    if len(real) > 1:
        raise RuntimeError(f"two HIP runtimes are mapped in this process: {sorted(real)}")


def load(path: Path = LIB_PATH):
    """The ctypes library with signatures set (no device initialisation)."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not path.exists():
            raise RuntimeError(
                f"{path.name} is not built; run `python -m etalon_amd.build` (needs ROCm hipcc)"
            )
        lib = C.CDLL(str(path))
        for name, argtypes in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = argtypes
            fn.restype = C.c_int
        _check_single_hip_runtime()
        _lib = lib
        return lib


class HipError(RuntimeError):
    """A libetalon_hip call returned an error code."""

    def __init__(self, name: str, code: int, message: str):
        super().__init__(f"{name} failed ({code}): {message}")
        self.code = code


def last_error(lib) -> str:
    buf = C.create_string_buffer(2048)
    lib.etl_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def call(name: str, *args) -> None:
    lib = load()
    rc = getattr(lib, name)(*args)
    if rc != 0:
        raise HipError(name, rc, last_error(lib))


def device_count() -> int:
    lib = load()
    n = C.c_int(0)
    lib.etl_device_count(C.byref(n))
    return n.value


def init(device: int = 0):
    """Initialise the library on ``device`` (raises RuntimeError without a GPU)."""
    global _device
    lib = load()
    if _device is None:
        if device_count() == 0:
            raise RuntimeError("backend 'hip' found no HIP device")
        call("etl_init", device)
        _device = device
    elif _device != device:
        raise RuntimeError(f"libetalon_hip is bound to device {_device}, not {device}")
    return lib


def device_info() -> str:
    lib = init(_device or 0)
    buf = C.create_string_buffer(512)
    call("etl_device_info", buf, len(buf))
    return buf.value.decode()
