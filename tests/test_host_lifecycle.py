"""Host-side lifecycle of libetalon_hip under AddressSanitizer, without a GPU.

The library's host code is built against a HIP runtime stub (tests/hoststub: device
memory is host heap, kernels do not run, streams/events/graphs are heap objects freed
on destruction) and driven through the package the way the benchmarks drive it
(tests/hoststub/lifecycle_workload.py): the autotuner's create/time/close loop over
every compiled shape and launch mode (including the multi-stream graphs of launch mode
3), batch and sequential runs, table refills, event timing, adjoint sessions, stale
tapes, co-plans and the batched adjoint.  A use after free, an out-of-bounds host
access, a graph exec replayed after its destruction or an event or stream used after
its plan tore it down aborts the run with a report.
"""

import os
import subprocess
import sys
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent


@pytest.fixture(scope="module")
def stub(tmp_path_factory):
    """(library, environment) of the stubbed, sanitised build."""
    sys.path.insert(0, str(HERE / "hoststub"))
    try:
        import build as hoststub
    finally:
        sys.path.pop(0)
    if not hoststub.available():
        pytest.skip("ROCm hipcc/clang with an AddressSanitizer runtime is not installed")
    out = tmp_path_factory.mktemp("hoststub")
    lib = hoststub.build(out)
    env = dict(os.environ, ETALON_AMD_LIB=str(lib), LD_PRELOAD=str(hoststub.asan_runtime()),
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               ETALON_AMD_TUNE_CACHE=str(out / "tune.json"))
    return lib, env


def _run(env, args, cwd):
    return subprocess.run([sys.executable, *args], env=env, cwd=cwd, stdout=subprocess.PIPE,
                          stderr=subprocess.STDOUT, text=True, timeout=900)


def test_host_lifecycle_under_asan(stub, tmp_path):
    _, env = stub
    r = _run(env, [str(HERE / "hoststub" / "lifecycle_workload.py")], tmp_path)
    assert "AddressSanitizer" not in r.stdout, r.stdout[-6000:]
    assert r.returncode == 0, r.stdout[-6000:]
    assert "lifecycle: done" in r.stdout


USE_AFTER_DESTROY = """
import sys, ctypes as C
sys.path.insert(0, {root!r})
from etalon_amd.solver import FDTDSolver
from etalon_amd.workloads import throughput_sim
from etalon_amd.hip import _lib
solver = FDTDSolver(throughput_sim("single", 200))
solver.advance(200)
plan = solver._plan
handle = plan.handle
_lib.call("etl_plan_destroy", handle)
parity = C.c_int()
print("advancing a destroyed plan", flush=True)
_lib.load().etl_plan_advance(handle, C.c_int64(200), C.c_int64(96), C.byref(parity))
"""


def test_harness_reports_a_use_after_destroy(stub, tmp_path):
    """The harness's own check: advancing a plan after etl_plan_destroy (a caller's
    mistake the C ABI cannot see) is reported as a heap-use-after-free."""
    _, env = stub
    r = _run(env, ["-c", USE_AFTER_DESTROY.format(root=str(HERE.parent))], tmp_path)
    assert "advancing a destroyed plan" in r.stdout, r.stdout[-4000:]
    assert r.returncode != 0
    assert "heap-use-after-free" in r.stdout, r.stdout[-4000:]
