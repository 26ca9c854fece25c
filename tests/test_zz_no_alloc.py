"""etl_field_energy_noalloc, the variant of etl_field_energy that allocates nothing
(its partial sums in the library's fixed device array), through the C ABI: two calls
in a row — a 2024 x 2024 state, then a smaller 301 x 187 one — each equal bit for bit
to etl_field_energy on the same state, in float32 and float64; and a grid beyond the
fixed array's limit is refused before anything runs."""

import ctypes as C

import numpy as np
import pytest

from etalon_amd.hip import _lib
from etalon_amd.hip.backend import DeviceArray, _precision_of, to_device

pytestmark = pytest.mark.gpu


def _state(nx, ny, dtype, seed):
    rng = np.random.default_rng(seed)
    arrays = {
        "ex": rng.standard_normal((nx, ny + 1)), "ey": rng.standard_normal((nx + 1, ny)),
        "hz": rng.standard_normal((nx, ny)) / 377.0,
        "cb_ex": rng.uniform(1e-9, 1e-8, (nx, ny + 1)), "cb_ey": rng.uniform(1e-9, 1e-8, (nx + 1, ny)),
    }
    return {k: to_device(v.astype(dtype)) for k, v in arrays.items()}


def _energy(entry, s, nx, ny):
    out = DeviceArray((3,), np.float64)
    _lib.call(entry, _precision_of(s["hz"].dtype), *(C.c_void_p(s[k].ptr) for k in
              ("ex", "ey", "hz", "cb_ex", "cb_ey")), nx, ny, C.c_void_p(out.ptr))
    return out.get()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_noalloc_energy_equals_the_allocating_entry_point(hip, dtype):
    for k, (nx, ny) in enumerate([(2024, 2024), (301, 187)]):  # the second call is smaller
        print(f"etl_field_energy_noalloc {np.dtype(dtype).name} {nx} x {ny}", flush=True)
        s = _state(nx, ny, dtype, seed=k)
        want = _energy("etl_field_energy", s, nx, ny)
        got = _energy("etl_field_energy_noalloc", s, nx, ny)
        assert np.all(want > 0)
        assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), (nx, ny)
        exact = [float(np.sum(s["ex"].get().astype(np.float64) ** 2
                              / s["cb_ex"].get().astype(np.float64))),
                 float(np.sum(s["ey"].get().astype(np.float64) ** 2
                              / s["cb_ey"].get().astype(np.float64))),
                 float(np.sum(s["hz"].get().astype(np.float64) ** 2))]
        np.testing.assert_allclose(got, exact, rtol=1e-12, atol=0)


def test_noalloc_energy_refuses_grids_beyond_its_scratch(hip):
    """16384 x 16384 Ex nodes exceed the 268435456 the fixed array serves: ETL_EARG with
    a message, before any kernel is launched (the buffers are never read)."""
    s = _state(8, 8, np.float32, seed=3)
    out = DeviceArray((3,), np.float64)
    with pytest.raises(_lib.HipError) as err:
        _lib.call("etl_field_energy_noalloc", _lib.F32, *(C.c_void_p(s[k].ptr) for k in
                  ("ex", "ey", "hz", "cb_ex", "cb_ey")), 16384, 16384, C.c_void_p(out.ptr))
    assert err.value.code == -1 and "at most 268435456" in str(err.value)
