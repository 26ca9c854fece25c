"""HipBackend: the reference's backend protocol over libetalon_hip.so.

Protocol (etalon/fdtd/backends/__init__.py:1-5; CupyBackend cupy_backend.py:30-98):
``name``, ``device``, ``xp``, ``to_device``, ``to_host``, ``synchronize``,
``update_h_te(s)``, ``update_e_te(s)``, ``accumulate_dft(acc, field, index, phasors)``.

Device arrays are ``DeviceArray`` objects: a device pointer owned through the
library's allocator plus shape and dtype.  They support what the reference's
solver and harnesses do with CuPy arrays:

* ``a[i, j]`` with Python ints reads one value (one D2H copy);
* ``a[row]`` on a 2D array is a row view (phasor tables);
* ``a[i, j] -= delta`` with device index arrays is one scatter-subtract kernel
  (FDTDSolver._inject, etalon/fdtd/solver.py:83-89);
* ``np.asarray(a)`` / ``backend.to_host(a)`` copy to the host.

``xp`` is NumPy: host-side math on device arrays goes through ``__array__``.
"""

from __future__ import annotations

import ctypes as C
import weakref
from typing import Any

import numpy as np

from etalon_amd.hip import _lib

_REAL = {np.dtype(np.float32): _lib.F32, np.dtype(np.float64): _lib.F64}
_COMPLEX_REAL = {np.dtype(np.complex64): _lib.F32, np.dtype(np.complex128): _lib.F64}


def _precision_of(dtype: np.dtype) -> int:
    dtype = np.dtype(dtype)
    if dtype in _REAL:
        return _REAL[dtype]
    if dtype in _COMPLEX_REAL:
        return _COMPLEX_REAL[dtype]
    raise TypeError(f"libetalon_hip computes in float32/float64, not {dtype}")


class _Buffer:
    """One device allocation; freed when the last array using it goes away.

    ``writes`` counts host writes into it (``DeviceArray.set`` and item assignment,
    through any view): a fused plan records the counts of its state buffers and
    re-checks its preconditions when one has changed."""

    __slots__ = ("ptr", "nbytes", "writes", "__weakref__")

    def __init__(self, nbytes: int):
        lib = _lib.init(_lib._device or 0)
        p = C.c_void_p()
        _lib.call("etl_malloc", C.byref(p), max(int(nbytes), 1))
        self.ptr = p.value
        self.nbytes = int(nbytes)
        self.writes = 0
        weakref.finalize(self, _free, lib, self.ptr)


# Host writes into any device buffer so far (every ``_Buffer.writes`` bump adds one):
# a plan that checked its buffers at epoch e skips the per-buffer comparison while the
# epoch is still e.
_write_epoch = 0


def host_write_epoch() -> int:
    """Host writes into device buffers so far, over all buffers of the process."""
    return _write_epoch


def _free(lib, ptr):
    try:
        lib.etl_free(C.c_void_p(ptr))
    except Exception:  # interpreter shutdown
        pass


class DeviceArray:
    """A C-contiguous array on the HIP device."""

    def __init__(self, shape, dtype, buffer: _Buffer | None = None, offset: int = 0):
        self.shape = tuple(int(s) for s in np.atleast_1d(np.asarray(shape, dtype=np.int64))) \
            if not isinstance(shape, tuple) else tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self._buf = buffer if buffer is not None else _Buffer(self.nbytes)
        self._offset = int(offset)

    # -- basic properties -------------------------------------------------
    @property
    def size(self) -> int:
        return int(np.prod(self.shape, dtype=np.int64))

    @property
    def ndim(self) -> int:
        return len(self.shape)

    @property
    def nbytes(self) -> int:
        return self.size * self.dtype.itemsize

    @property
    def ptr(self) -> int:
        return self._buf.ptr + self._offset

    def __len__(self) -> int:
        return self.shape[0]

    def __repr__(self) -> str:
        return f"DeviceArray(shape={self.shape}, dtype={self.dtype}, ptr=0x{self.ptr:x})"

    # -- host transfers ---------------------------------------------------
    def get(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            _lib.call("etl_memcpy_d2h", out.ctypes.data, C.c_void_p(self.ptr), self.nbytes)
        return out

    def rows(self, r0: int, r1: int) -> "DeviceArray":
        """A view of rows [r0, r1) along the first axis (contiguous in C order)."""
        if self.ndim < 1 or not 0 <= r0 <= r1 <= self.shape[0]:
            raise IndexError(f"rows [{r0}, {r1}) out of range for shape {self.shape}")
        stride = int(np.prod(self.shape[1:], dtype=np.int64)) * self.dtype.itemsize
        return DeviceArray((r1 - r0, *self.shape[1:]), self.dtype, self._buf,
                           self._offset + r0 * stride)

    @property
    def writes(self) -> int:
        """Host writes into this array's buffer so far (any view of it)."""
        return self._buf.writes

    def set(self, host: Any) -> None:
        host = np.ascontiguousarray(host, dtype=self.dtype)
        if host.shape != self.shape:
            host = np.ascontiguousarray(np.broadcast_to(host, self.shape))
        global _write_epoch
        self._buf.writes += 1
        _write_epoch += 1
        if self.nbytes:
            _lib.call("etl_memcpy_h2d", C.c_void_p(self.ptr), host.ctypes.data, self.nbytes)

    def __array__(self, dtype=None, copy=None):
        a = self.get()
        return a if dtype is None else a.astype(dtype)

    def astype(self, dtype) -> np.ndarray:
        """A host copy in ``dtype`` (``xp`` is NumPy: host code that does arithmetic on
        device arrays, such as etalon/studio/runner.py:258-280, works on the copy;
        etalon_amd.observe reduces on the device instead)."""
        return self.get().astype(dtype)

    def copy(self) -> "DeviceArray":
        out = DeviceArray(self.shape, self.dtype)
        _lib.call("etl_memcpy_d2d", C.c_void_p(out.ptr), C.c_void_p(self.ptr), self.nbytes)
        return out

    def fill_zero(self) -> None:
        _lib.call("etl_zero", C.c_void_p(self.ptr), self.nbytes)

    # -- indexing ---------------------------------------------------------
    def __getitem__(self, key):
        if isinstance(key, tuple) and len(key) == 2 and all(isinstance(k, DeviceArray)
                                                             for k in key):
            return _FancyRef(self, key[0], key[1])
        if isinstance(key, (int, np.integer)) and self.ndim >= 2:
            row = int(key)
            if row < 0:
                row += self.shape[0]
            if not 0 <= row < self.shape[0]:
                raise IndexError(f"row {key} out of range for shape {self.shape}")
            inner = self.shape[1:]
            stride = int(np.prod(inner, dtype=np.int64)) * self.dtype.itemsize
            return DeviceArray(inner, self.dtype, self._buf, self._offset + row * stride)
        if (isinstance(key, tuple) and len(key) == self.ndim
                and all(isinstance(k, (int, np.integer)) for k in key)):
            idx = tuple(int(k) + (s if int(k) < 0 else 0) for k, s in zip(key, self.shape))
            for k, s in zip(idx, self.shape):
                if not 0 <= k < s:
                    raise IndexError(f"index {key} out of range for shape {self.shape}")
            flat = int(np.ravel_multi_index(idx, self.shape))
            out = np.empty((), self.dtype)
            _lib.call("etl_memcpy_d2h", out.ctypes.data,
                      C.c_void_p(self.ptr + flat * self.dtype.itemsize), self.dtype.itemsize)
            return out[()]
        # Anything else: read the whole array back and index on the host.
        return self.get()[key]

    def __setitem__(self, key, value):
        if isinstance(value, _FancyRef) and value.applied and value.parent is self:
            return  # `a[i, j] -= d` already applied in place by _FancyRef.__isub__
        if key is Ellipsis or (isinstance(key, slice) and key == slice(None)):
            self.set(value)
            return
        host = self.get()
        host[key] = np.asarray(value)
        self.set(host)


class _FancyRef:
    """``parent[i, j]`` with device index arrays, pending an in-place operation."""

    def __init__(self, parent: DeviceArray, i: DeviceArray, j: DeviceArray):
        self.parent, self.i, self.j = parent, i, j
        self.applied = False

    def __isub__(self, delta):
        parent = self.parent
        if not isinstance(delta, DeviceArray):
            delta = to_device(np.asarray(delta, dtype=parent.dtype).reshape(-1))
        if delta.dtype != parent.dtype:
            raise TypeError(f"delta dtype {delta.dtype} != field dtype {parent.dtype}")
        n = self.i.size
        if delta.size != n or self.j.size != n:
            raise ValueError("index arrays and delta must have the same length")
        pitch = parent.shape[1]
        _lib.call("etl_scatter_sub", _precision_of(parent.dtype), C.c_void_p(parent.ptr),
                  pitch, C.c_void_p(self.i.ptr), C.c_void_p(self.j.ptr), C.c_void_p(delta.ptr), n)
        self.applied = True
        return self

    def materialize(self) -> DeviceArray:
        out = DeviceArray((self.i.size,), self.parent.dtype)
        _lib.call("etl_gather", _precision_of(self.parent.dtype), C.c_void_p(out.ptr),
                  C.c_void_p(self.parent.ptr), self.parent.shape[1], C.c_void_p(self.i.ptr),
                  C.c_void_p(self.j.ptr), self.i.size)
        return out

    def __array__(self, dtype=None, copy=None):
        return np.asarray(self.materialize(), dtype=dtype)


def to_device(a: Any) -> DeviceArray:
    host = np.ascontiguousarray(a)
    if host.dtype == np.int32:
        host = host.astype(np.int64)
    out = DeviceArray(host.shape, host.dtype)
    out.set(host)
    return out


def empty(shape, dtype) -> DeviceArray:
    return DeviceArray(tuple(shape), dtype)


def zeros(shape, dtype) -> DeviceArray:
    out = DeviceArray(tuple(shape), dtype)
    out.fill_zero()
    return out


class HipBackend:
    """One MI355X (the library's device): the protocol over hand-written HIP kernels."""

    name = "hip"
    xp = np

    def __init__(self, precision: str = "single", device: int = 0) -> None:
        if precision not in ("single", "double"):
            raise ValueError(f"precision must be 'single' or 'double', not {precision!r}")
        _lib.init(device)
        self.precision = precision
        self.device = f"hip:{device} {_lib.device_info()}"

    # -- memory -----------------------------------------------------------
    def to_device(self, a: Any) -> DeviceArray:
        return to_device(a)

    def to_host(self, a: Any) -> np.ndarray:
        if isinstance(a, DeviceArray):
            return a.get()
        return np.asarray(a)

    def synchronize(self) -> None:
        _lib.call("etl_synchronize")

    def zeros(self, shape, dtype) -> DeviceArray:
        return zeros(shape, dtype)

    # -- the protocol's compute --------------------------------------------
    def update_h_te(self, s) -> None:
        """Hz^(n−½) → Hz^(n+½) from Ex^n, Ey^n (numpy_backend.py:41-49)."""
        p = _precision_of(s.hz.dtype)
        _lib.call("etl_update_h_te", p, *(C.c_void_p(getattr(s, n).ptr) for n in (
            "hz", "ex", "ey", "psi_hz_x", "psi_hz_y", "bx_h", "cx_h", "ikx_h", "by_h", "cy_h",
            "iky_h")), float(s.db), float(s.inv_dx), float(s.inv_dy), int(s.nx), int(s.ny))

    def update_e_te(self, s) -> None:
        """Ex^n, Ey^n → Ex^(n+1), Ey^(n+1) from Hz^(n+½) (numpy_backend.py:51-61)."""
        p = _precision_of(s.ex.dtype)
        _lib.call("etl_update_e_te", p, *(C.c_void_p(getattr(s, n).ptr) for n in (
            "ex", "ey", "hz", "psi_ex_y", "psi_ey_x", "cb_ex", "cb_ey", "bx_e", "cx_e", "ikx_e",
            "by_e", "cy_e", "iky_e")), float(s.inv_dx), float(s.inv_dy), int(s.nx), int(s.ny))

    def accumulate_dft(self, acc: DeviceArray, field: DeviceArray, index: DeviceArray,
                       phasors: DeviceArray) -> None:
        """acc[f, k] += field.flat[index[k]] · phasors[f] (numpy_backend.py:63-67)."""
        n_f, n_nodes = acc.shape
        if np.dtype(index.dtype) != np.int64:
            raise TypeError("index must be int64")
        _lib.call("etl_accumulate_dft", _precision_of(field.dtype), C.c_void_p(acc.ptr),
                  C.c_void_p(field.ptr), C.c_void_p(index.ptr), C.c_void_p(phasors.ptr),
                  int(n_f), int(n_nodes))

    def inject(self, field: DeviceArray, i: DeviceArray, j: DeviceArray, delta: Any) -> None:
        """field[i, j] -= delta, one kernel (FDTDSolver._inject, solver.py:83-89)."""
        ref = _FancyRef(field, i, j)
        ref -= delta
