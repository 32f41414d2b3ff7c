"""Device arrays in memory that the HIP library allocates: the host code's GPU-resident state.

A :class:`DeviceArray` is a C-contiguous typed array in HBM, allocated and freed through the
library's C ABI (``bt_device_alloc`` / ``bt_device_free``); the kernels' device form takes
it wherever it takes a GPU array (:mod:`beamtracer_amd.kernels`), on the null stream. The
resident beam, the batched designs and the edge fields of the device-resident Vlasov loop keep
their state in these arrays; the few operations that state needs between traces -- fills,
copies, the largest magnitude of a field, the trajectories with a given status, gathers of
columns and plane records, scatters of statuses -- are the library's own kernels
(``bt_memory.hip``). PyTorch is not involved; it remains only the autograd wrapper's
(:mod:`beamtracer_amd.adjoint`).
"""

from __future__ import annotations

import ctypes
import math
from collections.abc import Sequence

import numpy as np

from beamtracer_amd import _lib

_DTYPES = (np.dtype(np.float64), np.dtype(np.int64), np.dtype(np.int32), np.dtype(np.uint8))


class _Block:
    """One allocation of the library, freed when the last array on it goes."""

    __slots__ = ("ptr", "nbytes")

    def __init__(self, nbytes: int) -> None:
        self.ptr = 0
        self.nbytes = int(nbytes)
        out = ctypes.c_void_p()
        _lib.check(_lib.lib().bt_device_alloc(int(nbytes), ctypes.byref(out)), "bt_device_alloc")
        self.ptr = int(out.value or 0)

    def __del__(self) -> None:
        if getattr(self, "ptr", 0):
            try:
                _lib.lib().bt_device_free(ctypes.c_void_p(self.ptr))
            except Exception:  # interpreter shutdown
                pass
            self.ptr = 0


class DeviceArray:
    """A C-contiguous array of ``dtype`` (float64, int64, int32 or uint8) in device memory.

    ``reshape`` and integer or slice indexing along the first axis give views of the same
    memory; everything else copies.
    """

    is_cuda = True

    def __init__(self, shape: int | Sequence[int], dtype=np.float64, *, _block: _Block | None = None,
                 _offset: int = 0) -> None:  # fmt: skip
        self.shape = (int(shape),) if np.isscalar(shape) else tuple(int(d) for d in shape)
        self.dtype = np.dtype(dtype)
        if self.dtype not in _DTYPES:
            raise TypeError(f"unsupported dtype {self.dtype}")
        if any(d < 0 for d in self.shape):
            raise ValueError(f"negative dimension in {self.shape}")
        self._block = _block if _block is not None else _Block(self.nbytes)
        self._offset = int(_offset)

    # -- shape ---------------------------------------------------------------------------

    @property
    def size(self) -> int:
        return math.prod(self.shape)

    @property
    def nbytes(self) -> int:
        return self.size * self.dtype.itemsize

    @property
    def ndim(self) -> int:
        return len(self.shape)

    def numel(self) -> int:
        return self.size

    def dim(self) -> int:
        return self.ndim

    def element_size(self) -> int:
        return self.dtype.itemsize

    def is_contiguous(self) -> bool:
        return True

    def __len__(self) -> int:
        return self.shape[0]

    def data_ptr(self) -> int:
        """The device address of the first element (0 for an empty array)."""
        return self._block.ptr + self._offset if self.size else 0

    def __repr__(self) -> str:
        return f"DeviceArray(shape={self.shape}, dtype={self.dtype})"

    def reshape(self, *shape) -> DeviceArray:
        """The same memory as ``shape`` (one entry may be -1)."""
        shape = tuple(shape[0]) if len(shape) == 1 and not np.isscalar(shape[0]) else tuple(shape)
        if shape.count(-1) == 1:
            known = math.prod(d for d in shape if d != -1)
            shape = tuple(self.size // known if d == -1 else d for d in shape)
        if math.prod(shape) != self.size:
            raise ValueError(f"cannot view {self.shape} as {shape}")
        return DeviceArray(shape, self.dtype, _block=self._block, _offset=self._offset)

    view = reshape

    def __getitem__(self, key) -> DeviceArray:
        """A row (int) or a run of rows (slice with step 1) along the first axis."""
        rows = self.shape[0]
        row = math.prod(self.shape[1:]) * self.dtype.itemsize
        if isinstance(key, (int, np.integer)):
            k = int(key) + (rows if key < 0 else 0)
            if not 0 <= k < rows:
                raise IndexError(key)
            return DeviceArray(self.shape[1:], self.dtype, _block=self._block, _offset=self._offset + k * row)
        if isinstance(key, slice):
            lo, hi, step = key.indices(rows)
            if step != 1:
                raise IndexError("only contiguous slices")
            hi = max(hi, lo)
            return DeviceArray((hi - lo, *self.shape[1:]), self.dtype, _block=self._block,
                               _offset=self._offset + lo * row)  # fmt: skip
        raise IndexError("a DeviceArray takes an int or a slice along its first axis")

    # -- creation and copies ---------------------------------------------------------------

    @classmethod
    def empty(cls, shape, dtype=np.float64) -> DeviceArray:
        return cls(shape, dtype)

    @classmethod
    def zeros(cls, shape, dtype=np.float64) -> DeviceArray:
        return cls(shape, dtype).zero_()

    @classmethod
    def full(cls, shape, value, dtype=np.float64) -> DeviceArray:
        return cls(shape, dtype).fill_(value)

    @classmethod
    def from_numpy(cls, array) -> DeviceArray:
        """A device copy of a host array."""
        a = np.ascontiguousarray(array)
        out = cls(a.shape, a.dtype)
        return out.copy_(a)

    def numpy(self) -> np.ndarray:
        """A host copy (after the work queued before it)."""
        out = np.empty(self.shape, self.dtype)
        if self.size:
            rc = _lib.lib().bt_copy_to_host(out.ctypes.data, ctypes.c_void_p(self.data_ptr()), self.nbytes, None)
            _lib.check(rc, "bt_copy_to_host")
        return out

    def copy_(self, src) -> DeviceArray:
        """Fill from ``src``: a host array (copied now) or a device array of the same size
        (queued on the null stream)."""
        if isinstance(src, DeviceArray):
            if src.size != self.size or src.dtype != self.dtype:
                raise ValueError(f"cannot copy {src} into {self}")
            rc = _lib.lib().bt_copy_on_device(ctypes.c_void_p(self.data_ptr()), ctypes.c_void_p(src.data_ptr()),
                                              self.nbytes, None)  # fmt: skip
            _lib.check(rc, "bt_copy_on_device")
            return self
        a = np.ascontiguousarray(src, dtype=self.dtype)
        if a.size != self.size:
            raise ValueError(f"cannot copy {a.shape} into {self.shape}")
        if self.size:
            rc = _lib.lib().bt_copy_to_device(ctypes.c_void_p(self.data_ptr()), a.ctypes.data, self.nbytes, None)
            _lib.check(rc, "bt_copy_to_device")
        return self

    def clone(self) -> DeviceArray:
        return DeviceArray(self.shape, self.dtype).copy_(self)

    def fill_(self, value) -> DeviceArray:
        """Every element ``value`` (queued on the null stream)."""
        if not self.size:
            return self
        L = _lib.lib()
        ptr = ctypes.c_void_p(self.data_ptr())
        if self.dtype == np.int64:
            _lib.check(L.bt_fill_i64(ptr, int(value), self.size, None), "bt_fill_i64")
        elif self.dtype == np.float64:
            _lib.check(L.bt_fill_f64(ptr, float(value), self.size, None), "bt_fill_f64")
        else:
            self.copy_(np.full(self.shape, value, self.dtype))
        return self

    def zero_(self) -> DeviceArray:
        return self.fill_(0)


def synchronize() -> None:
    """Wait for the work queued on the null stream."""
    _lib.check(_lib.lib().bt_synchronize(None), "bt_synchronize")


def abs_max(a: DeviceArray) -> float:
    """max |a| (NaN if any element is NaN, 0 for an empty array)."""
    _check(a, np.float64, "a")
    out = ctypes.c_double()
    _lib.check(_lib.lib().bt_abs_max_f64(ctypes.c_void_p(a.data_ptr()), a.size, ctypes.byref(out), None),
               "bt_abs_max_f64")  # fmt: skip
    return float(out.value)


def max_int(a: DeviceArray, empty: int = 0) -> int:
    """max of an int64 array (``empty`` for an empty one)."""
    _check(a, np.int64, "a")
    if not a.size:
        return int(empty)
    out = ctypes.c_int64()
    _lib.check(_lib.lib().bt_max_i64(ctypes.c_void_p(a.data_ptr()), a.size, ctypes.byref(out), None), "bt_max_i64")
    return int(out.value)


def select_equal(a: DeviceArray, value: int, scratch: DeviceArray | None = None) -> np.ndarray:
    """The indices (ascending, on the host) where the int64 array ``a`` equals ``value``;
    ``scratch`` (int64, at least ``a.size``) is used instead of a new allocation."""
    _check(a, np.int64, "a")
    if scratch is None or scratch.size < a.size or scratch.dtype != np.int64:
        scratch = DeviceArray(max(a.size, 1), np.int64)
    count = ctypes.c_int64()
    rc = _lib.lib().bt_select_equal_i64(ctypes.c_void_p(a.data_ptr()), a.size, int(value),
                                        ctypes.c_void_p(scratch.data_ptr()), ctypes.byref(count), None)  # fmt: skip
    _lib.check(rc, "bt_select_equal_i64")
    return np.sort(scratch[: int(count.value)].numpy())


def gather_columns(src: DeviceArray, idx: np.ndarray) -> np.ndarray:
    """``src[:, idx]`` of a 2D float64 array, on the host."""
    _check(src, np.float64, "src")
    rows, n = src.shape
    idx = np.ascontiguousarray(idx, dtype=np.int64)
    if not idx.size:
        return np.zeros((rows, 0))
    d_idx = DeviceArray.from_numpy(idx)
    out = DeviceArray((rows, idx.size), np.float64)
    rc = _lib.lib().bt_gather_columns_f64(ctypes.c_void_p(src.data_ptr()), rows, n, ctypes.c_void_p(d_idx.data_ptr()),
                                          idx.size, ctypes.c_void_p(out.data_ptr()), None)  # fmt: skip
    _lib.check(rc, "bt_gather_columns_f64")
    return out.numpy()


def scatter(dst: DeviceArray, idx: np.ndarray, values: np.ndarray) -> None:
    """``dst[idx] = values`` for an int64 array (queued, then awaited)."""
    _check(dst, np.int64, "dst")
    idx = np.ascontiguousarray(idx, dtype=np.int64)
    values = np.ascontiguousarray(np.broadcast_to(values, idx.shape), dtype=np.int64)
    if not idx.size:
        return
    d_idx, d_val = DeviceArray.from_numpy(idx), DeviceArray.from_numpy(values)
    rc = _lib.lib().bt_scatter_i64(ctypes.c_void_p(dst.data_ptr()), dst.size, ctypes.c_void_p(d_idx.data_ptr()),
                                   ctypes.c_void_p(d_val.data_ptr()), idx.size, None)  # fmt: skip
    _lib.check(rc, "bt_scatter_i64")


def gather_records(hits: DeviceArray, plane: int, trajectory: np.ndarray, slot: np.ndarray) -> np.ndarray:
    """The records ``hits[trajectory[r], plane, slot[r]]`` ``(m, 8)`` on the host."""
    _check(hits, np.float64, "hits")
    n, n_planes, capacity, _ = hits.shape
    trajectory = np.ascontiguousarray(trajectory, dtype=np.int64)
    slot = np.ascontiguousarray(slot, dtype=np.int64)
    if not trajectory.size:
        return np.zeros((0, 8))
    d_k, d_c = DeviceArray.from_numpy(trajectory), DeviceArray.from_numpy(slot)
    out = DeviceArray((trajectory.size, 8), np.float64)
    rc = _lib.lib().bt_gather_records(ctypes.c_void_p(hits.data_ptr()), n, n_planes, capacity, int(plane),
                                      ctypes.c_void_p(d_k.data_ptr()), ctypes.c_void_p(d_c.data_ptr()),
                                      trajectory.size, ctypes.c_void_p(out.data_ptr()), None)  # fmt: skip
    _lib.check(rc, "bt_gather_records")
    return out.numpy()


def _check(a, dtype, name: str) -> None:
    if not isinstance(a, DeviceArray):
        raise TypeError(f"{name} must be a DeviceArray, got {type(a).__name__}")
    if a.dtype != np.dtype(dtype):
        raise TypeError(f"{name} must be {np.dtype(dtype)}, got {a.dtype}")
