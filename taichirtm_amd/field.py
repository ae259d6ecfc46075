"""Device arrays with the part of the ti.field surface the reference path uses.

The reference builds its inputs as Taichi fields and moves data with
``from_numpy`` / ``to_numpy`` (rtm.py:904-972, forward_modeling.py:556-558,
580-582).  ``Field`` wraps a torch tensor in HIP device memory with the same
methods, so code written against the reference's field objects runs
unchanged.  Propagator arguments also accept plain numpy arrays and torch
tensors.
"""
from __future__ import annotations

import numpy as np
import torch

f32 = np.float32
i32 = np.int32


def device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("taichirtm_amd needs a HIP device (MI355X); none is visible")
    return torch.device("cuda", torch.cuda.current_device())


def free_device_bytes() -> int:
    """Device memory that new tensors can use: what the device reports free
    plus what torch's caching allocator holds without using (blocks freed by
    earlier batches stay cached and are reused)."""
    free, _ = torch.cuda.mem_get_info()
    return int(free + torch.cuda.memory_reserved() - torch.cuda.memory_allocated())


def stream_handle():
    return ctypes_void(torch.cuda.current_stream().cuda_stream)


def ctypes_void(x):
    import ctypes
    return ctypes.c_void_p(int(x)) if x else None


_TORCH = {np.dtype(np.float32): torch.float32, np.dtype(np.int32): torch.int32,
          np.dtype(np.float64): torch.float64}


class Field:
    """A device array: ``shape``, ``dtype``, ``from_numpy``, ``to_numpy``, ``fill``."""

    def __init__(self, dtype=f32, shape=None, tensor=None):
        if tensor is None:
            tensor = torch.zeros(tuple(shape), dtype=_TORCH[np.dtype(dtype)], device=device())
        self._t = tensor

    @property
    def tensor(self) -> torch.Tensor:
        return self._t

    @property
    def shape(self):
        return tuple(self._t.shape)

    @property
    def dtype(self):
        return np.dtype(str(self._t.dtype).replace("torch.", ""))

    def from_numpy(self, arr):
        a = np.ascontiguousarray(arr, dtype=self.dtype)
        if a.shape != self.shape:
            raise ValueError("shape %s does not match field shape %s" % (a.shape, self.shape))
        self._t.copy_(torch.from_numpy(a))

    def to_numpy(self):
        return self._t.detach().cpu().numpy()

    def to_torch(self):
        return self._t

    def fill(self, value):
        self._t.fill_(value)

    def data_ptr(self):
        return self._t.data_ptr()


class SnapshotField(Field):
    """Snapshots stored snapshot-major (num_snaps, nx, nz) so each save is one
    coalesced sweep; presented with the reference's logical shape
    (nx, nz, num_snaps) (forward_modeling.py:533-535)."""

    @property
    def shape(self):
        s, nx, nz = self._t.shape
        return (nx, nz, s)

    def to_numpy(self):
        return np.ascontiguousarray(self._t.detach().cpu().numpy().transpose(1, 2, 0))

    def from_numpy(self, arr):
        a = np.ascontiguousarray(np.asarray(arr, dtype=self.dtype).transpose(2, 0, 1))
        if a.shape != tuple(self._t.shape):
            raise ValueError("snapshot shape mismatch")
        self._t.copy_(torch.from_numpy(a))


class ZeroField(Field):
    """A field of a wave system that never leaves zero in this run: reads
    as zeros without holding device memory."""

    def __init__(self, shape, dtype=f32):
        self._shape = tuple(shape)
        self._dtype = np.dtype(dtype)
        self._t = None

    @property
    def tensor(self):
        return None

    @property
    def shape(self):
        return self._shape

    @property
    def dtype(self):
        return self._dtype

    def to_numpy(self):
        return np.zeros(self._shape, self._dtype)

    def from_numpy(self, arr):
        raise ValueError("this field belongs to an inactive wave system")

    def data_ptr(self):
        return 0


def field(dtype=f32, shape=None):
    """Counterpart of ``ti.field(dtype=..., shape=...)``: a zeroed device field."""
    return Field(dtype, shape)


def as_tensor(x, dtype=f32, shape=None, name="field") -> torch.Tensor:
    """Accept a Field, torch tensor or array-like; return a contiguous device
    tensor of ``dtype`` (no copy when it already is one)."""
    tdt = _TORCH[np.dtype(dtype)]
    if isinstance(x, ZeroField):
        t = torch.zeros(x.shape, dtype=tdt, device=device())
    elif isinstance(x, Field):
        t = x.tensor
    elif isinstance(x, torch.Tensor):
        t = x
    else:
        t = torch.from_numpy(np.ascontiguousarray(np.asarray(x), dtype=dtype))
    if t.device.type != "cuda" or t.dtype != tdt or not t.is_contiguous():
        t = t.to(device=device(), dtype=tdt).contiguous()
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError("%s has shape %s, expected %s" % (name, tuple(t.shape), tuple(shape)))
    return t


def as_host(x, dtype=None) -> np.ndarray:
    if isinstance(x, Field):
        a = x.to_numpy()
    elif isinstance(x, torch.Tensor):
        a = x.detach().cpu().numpy()
    else:
        a = np.asarray(x)
    return a if dtype is None else np.ascontiguousarray(a, dtype=dtype)
