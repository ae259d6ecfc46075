"""Captured time stepping: a range of blocked steps recorded once into a HIP
graph and replayed (``trtm_plan_create / _launch / _destroy``).

The reference issues 6-7 Taichi kernels per step from a Python loop
(forward_modeling.py:539-572, backward_modeling.py:539-576).  A batch call of
the library already replaces that loop with one launch per block of K steps
for every shot; a ``StepPlan`` records that launch sequence for one range
[it_begin, it_end) into a graph, so a caller can replay it with one graph
launch, embed it in a graph of its own, or rerun the same steps after
changing buffer contents (state, wavelets, data, medium) in place.

The plan keeps the descriptors' pointers: every buffer they name, and the
per-shot and batch workspaces (which hold the tile lists and shot tables),
must stay alive and untouched until ``close``.  This object holds references
to the tensors it is given for that purpose.
"""
from __future__ import annotations

import ctypes
from typing import Sequence

import torch

from . import _lib
from .field import device, stream_handle


class StepPlan:
    """Steps [it_begin, it_end) of the shots ``descs`` (``StepDesc``
    instances sharing one medium), forward (with ``save``) or backward,
    captured once.  ``keep`` holds whatever tensors the descriptors point at,
    so they live as long as the plan."""

    def __init__(self, descs: Sequence[_lib.StepDesc], it_begin: int, it_end: int,
                 forward: bool, save: bool = False, keep=()):
        n = len(descs)
        if n < 1:
            raise ValueError("no shots")
        self._arr = (_lib.StepDesc * n)()
        for q, d in enumerate(descs):
            ctypes.memmove(ctypes.byref(self._arr[q]), ctypes.byref(d), ctypes.sizeof(_lib.StepDesc))
        nbytes = int(_lib.load().trtm_batch_workspace_bytes(n))
        self._bws = torch.empty(nbytes, dtype=torch.uint8, device=device())
        self._keep = list(keep)
        self.range = (int(it_begin), int(it_end))
        self.forward = bool(forward)
        handle = ctypes.c_void_p()
        _lib.call("trtm_plan_create", self._arr, n, ctypes.c_void_p(self._bws.data_ptr()), nbytes,
                  int(it_begin), int(it_end), int(bool(forward)), int(bool(save)),
                  stream_handle(), ctypes.byref(handle))
        self._handle = handle

    def launch(self):
        """Replay the steps on the current stream."""
        if self._handle is None:
            raise RuntimeError("plan is closed")
        _lib.call("trtm_plan_launch", self._handle, stream_handle())

    def close(self):
        """Release the graph (after the work launched so far has finished)."""
        if self._handle is not None:
            torch.cuda.synchronize()
            _lib.call("trtm_plan_destroy", self._handle)
            self._handle = None
            self._keep = []

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            if self._handle is not None and _lib._LIB is not None:
                torch.cuda.synchronize()
                _lib._LIB.trtm_plan_destroy(self._handle)
                self._handle = None
        except Exception:
            pass
