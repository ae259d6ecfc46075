"""Autotuner of the temporally blocked step kernels.

Every compiled shape of the two temporally blocked kernel families (LDS
tiles, ``"blocked"``; one wavefront per tile with the state in registers,
``"wave"``) gives bit-identical results, so the choice only changes the speed.
For a problem, ``tune`` runs a short window of time steps with every compiled
shape (tile rows, tile columns, steps per launch) on the problem's own
buffers, timed with events on the launch stream, and keeps the fastest.  The
choice is cached per problem key for the life of the process; the caller
restores whatever the trial steps wrote.
"""
from __future__ import annotations

import ctypes
from typing import Callable, Dict, List, Tuple

import torch

from . import _lib

Shape = Tuple[str, int, int, int]  # kernel family, tile_x, tile_z, block_steps

FAMILIES = (("blocked", _lib.KERNEL_BLOCKED), ("wave", _lib.KERNEL_WAVE))
_CACHE: Dict[tuple, Shape] = {}
TIMINGS: Dict[tuple, List[Tuple[Shape, float]]] = {}  # last tuning of each key, ms per window


def candidates(systems: int) -> List[Shape]:
    """Every compiled shape of both families for the given wave systems."""
    out = []
    for name, kid in FAMILIES:
        buf = (ctypes.c_int32 * (4 * 64))()
        n = _lib.call("trtm_block_configs", systems, kid, buf, 64)
        for i in range(min(n, 64)):
            tx, tz, _threads, k = buf[4 * i:4 * i + 4]
            out.append((name, int(tx), int(tz), int(k)))
    return out


FWI_PASSES = {"forward": 0, "history": 0, "adjoint": 1, "scattered": 2}


def fwi_candidates(systems: int, pass_name: str = "adjoint") -> List[Shape]:
    """Every compiled shape of the blocked FWI kernels of a pass ("forward",
    "history", "adjoint" or "scattered") that fits these systems."""
    buf = (ctypes.c_int32 * (4 * 64))()
    n = _lib.call("trtm_fwi_block_configs", systems, FWI_PASSES[pass_name], buf, 64)
    return [("fwi", int(buf[4 * i]), int(buf[4 * i + 1]), int(buf[4 * i + 3]))
            for i in range(min(n, 64))]


def tune(key: tuple, systems: int, trial: Callable[[Shape], None], repeats: int = 2,
         shapes: List[Shape] = None) -> Shape:
    """Fastest shape for ``key`` among ``shapes`` (default: both blocked
    step-kernel families); ``trial(shape)`` runs the timing window once with
    that shape on the current stream."""
    if key in _CACHE:
        return _CACHE[key]
    if not systems and shapes is None:
        # no wave system propagates (e.g. a backward pass with no data): no
        # step kernel runs, the library's default shape stands
        return ("blocked", 0, 0, 0)
    times = []
    for shape in (shapes if shapes is not None else candidates(systems)):
        trial(shape)  # first call: tile lists, caches
        best = None
        for _ in range(repeats):
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            trial(shape)
            e1.record()
            e1.synchronize()
            ms = e0.elapsed_time(e1)
            best = ms if best is None else min(best, ms)
        times.append((shape, best))
    times.sort(key=lambda x: x[1])
    TIMINGS[key] = times
    _CACHE[key] = times[0][0]
    return times[0][0]


def clear():
    _CACHE.clear()
    TIMINGS.clear()
