"""Reading one shot of the reference's field-data npz format.

A file holds ``distance`` (receiver positions along the line, m), ``source_x``
(source position, m) and the three-component records ``x``, ``y``, ``z``, each
(receivers, samples) at 100 kHz.  ``load_shot`` subsamples receivers and time
and picks the source trace exactly as the reference's driver does
(examples/example_2nd_step.py:68-69, 109-133), giving the arrays its
``ReverseTimeMigration`` fluent calls take (set_observed_data, set_source,
set_receivers, set_frequency).  Files are read with ``allow_pickle=False``.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Union

import numpy as np


@dataclass
class Shot:
    observed_u: np.ndarray  # (receivers, samples) float32
    observed_v: np.ndarray
    observed_w: np.ndarray
    source_u: np.ndarray    # (1, samples) float32: the trace nearest the source
    source_v: np.ndarray
    source_w: np.ndarray
    source_x: float
    distance: np.ndarray    # receiver positions
    fs: float


def source_channel(distance, source_x) -> List[int]:
    """Index (as a one-element list) of the receiver nearest the source
    (example_2nd_step.py:68-69)."""
    return [int(np.argmin(np.abs(np.asarray(distance) - source_x)))]


def load_shot(npz: Union[str, dict], num_receivers: int, sampling_freq: float, time_to: float,
              original_fs: float = 100000.0) -> Shot:
    """Receivers every len(distance) // num_receivers (at most num_receivers of
    them), samples every int(original_fs / sampling_freq) up to
    int(original_fs * time_to) (example_2nd_step.py:112-133)."""
    d = np.load(npz, allow_pickle=False) if isinstance(npz, str) else npz
    distance_all = np.asarray(d["distance"])
    step = max(1, len(distance_all) // num_receivers)
    distance = distance_all[::step][:num_receivers]
    source_x = float(d["source_x"])
    ch = source_channel(distance_all, source_x)
    fs = float(sampling_freq)
    down = max(1, int(original_fs / fs))
    n_orig = int(original_fs * time_to)
    rec = slice(None, None, step)
    tim = slice(None, n_orig, down)
    obs = [np.asarray(d[k])[rec, tim][:num_receivers].astype(np.float32) for k in ("x", "y", "z")]
    src = [np.asarray(d[k])[ch, tim].astype(np.float32) for k in ("x", "y", "z")]
    return Shot(*obs, *src, source_x, distance, fs)
