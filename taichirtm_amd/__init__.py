"""taichirtm_amd — MI355X-native elastic FD / RTM hot path of TaichiRTM.

Keeps the reference package's public names (src/__init__.py:24-55): the
propagators, the RTM orchestrator and the imaging utilities, with the time
stepping in hand-written HIP kernels for gfx950 behind a C ABI
(include/taichirtm_hip.h).
"""
from .backward_modeling import BackwardModeling
from .field import Field, field
from .forward_modeling import ForwardModeling
from .imaging import (
    compute_display_limits,
    create_visualization_data,
    load_rtm_results,
    prepare_image_data,
    save_stacked_results,
    stack_rtm_images,
)
from .rtm import ReverseTimeMigration, init_taichi, reset_taichi

__version__ = "2.0.0"

__all__ = [
    "ReverseTimeMigration",
    "init_taichi",
    "reset_taichi",
    "ForwardModeling",
    "BackwardModeling",
    "load_rtm_results",
    "stack_rtm_images",
    "compute_display_limits",
    "prepare_image_data",
    "create_visualization_data",
    "save_stacked_results",
    "Field",
    "field",
]
