"""bismit_amd: MI355X-native engine of bismit's per-cycle cortical pipeline.

The compute runs in libbismit_hip.so (hand-written HIP for gfx950 behind the C
ABI in include/bismit_hip.h); this package holds the host-side mirror of the
reference's Cortex/CorticalArea API, area recipes, input generators and the
batched/serving loops.
"""
from . import _lib  # noqa: F401
from .recipes import AreaScheme, LayerScheme, SliceScheme, TuftScheme, calc_scale  # noqa: F401

__all__ = ["AreaScheme", "LayerScheme", "SliceScheme", "TuftScheme", "calc_scale"]
