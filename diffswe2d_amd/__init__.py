"""MI355X-native (gfx950 HIP) drop-in for DiffSWE2d's per-time-step update.

Public names mirror the reference package (src/__init__.py:9-20) for the
time-step path: ``SWEConfig``, ``SWE2D``, ``RainfallForcing`` and
``load_dynamic_model``.
"""
from .config import SWEConfig
from .forcing import RainfallForcing
from .solver import SWE2D
from .model_loader import load_dynamic_model

__all__ = ["SWEConfig", "SWE2D", "RainfallForcing", "load_dynamic_model"]
