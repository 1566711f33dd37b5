"""MI355X-native (gfx950 HIP) drop-in for DiffSWE2d's per-time-step update.

Public names mirror the reference package (src/__init__.py:1-20) for the
time-step path and the file inputs around it: ``SWEConfig``, ``SWE2D``,
``RainfallForcing``, ``ModelGrid``, ``GridData``,
``load_dem_and_roughness_to_grid``, ``load_rainfall_to_grid``,
``ascii_to_tensor``, ``parse_model_grid``, ``load_dynamic_model`` and
``setup_logger``.
"""
from .config import SWEConfig
from .forcing import RainfallForcing
from .grid import ModelGrid
from .io_ascii import ascii_to_tensor, parse_model_grid
from .io_netcdf import GridData, load_dem_and_roughness_to_grid, load_rainfall_to_grid
from .logger import setup_logger
from .solver import SWE2D
from .model_loader import load_dynamic_model

__all__ = ["SWEConfig", "SWE2D", "RainfallForcing", "ModelGrid", "GridData", "load_dem_and_roughness_to_grid",
           "load_rainfall_to_grid", "ascii_to_tensor", "parse_model_grid", "load_dynamic_model", "setup_logger"]
