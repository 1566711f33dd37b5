"""``load_dynamic_model`` with the reference signature and defaults
(src/model_loader.py:6-20).

DEM file loading (NetCDF / ESRI-ASCII, model_loader.py:29-121) is outside the
HIP time-step build: pass the bed and roughness as tensors through
``dem_filepath`` / ``model_grid`` as ``(bed, roughness, dx, dy)``, or build
``SWE2D`` directly.  The text rainfall (:116-118) and tide series (:124-130)
branches are kept.
"""
from __future__ import annotations

import numpy as np
import torch

from .config import SWEConfig
from .forcing import load_rainfall_txt
from .solver import SWE2D


def load_dynamic_model(dem_filepath, model_grid, device, rain_filepath=None, tide_filepath=None,
                       boundary_left="transmissive", boundary_right="wall", boundary_top="wall",
                       boundary_bottom="water_level", frictionmodel="manning", roughness_filepath=None,
                       roughness_var_name="z0", default_roughness_value=0.025):
    """Build an SWE2D with the reference's boundary defaults; returns
    (model, use_txt_rainfall, rain_times, rain_rate_ms)."""
    if isinstance(dem_filepath, str):
        if dem_filepath.endswith((".nc", ".asc")):
            raise NotImplementedError(
                "DEM file loading is not part of the HIP time-step build; pass (bed, roughness, dx, dy) "
                "tensors instead of a file path.")
        raise ValueError("Unsupported DEM format! Please provide a .nc or .asc file.")
    rain_times, rain_rate_ms, use_txt_rainfall = None, None, False
    if rain_filepath:
        rain_times, rain_rate_ms = load_rainfall_txt(rain_filepath)
        use_txt_rainfall = True
    bed, rough, dx, dy = dem_filepath
    bed = torch.as_tensor(bed, dtype=torch.float64, device=device)
    rough = (torch.full_like(bed, float(default_roughness_value)) if rough is None
             else torch.as_tensor(rough, dtype=torch.float64, device=device))
    cfg = SWEConfig(boundary_left=boundary_left, boundary_right=boundary_right, boundary_top=boundary_top,
                    boundary_bottom=boundary_bottom, frictionmodel=frictionmodel)
    model = SWE2D(bed, rough, dx, dy, config=cfg, train_dem=False, maximum_dem_correction=1.0).to(device)
    if tide_filepath:
        tide = np.loadtxt(tide_filepath)
        model.bc_times = torch.tensor(tide[:, 0], dtype=torch.float64, device=device)
        model.bc_wls = torch.tensor(tide[:, 1], dtype=torch.float64, device=device)
        model.use_dynamic_bc = True
    return model, use_txt_rainfall, rain_times, rain_rate_ms
