"""``load_dynamic_model`` with the reference signature and defaults
(src/model_loader.py:6-133).

* ``.asc`` DEMs: ESRI-ASCII through ``SWE2D.from_ascii`` with the reference's
  choices (linear DEM / nearest roughness regridding, nearest extrapolation,
  NODATA filled, FP64, ``maximum_dem_correction=1.0``; model_loader.py:68-111);
  the text rainfall series (:116-118) is read and returned for the caller to
  apply.
* ``.nc`` DEMs: NetCDF-4 through ``SWE2D.from_netcdf`` (linear DEM and
  roughness remapping, nearest extrapolation, FP64; model_loader.py:29-64);
  a ``.nc`` rainfall file becomes gridded forcing (mm/h), a ``.txt`` one is
  returned for the caller to apply.
* A tuple ``(bed, roughness, dx, dy)`` of arrays in place of a path builds the
  same model from tensors (no file).
* The optional tide series (:124-130) sets the dynamic boundary level.
"""
from __future__ import annotations

from pathlib import Path

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
    cfg = SWEConfig(boundary_left=boundary_left, boundary_right=boundary_right, boundary_top=boundary_top,
                    boundary_bottom=boundary_bottom, frictionmodel=frictionmodel)
    rain_times, rain_rate_ms, use_txt_rainfall = None, None, False
    if isinstance(dem_filepath, (str, Path)):
        path = str(dem_filepath)
        if path.endswith(".nc"):
            rough_kw = ({"roughness_path": roughness_filepath, "roughness_variable": roughness_var_name}
                        if roughness_filepath else {"default_roughness": default_roughness_value})
            rain_kw = {}
            if rain_filepath and str(rain_filepath).endswith(".nc"):
                rain_kw = {"rainfall_path": rain_filepath, "rainfall_variable": "rainfall",
                           "rainfall_units": "mm/h"}
            elif rain_filepath and str(rain_filepath).endswith(".txt"):
                rain_times, rain_rate_ms = load_rainfall_txt(rain_filepath)
                use_txt_rainfall = True
            model = SWE2D.from_netcdf(path, model_grid=model_grid, config=cfg, dem_method="linear",
                                      roughness_method="linear", dem_outside_domain="nearest", train_dem=False,
                                      maximum_dem_correction=1.0, **rough_kw, **rain_kw).to(device)
        elif not path.endswith(".asc"):
            raise ValueError("Unsupported DEM format! Please provide a .nc or .asc file.")
        else:
            rough_kw = ({"roughness_path": roughness_filepath} if roughness_filepath
                        else {"default_roughness": default_roughness_value})
            model = SWE2D.from_ascii(path, model_grid=model_grid, config=cfg, dem_method="linear",
                                     roughness_method="nearest", dem_outside_domain="nearest",
                                     roughness_outside_domain="nearest", fill_dem_nodata=True,
                                     fill_roughness_nodata=True, train_dem=False, maximum_dem_correction=1.0,
                                     device=device, dtype=torch.float64, **rough_kw).to(device)
            if rain_filepath:
                rain_times, rain_rate_ms = load_rainfall_txt(rain_filepath)
                use_txt_rainfall = True
    else:
        bed, rough, dx, dy = dem_filepath
        bed = torch.as_tensor(bed, dtype=torch.float64, device=device)
        rough = (torch.full_like(bed, float(default_roughness_value)) if rough is None
                 else torch.as_tensor(rough, dtype=torch.float64, device=device))
        model = SWE2D(bed, rough, dx, dy, config=cfg, train_dem=False, maximum_dem_correction=1.0).to(device)
        if rain_filepath:
            rain_times, rain_rate_ms = load_rainfall_txt(rain_filepath)
            use_txt_rainfall = True
    if tide_filepath:
        tide = np.loadtxt(tide_filepath)
        model.bc_times = torch.tensor(tide[:, 0], dtype=torch.float64, device=device)
        model.bc_wls = torch.tensor(tide[:, 1], dtype=torch.float64, device=device)
        model.use_dynamic_bc = True
    return model, use_txt_rainfall, rain_times, rain_rate_ms
