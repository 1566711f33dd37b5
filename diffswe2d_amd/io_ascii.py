"""ESRI-ASCII rasters regridded onto a model grid (src/io_ascii.py).

Conventions, as in the reference module:

* a model grid is given by its OUTER edges and one square resolution; cell
  centres sit at ``xmin + (i + 0.5) * resolution``;
* a raster's origin is ``xllcorner/yllcorner`` (outer corner) or
  ``xllcenter/yllcenter`` (centre of the lower-left cell, half a cell in);
* the file stores the northern row first; arrays here run south to north,
  so row index grows with y;
* NODATA cells become NaN, are optionally filled from the nearest valid
  source cell, and the raster is then interpolated (linear or nearest) to the
  model-cell centres; target centres outside the raster raise, take the
  nearest source value, or stay NaN.

Parsing and interpolation use the same library routines as the reference
(pandas' C tokenizer, scipy's RegularGridInterpolator and
NearestNDInterpolator), so a bed read here is bit-identical to the
reference's for the same file and grid.
"""
from __future__ import annotations

from dataclasses import dataclass
from pathlib import Path
from typing import Any, Mapping, Optional, Union

import numpy as np
import pandas as pd
import torch
from scipy.interpolate import NearestNDInterpolator, RegularGridInterpolator

from .forcing import load_rainfall_txt

PathLike = Union[str, Path]

_HEADER_KEYS = ("ncols", "nrows", "xllcorner", "yllcorner", "xllcenter", "yllcenter", "cellsize",
                "nodata_value")


def _finite_float(value: Any, name: str) -> float:
    """One finite float from a Python / NumPy / torch scalar (io_ascii.py:280-311)."""
    if isinstance(value, (torch.Tensor, np.ndarray)):
        n = value.numel() if isinstance(value, torch.Tensor) else value.size
        if n != 1:
            raise ValueError(f"{name} must contain one value, but its shape is {tuple(value.shape)}.")
        value = value.detach().cpu().item() if isinstance(value, torch.Tensor) else value.item()
    out = float(value)
    if not np.isfinite(out):
        raise ValueError(f"{name} must be finite, got {out}.")
    return out


def _cell_count(lo: float, hi: float, res: float, axis: str) -> int:
    exact = (hi - lo) / res
    n = int(round(exact))
    if not np.isclose(exact, n, rtol=0.0, atol=max(1.0e-9, abs(exact) * 1.0e-10)):
        raise ValueError(f"The {axis}-extent is not divisible by resolution: lower={lo}, upper={hi}, "
                         f"resolution={res}, cell count={exact}")
    if n < 1:
        raise ValueError(f"The grid must contain at least one {axis}-cell.")
    return n


def _centres(lo: float, n: int, res: float) -> np.ndarray:
    return lo + (np.arange(n, dtype=np.float64) + 0.5) * res


def _edges(lo: float, n: int, res: float) -> np.ndarray:
    return lo + np.arange(n + 1, dtype=np.float64) * res


@dataclass(frozen=True)
class GridSpec:
    """CPU description of a square-cell model grid by its outer edges (io_ascii.py:89-277)."""
    xmin: float
    xmax: float
    ymin: float
    ymax: float
    resolution: float

    def __post_init__(self) -> None:
        for name in ("xmin", "xmax", "ymin", "ymax", "resolution"):
            if not np.isfinite(getattr(self, name)):
                raise ValueError(f"{name} must be finite, got {getattr(self, name)}.")
        if self.xmax <= self.xmin:
            raise ValueError(f"xmax must be greater than xmin: xmin={self.xmin}, xmax={self.xmax}.")
        if self.ymax <= self.ymin:
            raise ValueError(f"ymax must be greater than ymin: ymin={self.ymin}, ymax={self.ymax}.")
        if self.resolution <= 0.0:
            raise ValueError(f"resolution must be positive, got {self.resolution}.")
        self.nx, self.ny  # noqa: B018 - validates both extents

    @property
    def nx(self) -> int:
        return _cell_count(self.xmin, self.xmax, self.resolution, "x")

    @property
    def ny(self) -> int:
        return _cell_count(self.ymin, self.ymax, self.resolution, "y")

    @property
    def shape(self) -> tuple:
        return self.ny, self.nx

    @property
    def x(self) -> np.ndarray:
        return _centres(self.xmin, self.nx, self.resolution)

    @property
    def y(self) -> np.ndarray:
        return _centres(self.ymin, self.ny, self.resolution)

    @property
    def x_edges(self) -> np.ndarray:
        return _edges(self.xmin, self.nx, self.resolution)

    @property
    def y_edges(self) -> np.ndarray:
        return _edges(self.ymin, self.ny, self.resolution)

    def as_dict(self) -> dict:
        return dict(xmin=self.xmin, xmax=self.xmax, ymin=self.ymin, ymax=self.ymax, resolution=self.resolution,
                    nx=self.nx, ny=self.ny, shape=self.shape)


def _grid_resolution(grid: Any) -> Optional[float]:
    """Square cell size from ``resolution``/``dx``/``cellsize``/``cell_size`` (io_ascii.py:314-376)."""
    for name in ("resolution", "dx", "cellsize", "cell_size"):
        if not hasattr(grid, name):
            continue
        value = getattr(grid, name)
        vals = (value.detach().cpu().numpy() if isinstance(value, torch.Tensor) else np.asarray(value)).reshape(-1)
        if vals.size == 1:
            return _finite_float(vals[0], name)
        if vals.size == 2:
            dx, dy = _finite_float(vals[0], f"{name}[0]"), _finite_float(vals[1], f"{name}[1]")
            if not np.isclose(dx, dy, rtol=0.0, atol=max(1.0e-12, abs(dx) * 1.0e-10)):
                raise ValueError(f"This ASCII reader currently requires square model cells, but dx={dx} "
                                 f"and dy={dy}.")
            return dx
        raise ValueError(f"{name} must be scalar or contain dx and dy, but it contains {vals.size} values.")
    return None


def _coordinate_vector(value: Any, name: str) -> np.ndarray:
    arr = value.detach().cpu().numpy() if isinstance(value, torch.Tensor) else np.asarray(value)
    arr = np.asarray(arr, dtype=np.float64).squeeze()
    if arr.ndim != 1:
        raise ValueError(f"{name} must be a one-dimensional coordinate vector, but its shape is {arr.shape}.")
    if arr.size < 1:
        raise ValueError(f"{name} cannot be empty.")
    if not np.isfinite(arr).all():
        raise ValueError(f"{name} contains NaN or infinite values.")
    return arr


def parse_model_grid(model_grid: Any) -> GridSpec:
    """GridSpec from a GridSpec, a bounds mapping, an object with bounds + resolution,
    or an object with centre vectors x, y + resolution (io_ascii.py:419-606)."""
    if isinstance(model_grid, GridSpec):
        return model_grid
    keys = ("xmin", "xmax", "ymin", "ymax", "resolution")
    if isinstance(model_grid, Mapping):
        missing = [k for k in keys if k not in model_grid]
        if missing:
            raise KeyError(f"model_grid is missing required entries: {missing}")
        return GridSpec(*(_finite_float(model_grid[k], k) for k in keys))
    res = _grid_resolution(model_grid)
    if all(hasattr(model_grid, k) for k in keys[:4]):
        if res is None:
            raise AttributeError("The supplied model_grid contains bounds but does not provide resolution, dx, "
                                 "cellsize, or cell_size.")
        return GridSpec(*(_finite_float(getattr(model_grid, k), k) for k in keys[:4]), res)
    if hasattr(model_grid, "x") and hasattr(model_grid, "y"):
        if res is None:
            raise AttributeError("The supplied model_grid contains x and y coordinates but does not provide "
                                 "resolution, dx, cellsize, or cell_size.")
        x = _coordinate_vector(model_grid.x, "model_grid.x")
        y = _coordinate_vector(model_grid.y, "model_grid.y")
        grid = GridSpec(float(np.min(x) - 0.5 * res), float(np.max(x) + 0.5 * res),
                        float(np.min(y) - 0.5 * res), float(np.max(y) + 0.5 * res), res)
        for axis, n, c in (("x", grid.nx, x), ("y", grid.ny, y)):
            if n != c.size:
                raise ValueError(f"The number of {axis} coordinates does not agree with the inferred grid bounds "
                                 f"and resolution: coordinate count = {c.size}, inferred n{axis} = {n}")
        return grid
    raise TypeError("Unsupported model_grid object. The supplied grid must be a GridSpec, a dictionary with xmin, "
                    "xmax, ymin, ymax, resolution, an object with those bounds and a resolution, or an object "
                    "with x, y and a resolution")


@dataclass
class AsciiRaster:
    """Raster values (south row first) with their outer bounds (io_ascii.py:614-719)."""
    values: np.ndarray
    xmin: float
    xmax: float
    ymin: float
    ymax: float
    resolution: float
    ncols: int
    nrows: int
    nodata_value: Optional[float] = None
    path: Optional[Path] = None

    @property
    def shape(self) -> tuple:
        return self.nrows, self.ncols

    @property
    def x(self) -> np.ndarray:
        return _centres(self.xmin, self.ncols, self.resolution)

    @property
    def y(self) -> np.ndarray:
        return _centres(self.ymin, self.nrows, self.resolution)

    @property
    def x_edges(self) -> np.ndarray:
        return _edges(self.xmin, self.ncols, self.resolution)

    @property
    def y_edges(self) -> np.ndarray:
        return _edges(self.ymin, self.nrows, self.resolution)

    def as_dict(self) -> dict:
        return dict(path=None if self.path is None else str(self.path), xmin=self.xmin, xmax=self.xmax,
                    ymin=self.ymin, ymax=self.ymax, resolution=self.resolution, ncols=self.ncols,
                    nrows=self.nrows, shape=self.shape, nodata_value=self.nodata_value)


def _read_header(path: Path):
    """Header entries (lower-case keys) and the number of raw lines they occupy (io_ascii.py:722-840)."""
    if not path.exists():
        raise FileNotFoundError(f"ASCII raster was not found: {path}")
    if not path.is_file():
        raise ValueError(f"ASCII raster path is not a file: {path}")
    header, consumed = {}, 0
    with path.open("r", encoding="utf-8") as f:
        for raw in f:
            words = raw.split()
            if not words:
                consumed += 1
                continue
            if len(words) < 2 or words[0].lower() not in _HEADER_KEYS:
                break
            try:
                header[words[0].lower()] = float(words[1])
            except ValueError as exc:
                raise ValueError(f"Invalid ASCII header line in {path}:\n    {raw.strip()}") from exc
            consumed += 1
    missing = {"ncols", "nrows", "cellsize"}.difference(header)
    if missing:
        raise ValueError(f"ASCII header in {path} is missing entries: {sorted(missing)}")
    if not ({"xllcorner", "yllcorner"} <= header.keys() or {"xllcenter", "yllcenter"} <= header.keys()):
        raise ValueError(f"ASCII header in {path} must contain either xllcorner and yllcorner or xllcenter "
                         "and yllcenter")
    return header, consumed


def read_ascii(filepath: PathLike, *, convert_nodata_to_nan: bool = True, flip_y: bool = True,
               dtype=np.float64) -> AsciiRaster:
    """Read an ESRI-ASCII raster (io_ascii.py:843-1004)."""
    path = Path(filepath)
    header, skip = _read_header(path)
    counts = []
    for key in ("ncols", "nrows"):
        n = int(round(header[key]))
        if not np.isclose(header[key], n):
            raise ValueError(f"{key} must be an integer, got {header[key]}.")
        counts.append(n)
    ncols, nrows = counts
    if ncols < 1 or nrows < 1:
        raise ValueError(f"ncols and nrows must be positive, got ncols={ncols}, nrows={nrows}.")
    res = float(header["cellsize"])
    if not np.isfinite(res) or res <= 0.0:
        raise ValueError(f"ASCII cellsize must be finite and positive, got {res}.")
    values = pd.read_csv(path, sep=r"\s+", skiprows=skip, header=None).values
    if values.ndim == 1:
        values = values.reshape(nrows, ncols) if values.size == nrows * ncols else values
    if values.shape != (nrows, ncols):
        raise ValueError(f"ASCII raster dimensions do not match the header: file={path}, loaded shape="
                         f"{values.shape}, expected shape={(nrows, ncols)}")
    xmin = header["xllcorner"] if "xllcorner" in header else header["xllcenter"] - 0.5 * res
    ymin = header["yllcorner"] if "yllcorner" in header else header["yllcenter"] - 0.5 * res
    values = np.array(values, dtype=dtype)
    nodata = header.get("nodata_value")
    if convert_nodata_to_nan and nodata is not None:
        values[np.isclose(values, nodata, rtol=0.0, atol=max(1.0e-12, abs(nodata) * 1.0e-12))] = np.nan
    values = np.flipud(values).copy() if flip_y else np.ascontiguousarray(values)
    return AsciiRaster(values=values, xmin=float(xmin), xmax=float(xmin) + ncols * res, ymin=float(ymin),
                       ymax=float(ymin) + nrows * res, resolution=res, ncols=ncols, nrows=nrows,
                       nodata_value=nodata, path=path)


def fill_nodata_nearest(raster: AsciiRaster) -> np.ndarray:
    """Source values with NaN holes taken from the nearest valid source cell (io_ascii.py:1009-1063)."""
    values = np.asarray(raster.values, dtype=np.float64).copy()
    holes = ~np.isfinite(values)
    if not holes.any():
        return values
    if holes.all():
        raise ValueError(f"The ASCII raster contains no valid finite cells: {raster.path}")
    xx, yy = np.meshgrid(raster.x, raster.y, indexing="xy")
    good = ~holes
    lookup = NearestNDInterpolator(np.column_stack((yy[good], xx[good])), values[good])
    values[holes] = lookup(yy[holes], xx[holes])
    return values


def regrid_ascii(raster: AsciiRaster, model_grid: Any, *, method: str = "linear", outside_domain: str = "error",
                 fill_internal_nodata: bool = True) -> np.ndarray:
    """Interpolate a raster to the model-cell centres; returns [ny, nx] (io_ascii.py:1069-1263)."""
    grid = parse_model_grid(model_grid)
    method, outside_domain = str(method).lower(), str(outside_domain).lower()
    if method not in ("linear", "nearest"):
        raise ValueError(f"method must be 'linear' or 'nearest', got {method!r}.")
    if outside_domain not in ("error", "nearest", "nan"):
        raise ValueError(f"outside_domain must be 'error', 'nearest', or 'nan', got {outside_domain!r}.")
    if method == "linear" and (raster.ncols < 2 or raster.nrows < 2):
        raise ValueError("Linear interpolation requires at least two source cells in both x and y. Use "
                         "method='nearest' for a one-row or one-column raster.")
    src = fill_nodata_nearest(raster) if fill_internal_nodata else np.asarray(raster.values, np.float64).copy()
    if method == "linear" and not np.isfinite(src).all():
        raise ValueError("The source raster contains NaN or infinite values. Use fill_internal_nodata=True "
                         "before linear interpolation.")
    tx, ty = np.meshgrid(grid.x, grid.y, indexing="xy")
    pts = np.column_stack((ty.ravel(), tx.ravel()))
    axes = (raster.y, raster.x)
    out = RegularGridInterpolator(axes, src, method=method, bounds_error=False, fill_value=np.nan)(pts)
    out = out.reshape(grid.shape)
    bad = ~np.isfinite(out)
    if bad.any() and outside_domain == "nearest":
        near = RegularGridInterpolator(axes, src, method="nearest", bounds_error=False, fill_value=None)(pts)
        out[bad] = near.reshape(grid.shape)[bad]
        bad = ~np.isfinite(out)
    if bad.any() and outside_domain == "error":
        raise ValueError(
            f"{int(bad.sum())} target cells could not be assigned a valid raster value. Source ASCII outer "
            f"bounds x [{raster.xmin}, {raster.xmax}], y [{raster.ymin}, {raster.ymax}]; target model-grid "
            f"outer bounds x [{grid.xmin}, {grid.xmax}], y [{grid.ymin}, {grid.ymax}]; invalid target "
            f"cell-centre range x [{tx[bad].min()}, {tx[bad].max()}], y [{ty[bad].min()}, {ty[bad].max()}]. "
            "Use outside_domain='nearest' if nearest-neighbour extrapolation is acceptable.")
    return np.ascontiguousarray(out)


def read_ascii_on_grid(filepath: PathLike, model_grid: Any, *, method: str = "linear",
                       outside_domain: str = "error", fill_internal_nodata: bool = True, dtype=np.float32):
    """Read and regrid; returns (values [ny, nx], GridSpec) (io_ascii.py:1266-1308)."""
    grid = parse_model_grid(model_grid)
    raster = read_ascii(filepath, convert_nodata_to_nan=True, flip_y=True, dtype=np.float64)
    values = regrid_ascii(raster, grid, method=method, outside_domain=outside_domain,
                          fill_internal_nodata=fill_internal_nodata)
    return np.ascontiguousarray(np.asarray(values, dtype=dtype)), grid


def ascii_to_tensor(filepath: PathLike, model_grid: Any, *, method: str = "linear", outside_domain: str = "error",
                    fill_internal_nodata: bool = True, device=None, dtype: torch.dtype = torch.float32):
    """Read, regrid (in float64) and convert; returns (tensor [ny, nx], GridSpec) (io_ascii.py:1311-1386)."""
    values, grid = read_ascii_on_grid(filepath, model_grid, method=method, outside_domain=outside_domain,
                                      fill_internal_nodata=fill_internal_nodata, dtype=np.float64)
    tensor = torch.as_tensor(values, dtype=dtype, device=device)
    if tuple(tensor.shape) != grid.shape:
        raise RuntimeError(f"Unexpected output tensor shape: actual={tuple(tensor.shape)}, expected={grid.shape}")
    return tensor, grid


def read_ascii_native_tensor(filepath: PathLike, *, convert_nodata_to_nan: bool = True, device=None,
                             dtype: torch.dtype = torch.float32):
    """Raster on its own grid, south row first; returns (tensor, GridSpec) (io_ascii.py:1389-1436)."""
    r = read_ascii(filepath, convert_nodata_to_nan=convert_nodata_to_nan, flip_y=True, dtype=np.float64)
    grid = GridSpec(r.xmin, r.xmax, r.ymin, r.ymax, r.resolution)
    return torch.as_tensor(np.ascontiguousarray(r.values), dtype=dtype, device=device), grid


def grids_are_equal(raster: AsciiRaster, model_grid: Any, *, atol: float = 1.0e-8) -> bool:
    """True when the raster already lies on the model grid cell for cell (io_ascii.py:1439-1484)."""
    g = parse_model_grid(model_grid)
    same = all(np.isclose(a, b, rtol=0.0, atol=atol) for a, b in
               ((raster.xmin, g.xmin), (raster.xmax, g.xmax), (raster.ymin, g.ymin), (raster.ymax, g.ymax),
                (raster.resolution, g.resolution)))
    return bool(same and raster.ncols == g.nx and raster.nrows == g.ny)


def describe_ascii(filepath: PathLike) -> dict:
    """Raster metadata plus valid/invalid counts and min/max/mean of the valid cells (io_ascii.py:1487-1530)."""
    r = read_ascii(filepath, convert_nodata_to_nan=True, flip_y=True, dtype=np.float64)
    ok = np.isfinite(r.values)
    d = r.as_dict()
    any_ok = bool(ok.any())
    d.update(valid_cell_count=int(ok.sum()), invalid_cell_count=int((~ok).sum()),
             minimum=float(np.nanmin(r.values)) if any_ok else None,
             maximum=float(np.nanmax(r.values)) if any_ok else None,
             mean=float(np.nanmean(r.values)) if any_ok else None)
    return d


def describe_model_grid(model_grid: Any) -> dict:
    """GridSpec.as_dict of any accepted model-grid form (io_ascii.py:1533-1540)."""
    return parse_model_grid(model_grid).as_dict()


__all__ = ["AsciiRaster", "GridSpec", "ascii_to_tensor", "describe_ascii", "describe_model_grid",
           "fill_nodata_nearest", "grids_are_equal",
           "load_rainfall_txt", "parse_model_grid", "read_ascii", "read_ascii_native_tensor",
           "read_ascii_on_grid", "regrid_ascii"]
