"""NetCDF DEM, roughness and rainfall inputs remapped to a model grid (src/io_netcdf.py).

The reference opens files with xarray (``decode_cf=True, mask_and_scale=True``)
and interpolates with ``DataArray.interp``.  Neither xarray nor netCDF4 is
installed here, so NetCDF-4 files are read by ``diffswe2d_amd.hdf5`` (NetCDF-4
is HDF5 underneath), NetCDF-3 files by scipy's classic-format reader, and both
are decoded the way xarray decodes them:

* variables are the root group's datasets other than pure dimensions; a
  variable's dimension names come from its ``_Netcdf4Coordinates`` dimension
  ids, matched against the ``_Netcdf4Dimid`` of the dimension-scale datasets;
* ``_FillValue`` and ``missing_value`` become NaN; ``scale_factor`` and
  ``add_offset`` are applied;
* ``interp`` over two independent 1-D target coordinates is two 1-D
  interpolations, x first, then y (scipy ``interp1d``; NaN outside the source
  range for ``linear``; ``nearest`` with extrapolation for the outside fill).

On top of that sit the reference's functions with its names and rules:
``load_dem_and_roughness_to_grid`` (DEM and log-space roughness),
``conservative_rectilinear_remap`` and ``load_rainfall_to_grid``.
"""
from __future__ import annotations

from dataclasses import dataclass
from pathlib import Path
from typing import Dict, Optional, Sequence

import numpy as np
import torch
from scipy.interpolate import interp1d

from . import hdf5
from .forcing import RainfallForcing
from .grid import ModelGrid

_NOT_A_VARIABLE = "This is a netCDF dimension but not a netCDF variable."


@dataclass(frozen=True)
class GridData:
    """Remapped DEM and roughness on the model grid (io_netcdf.py:63-72)."""
    bed: torch.Tensor
    roughness_length: torch.Tensor
    x: torch.Tensor
    y: torch.Tensor
    dx: float
    dy: float
    valid_mask: torch.Tensor
    crs: Optional[str]


class Variable:
    """A decoded NetCDF variable: ``values`` (numpy), ``dims`` (names), ``attrs``."""

    def __init__(self, name, values, dims, attrs, coords=None):
        self.name, self.values, self.dims, self.attrs = name, values, tuple(dims), dict(attrs)
        self.coords = {} if coords is None else coords
        if self.values.ndim != len(self.dims):
            raise ValueError(f"variable {name!r}: {self.values.ndim} axes but dimensions {self.dims}")

    def __getitem__(self, dim):
        """Coordinate vector along one of this variable's dimensions (set by the dataset)."""
        return self.coords[dim]

    def squeeze(self):
        keep = [i for i, n in enumerate(self.values.shape) if n != 1]
        v = Variable(self.name, self.values.reshape([self.values.shape[i] for i in keep]),
                     [self.dims[i] for i in keep], self.attrs)
        v.coords = self.coords
        return v

    def transpose(self, *dims):
        if sorted(dims) != sorted(self.dims):
            raise ValueError(f"variable {self.name!r} has dimensions {self.dims}, not {dims}")
        v = Variable(self.name, np.transpose(self.values, [self.dims.index(d) for d in dims]), dims, self.attrs)
        v.coords = self.coords
        return v

    def flipped(self, dim):
        axis = self.dims.index(dim)
        v = Variable(self.name, np.flip(self.values, axis=axis).copy(), self.dims, self.attrs)
        v.coords = dict(self.coords)
        v.coords[dim] = self.coords[dim][::-1].copy()
        return v


class Dataset:
    """Root-group variables of a NetCDF-4 or NetCDF-3 file, CF-decoded (xarray's mask_and_scale)."""

    def __init__(self, path):
        self.path = Path(path)
        if not self.path.exists():
            raise FileNotFoundError(f"NetCDF file not found: {self.path}")
        with open(self.path, "rb") as fh:
            magic = fh.read(4)
        if magic[:3] == b"CDF":
            self._open_classic()
        else:
            self._open_hdf5()
        for v in self.variables.values():
            v.coords = self.coords
        self.data_vars = {n: v for n, v in self.variables.items() if n not in self.coords}

    def _open_classic(self):
        """NetCDF-3 (classic / 64-bit offset) through scipy's reader."""
        from scipy.io import netcdf_file

        def text(v):
            return v.decode("utf-8", "replace") if isinstance(v, bytes) else v

        with netcdf_file(self.path, "r", mmap=False, maskandscale=False) as f:
            self.attrs = {k: text(v) for k, v in f._attributes.items()}
            self.variables = {}
            for name, var in f.variables.items():
                attrs = {k: text(v) for k, v in var._attributes.items()}
                values = np.array(var.data, copy=True)
                self.variables[name] = Variable(name, self._decode(values, attrs), var.dimensions, attrs)
        self.coords = {n: v.values for n, v in self.variables.items() if v.dims == (n,)}

    def _open_hdf5(self):
        """NetCDF-4: HDF5 datasets of the root group."""
        with hdf5.File(self.path) as f:
            self.attrs = dict(f.root_attrs)
            raw = {name: f.dataset(name) for name in f.links}
            dim_names, dimids, coords = {}, {}, {}
            for name, ds in raw.items():
                a = ds.attrs
                if a.get("CLASS") == "DIMENSION_SCALE":
                    if "_Netcdf4Dimid" in a:
                        dimids[int(np.asarray(a["_Netcdf4Dimid"]).reshape(-1)[0])] = name
                    dim_names[name] = len(ds.shape) and ds.shape[0]
            self.variables: Dict[str, Variable] = {}
            for name, ds in raw.items():
                a = ds.attrs
                if str(a.get("NAME", "")).startswith(_NOT_A_VARIABLE):
                    continue
                if a.get("CLASS") == "DIMENSION_SCALE":
                    dims = (name,)
                elif "_Netcdf4Coordinates" in a and all(int(i) in dimids for i in
                                                        np.asarray(a["_Netcdf4Coordinates"]).reshape(-1)):
                    dims = tuple(dimids[int(i)] for i in np.asarray(a["_Netcdf4Coordinates"]).reshape(-1))
                else:
                    dims = self._dims_by_size(name, ds.shape, dim_names)
                self.variables[name] = Variable(name, self._decode(ds.read(), a), dims,
                                                {k: v for k, v in a.items() if not k.startswith("_Netcdf4")
                                                 and k not in ("CLASS", "NAME", "REFERENCE_LIST",
                                                               "DIMENSION_LIST")})
            for name in dim_names:
                if name in self.variables:
                    coords[name] = self.variables[name].values
        self.coords = coords

    @staticmethod
    def _dims_by_size(name, shape, dim_names):
        dims = []
        for n in shape:
            match = [d for d, size in dim_names.items() if size == n]
            if len(match) != 1:
                raise ValueError(f"cannot tell the dimensions of variable {name!r} (shape {shape})")
            dims.append(match[0])
        return tuple(dims)

    @staticmethod
    def _decode(values, attrs):
        masks = [attrs[k] for k in ("_FillValue", "missing_value") if k in attrs and attrs[k] is not None]
        scale, offset = attrs.get("scale_factor"), attrs.get("add_offset")
        if not masks and scale is None and offset is None:
            return values
        out = values.astype(np.float64) if values.dtype.kind in "iu" else values.copy()
        for m in masks:
            for fill in np.asarray(m).reshape(-1):
                if np.isnan(fill):
                    continue
                out[values == fill] = np.nan
        if scale is not None:
            out = out * np.asarray(scale).reshape(-1)[0]
        if offset is not None:
            out = out + np.asarray(offset).reshape(-1)[0]
        return out

    def __getitem__(self, name) -> Variable:
        return self.variables[name]

    def __contains__(self, name) -> bool:
        return name in self.variables

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        return False


def open_dataset(path) -> Dataset:
    return Dataset(path)


def _find_variable(ds: Dataset, requested: Optional[str], candidates: Sequence[str]) -> str:
    """Requested name, else the first candidate present among the data variables (io_netcdf.py:75-83)."""
    if requested is not None:
        if requested not in ds.data_vars:
            raise KeyError(f"Variable {requested!r} not found; available={list(ds.data_vars)}")
        return requested
    lookup = {n.lower(): n for n in ds.data_vars}
    for c in candidates:
        if c.lower() in lookup:
            return lookup[c.lower()]
    raise KeyError(f"Could not infer variable from {candidates}; available={list(ds.data_vars)}")


def _ascending(var: Variable, x_name: str, y_name: str) -> Variable:
    """(y, x) order with increasing coordinates (io_netcdf.py:86-90)."""
    var = var.transpose(y_name, x_name)
    for d in (x_name, y_name):
        c = var.coords[d]
        if float(c[0]) > float(c[-1]):
            var = var.flipped(d)
    return var


def _crs_text(ds: Dataset, var: Variable):
    mapping = var.attrs.get("grid_mapping")
    if mapping and mapping in ds.variables:
        a = ds[mapping].attrs
        return a.get("crs_wkt") or a.get("spatial_ref") or str(a)
    return ds.attrs.get("crs") or var.attrs.get("crs")


def _interp_axis(values, src, dst, axis, method, extrapolate):
    f = interp1d(src, values, kind=method, axis=axis, bounds_error=False, copy=False,
                 fill_value="extrapolate" if extrapolate else np.nan, assume_sorted=True)
    return f(dst)


def _interp(var: Variable, tx, ty, x_name, y_name, method, extrapolate=False):
    out = _interp_axis(var.values, var.coords[x_name], tx, 1, method, extrapolate)
    return _interp_axis(out, var.coords[y_name], ty, 0, method, extrapolate)


def _remap(var: Variable, grid: ModelGrid, x_name, y_name, method, outside, field_name):
    """Interpolate a (y, x) variable onto model-cell centres (io_netcdf.py:117-141)."""
    if method not in ("linear", "nearest"):
        raise ValueError("method must be linear or nearest")
    if outside not in ("error", "nearest", "nan", "zero"):
        raise ValueError("invalid outside policy")
    tx = np.asarray(grid.x.detach().cpu())
    ty = np.asarray(grid.y.detach().cpu())
    sx, sy = np.asarray(var.coords[x_name]), np.asarray(var.coords[y_name])
    if (tx.min() < sx.min() or tx.max() > sx.max() or ty.min() < sy.min() or ty.max() > sy.max()) \
            and outside == "error":
        raise ValueError(f"Model grid extends beyond {field_name} source coverage")
    out = _interp(var, tx, ty, x_name, y_name, method)
    if outside == "nearest":
        near = _interp(var, tx, ty, x_name, y_name, "nearest", extrapolate=True)
        out = np.where(np.isnan(out), near, out)
    elif outside == "zero":
        out = np.where(np.isnan(out), 0.0, out)
    return out


def load_dem_and_roughness_to_grid(path, model_grid: ModelGrid, dem_variable=None, roughness_path=None,
                                   roughness_variable=None, x_name="x", y_name="y", dem_method="linear",
                                   roughness_method="linear", outside_domain="error", dtype=torch.float64,
                                   device=None) -> GridData:
    """DEM (and roughness, interpolated in log space) on the model grid (io_netcdf.py:144-198)."""
    ds = open_dataset(path)
    zn = _find_variable(ds, dem_variable, ("zb", "bed", "bed_elevation", "elevation", "dem", "z"))
    z = _ascending(ds[zn].squeeze(), x_name, y_name)
    if not np.isfinite(z.values).all():
        raise ValueError("Source DEM contains missing values")
    z_values = np.asarray(_remap(z, model_grid, x_name, y_name, dem_method, outside_domain, "DEM"),
                          dtype=np.float64)
    try:
        ds_r = open_dataset(roughness_path) if roughness_path is not None else ds
        rn = _find_variable(ds_r, roughness_variable, ("roughness_length", "roughness", "z0", "zo", "n"))
        r = _ascending(ds_r[rn].squeeze(), x_name, y_name)
        if not np.isfinite(r.values).all() or np.any(r.values <= 0):
            raise ValueError("Source roughness length must be finite and positive")
        log_r = Variable(r.name, np.log(r.values), r.dims, r.attrs)
        log_r.coords = r.coords
        r_model = np.exp(np.asarray(_remap(log_r, model_grid, x_name, y_name, roughness_method, outside_domain,
                                           "roughness"), dtype=np.float64))
    except (KeyError, FileNotFoundError, OSError):
        r_model = np.ones_like(z_values)
    valid = np.isfinite(z_values) & np.isfinite(r_model)
    if not valid.all():
        raise ValueError("Model grid contains cells not covered by DEM/roughness")
    crs = _crs_text(ds, z)
    return GridData(torch.as_tensor(z_values, dtype=dtype, device=device),
                    torch.as_tensor(r_model, dtype=dtype, device=device), model_grid.x, model_grid.y,
                    model_grid.dx, model_grid.dy, torch.as_tensor(valid, dtype=torch.bool, device=device), crs)


def _centres_to_edges(centres, name):
    c = np.asarray(centres, dtype=np.float64)
    if c.ndim != 1 or c.size < 2 or np.any(np.diff(c) <= 0):
        raise ValueError(f"{name} must be strictly increasing 1-D cell centres")
    e = np.empty(c.size + 1, dtype=np.float64)
    e[1:-1] = 0.5 * (c[:-1] + c[1:])
    e[0] = c[0] - 0.5 * (c[1] - c[0])
    e[-1] = c[-1] + 0.5 * (c[-1] - c[-2])
    return e


def _overlap_matrix(source_edges, target_edges):
    """Lengths of the intersections of target and source cells, [target, source]."""
    lo = np.maximum(target_edges[:-1, None], source_edges[None, :-1])
    hi = np.minimum(target_edges[1:, None], source_edges[None, 1:])
    return np.maximum(0.0, hi - lo)


def conservative_rectilinear_remap(values, source_x, source_y, target_grid, outside_domain="zero"):
    """Area-weighted remap of cell-average rates [time, y, x] to model cells (io_netcdf.py:219-246).

    Each target cell takes sum(R_S * |S n T|) / |T|; area outside the source
    contributes zero, or raises with ``outside_domain='error'``."""
    if outside_domain not in ("zero", "error"):
        raise ValueError("Conservative rainfall supports outside_domain='zero' or 'error'")
    sxe, sye = _centres_to_edges(source_x, "rainfall x"), _centres_to_edges(source_y, "rainfall y")
    txe = _centres_to_edges(np.asarray(target_grid.x.detach().cpu()), "model x")
    tye = _centres_to_edges(np.asarray(target_grid.y.detach().cpu()), "model y")
    if outside_domain == "error" and (txe[0] < sxe[0] or txe[-1] > sxe[-1] or tye[0] < sye[0]
                                      or tye[-1] > sye[-1]):
        raise ValueError("Model grid extends outside rainfall cell-edge coverage")
    wx, wy = _overlap_matrix(sxe, txe), _overlap_matrix(sye, tye)
    area = np.diff(tye)[:, None] * np.diff(txe)[None, :]
    return np.einsum("ai,tij,bj->tab", wy, values, wx, optimize=True) / area[None, :, :]


_RAIN_UNITS = {"m/s": 1.0, "ms-1": 1.0, "m.s-1": 1.0, "mm/s": 1e-3, "mms-1": 1e-3, "mm/h": 1e-3 / 3600,
               "mm/hr": 1e-3 / 3600, "mmh-1": 1e-3 / 3600, "mmhour-1": 1e-3 / 3600, "m/h": 1 / 3600,
               "m/hr": 1 / 3600, "mh-1": 1 / 3600, "kgm-2s-1": 1e-3}


def _unit_factor(units: str) -> float:
    """Rainfall unit -> m/s factor (io_netcdf.py:100-107)."""
    key = units.lower().replace(" ", "").replace("**", "^")
    if key not in _RAIN_UNITS:
        raise ValueError(f"Unsupported rainfall units {units!r}")
    return _RAIN_UNITS[key]


def _time_seconds(var: Variable):
    """Seconds since the first record (io_netcdf.py:110-117); CF "<unit> since <date>" units included."""
    v = np.asarray(var.values, dtype=np.float64)
    units = str(var.attrs.get("units", "seconds")).lower()
    factor = 1.0
    if units.startswith("minute"):
        factor = 60.0
    elif units.startswith("hour"):
        factor = 3600.0
    elif units.startswith("day"):
        factor = 86400.0
    elif units.startswith("millisecond"):
        factor = 1e-3
    return (v - v[0]) * factor


def load_rainfall_to_grid(path, model_grid: ModelGrid, variable=None, time_name="time", x_name="x", y_name="y",
                          units=None, outside_domain="zero", expected_crs=None, dtype=torch.float64,
                          device=None) -> RainfallForcing:
    """Gridded rainfall, conservatively remapped, as a RainfallForcing (io_netcdf.py:249-275)."""
    ds = open_dataset(path)
    name = _find_variable(ds, variable, ("rainfall", "rain", "precipitation_rate", "precipitation"))
    rain = ds[name].transpose(time_name, y_name, x_name)
    for d in (x_name, y_name):
        if float(rain.coords[d][0]) > float(rain.coords[d][-1]):
            rain = rain.flipped(d)
    rain_crs = _crs_text(ds, rain)
    if expected_crs and rain_crs and str(expected_crs) != str(rain_crs):
        raise ValueError("Rainfall and DEM CRS metadata differ; reproject before loading")
    src = np.asarray(rain.values, dtype=np.float64)
    if not np.isfinite(src).all():
        raise ValueError("Source rainfall contains missing values")
    values = conservative_rectilinear_remap(src, np.asarray(rain.coords[x_name]), np.asarray(rain.coords[y_name]),
                                            model_grid, outside_domain=outside_domain)
    factor = _unit_factor(units or rain.attrs.get("units", ""))
    times = _time_seconds(ds[time_name])
    return RainfallForcing(torch.as_tensor(times, dtype=dtype, device=device),
                           torch.as_tensor(values * factor, dtype=dtype, device=device)[:, None])


__all__ = ["GridData", "ModelGrid", "conservative_rectilinear_remap", "load_dem_and_roughness_to_grid",
           "load_rainfall_to_grid", "open_dataset"]
