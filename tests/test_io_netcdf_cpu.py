"""NetCDF input path (src/io_netcdf.py, src/solver.py:149-238, model_loader.py:29-64) on the CPU.

The HDF5 reader is pinned by the reference's own files: the Monai bathymetry
(chunked, shuffle + deflate) must reproduce the range its ``actual_range``
attribute records and the bed that the reference's printed gauge records
imply at nine gauge cells of two grids; the 0.08 m run's hmax.nc (contiguous)
must equal, at each gauge cell, the largest depth printed in that gauge's
per-step record of the same run.  Decoding and remapping rules are pinned by
closed forms (a plane survives linear remapping; conservative remapping keeps
the total; CF masks and scales), on files of both NetCDF-4 and NetCDF-3 (written
here with scipy's classic-format writer), and the HDF5 filter pieces the
shipped files do not use (Fletcher-32, filter masks, version-1 pipeline
messages) on hand-built bytes.
"""
from pathlib import Path

import numpy as np
import pytest
import torch

import diffswe2d_amd as swe
from diffswe2d_amd import hdf5, io_netcdf
from diffswe2d_amd.grid import ModelGrid

GOLDEN = Path(__file__).resolve().parent / "golden"
BATHY = GOLDEN / "monai_bathy.nc"


def test_monai_bathymetry_reads_as_recorded():
    with hdf5.File(BATHY) as f:
        assert set(f.links) == {"x", "y", "z"}
        assert f.root_attrs["Conventions"] == "CF-1.7"
        z = f.dataset("z")
        assert z.shape == (246, 394) and z.dtype == np.float32
        zr = z.read()
        lo, hi = z.attrs["actual_range"]
    assert np.isfinite(zr).all() and zr.min() == np.float32(lo) and zr.max() == np.float32(hi)
    ds = io_netcdf.open_dataset(BATHY)
    assert list(ds.data_vars) == ["z"] and ds["z"].dims == ("y", "x")
    x, y = ds.coords["x"], ds.coords["y"]
    assert x.shape == (394,) and y.shape == (246,)
    assert np.allclose(np.diff(x), 0.014, atol=1e-12) and np.allclose(np.diff(y), 0.014, atol=1e-12)
    assert x[0] == 0.0 and abs(x[-1] - 5.502) < 1e-12 and y[0] == 0.0 and abs(y[-1] - 3.43) < 1e-12


@pytest.mark.parametrize("tag,xmax,ymax,res", [("r1", 5.5, 3.43, 0.01), ("r8", 5.44, 3.36, 0.08)])
def test_monai_model_bed_matches_the_reference_gauges(tag, xmax, ymax, res):
    d = np.load(GOLDEN / "monai.npz")
    grid = ModelGrid.from_bounds(xmin=0.0, xmax=xmax, ymin=0.0, ymax=ymax, resolution=res)
    model, use_rain, _, _ = swe.load_dynamic_model(
        dem_filepath=str(BATHY), model_grid=grid, device="cpu", boundary_left="water_level",
        boundary_right="wall", boundary_top="wall", boundary_bottom="wall", frictionmodel="manning",
        default_roughness_value=0.01, tide_filepath=str(GOLDEN / "monai_input_wave.txt"))
    assert not use_rain and model.use_dynamic_bc and model.bc_times.numel() == 452
    assert tuple(model.bed.shape[-2:]) == (grid.ny, grid.nx) and model.bed.dtype == torch.float64
    assert torch.all(model.roughness_length == 0.01) and bool(model.valid_mask.all())
    assert torch.equal(model.x, grid.x) and model.maximum_dem_correction == 1.0
    # each record's first row is after one step from rest: depth moved by at most a few 1e-6
    z_ref = (d[f"{tag}_level_u"][0] - d[f"{tag}_depth_u"][0]) * 1e-6
    for g, (gx, gy) in enumerate(d[f"{tag}_gauge_xy"]):
        ix = int(torch.argmin(torch.abs(grid.x - gx)))
        iy = int(torch.argmin(torch.abs(grid.y - gy)))
        assert abs(float(model.bed[0, 0, iy, ix]) - z_ref[g]) <= 1.0e-6, (tag, g)


def test_contiguous_hmax_agrees_with_the_gauge_records():
    d = np.load(GOLDEN / "monai.npz")
    with hdf5.File(GOLDEN / "monai_hmax_res008.nc") as f:
        hmax = f.dataset("hmax").read()
        x, y = f.dataset("x").read(), f.dataset("y").read()
    assert hmax.shape == (42, 68) and hmax.dtype == np.float64
    assert np.allclose(x, (np.arange(68) + 0.5) * 0.08) and np.allclose(y, (np.arange(42) + 0.5) * 0.08)
    peak = d["r8_depth_u"].max(axis=0) * 1e-6
    for g, (gx, gy) in enumerate(d["r8_gauge_xy"]):
        ix, iy = int(np.argmin(np.abs(x - gx))), int(np.argmin(np.abs(y - gy)))
        assert abs(hmax[iy, ix] - peak[g]) <= 5.0e-7 + 1e-12, g


def test_cf_decoding_rules():
    raw = np.array([[1, -999, 3]], dtype=np.int16)
    out = io_netcdf.Dataset._decode(raw, {"_FillValue": np.array([-999], dtype=np.int16),
                                          "scale_factor": np.array([0.5]), "add_offset": np.array([10.0])})
    assert np.isnan(out[0, 1]) and out[0, 0] == 10.5 and out[0, 2] == 11.5
    f32 = np.array([1.0, 9.0], dtype=np.float32)
    out = io_netcdf.Dataset._decode(f32, {"missing_value": np.array([9.0], dtype=np.float32)})
    assert out.dtype == np.float32 and np.isnan(out[1]) and out[0] == 1.0
    assert io_netcdf.Dataset._decode(f32, {"_FillValue": np.array([np.nan], dtype=np.float32)}) is not f32


def synthetic(values, x, y, dims=("y", "x")):
    v = io_netcdf.Variable("z", np.asarray(values, dtype=np.float64), dims, {})
    v.coords = {"x": np.asarray(x, dtype=np.float64), "y": np.asarray(y, dtype=np.float64)}
    return v


def test_remap_rules():
    x, y = np.linspace(0.0, 4.0, 9), np.linspace(0.0, 3.0, 7)
    plane = 0.3 + 0.5 * x[None, :] - 0.2 * y[:, None]
    v = synthetic(plane, x, y)
    inside = ModelGrid.from_bounds(0.5, 3.5, 0.5, 2.5, 0.25)
    out = io_netcdf._remap(v, inside, "x", "y", "linear", "error", "DEM")
    tx, ty = inside.x.numpy(), inside.y.numpy()
    assert np.allclose(out, 0.3 + 0.5 * tx[None, :] - 0.2 * ty[:, None], atol=1e-14, rtol=0)
    wide = ModelGrid.from_bounds(-1.0, 5.0, -1.0, 4.0, 0.5)
    with pytest.raises(ValueError, match="beyond DEM"):
        io_netcdf._remap(v, wide, "x", "y", "linear", "error", "DEM")
    nan = io_netcdf._remap(v, wide, "x", "y", "linear", "nan", "DEM")
    near = io_netcdf._remap(v, wide, "x", "y", "linear", "nearest", "DEM")
    zero = io_netcdf._remap(v, wide, "x", "y", "linear", "zero", "DEM")
    holes = np.isnan(nan)
    assert holes.any() and np.isfinite(near).all() and np.all(zero[holes] == 0.0)
    assert near[0, 0] == plane[0, 0] and near[-1, -1] == plane[-1, -1]
    # descending source coordinates and (x, y) storage order are normalised first
    flipped = io_netcdf._ascending(synthetic(plane[::-1].T.copy(), x, y[::-1], dims=("x", "y")), "x", "y")
    assert np.array_equal(flipped.values, plane) and np.array_equal(flipped.coords["y"], y)


def test_conservative_rainfall_remap():
    rng = np.random.default_rng(0)
    sx, sy = (np.arange(6) + 0.5) * 0.5, (np.arange(4) + 0.5) * 0.5
    vals = rng.uniform(0.0, 2.0, size=(3, 4, 6))
    target = ModelGrid.from_bounds(0.0, 3.0, 0.0, 2.0, 0.2)
    out = io_netcdf.conservative_rectilinear_remap(vals, sx, sy, target)
    assert out.shape == (3, 10, 15)
    assert np.allclose((out * 0.04).sum(axis=(1, 2)), (vals * 0.25).sum(axis=(1, 2)), rtol=1e-12)
    flat = io_netcdf.conservative_rectilinear_remap(np.full((1, 4, 6), 7.0), sx, sy, target)
    assert np.allclose(flat, 7.0, rtol=1e-13)
    with pytest.raises(ValueError, match="outside"):
        io_netcdf.conservative_rectilinear_remap(vals, sx, sy, ModelGrid.from_bounds(-0.2, 3.0, 0.0, 2.0, 0.2),
                                                 outside_domain="error")
    assert io_netcdf._unit_factor("mm/h") == 1e-3 / 3600 and io_netcdf._unit_factor("kg m-2 s-1") == 1e-3
    with pytest.raises(ValueError):
        io_netcdf._unit_factor("furlongs")
    t = io_netcdf.Variable("time", np.array([2.0, 3.0, 5.0]), ("time",), {"units": "hours since 2000-01-01"})
    assert np.array_equal(io_netcdf._time_seconds(t), [0.0, 3600.0, 10800.0])


def test_from_netcdf_constructor_and_errors(tmp_path):
    m = swe.SWE2D.from_netcdf(BATHY, model_grid={"xmin": 0.0, "xmax": 5.44, "ymin": 0.0, "ymax": 3.36,
                                                 "resolution": 0.08}, dem_outside_domain="nearest")
    assert tuple(m.bed.shape) == (1, 1, 42, 68) and m.dem_correction_parameter.requires_grad
    assert torch.all(m.roughness_length == 0.025) and m.dem_path == str(BATHY)
    with pytest.raises(KeyError, match="not found"):
        swe.SWE2D.from_netcdf(BATHY, model_grid=ModelGrid.from_bounds(0.0, 5.44, 0.0, 3.36, 0.08),
                              dem_variable="elevation")
    not_h5 = tmp_path / "plain.nc"
    not_h5.write_bytes(b"CDF\x01" + bytes(60))
    with pytest.raises(hdf5.HDF5Error):
        hdf5.File(not_h5)


def write_classic(path, variables, dims, attrs=None):
    """A NetCDF-3 file through scipy's writer: variables = {name: (dims, array, attrs)}."""
    from scipy.io import netcdf_file

    with netcdf_file(path, "w") as f:
        for k, v in (attrs or {}).items():
            setattr(f, k, v)
        for name, size in dims.items():
            f.createDimension(name, size)
        for name, (vdims, arr, vattrs) in variables.items():
            var = f.createVariable(name, arr.dtype.char, vdims)
            var[:] = arr
            for k, v in vattrs.items():
                setattr(var, k, v)
    return path


def test_classic_netcdf_dem_roughness_and_rainfall(tmp_path):
    x = np.linspace(0.0, 4.0, 9)
    y = np.linspace(3.0, 0.0, 7)  # stored north to south
    z = (0.3 + 0.5 * x[None, :] - 0.2 * y[:, None]).astype(np.float32)
    n = np.full((7, 9), 0.02, dtype=np.float32)
    dem = write_classic(tmp_path / "dem.nc", {"x": (("x",), x, {}), "y": (("y",), y, {}),
                                              "elevation": (("y", "x"), z, {"units": "m"}),
                                              "n": (("y", "x"), n, {})}, {"x": 9, "y": 7})
    ds = io_netcdf.open_dataset(dem)
    assert sorted(ds.data_vars) == ["elevation", "n"] and ds["elevation"].dims == ("y", "x")
    grid = ModelGrid.from_bounds(0.5, 3.5, 0.5, 2.5, 0.25)
    g = io_netcdf.load_dem_and_roughness_to_grid(dem, grid, roughness_variable="n")
    tx, ty = grid.x.numpy(), grid.y.numpy()
    plane = 0.3 + 0.5 * tx[None, :] - 0.2 * ty[:, None]
    assert np.allclose(g.bed.numpy(), plane, atol=2e-7, rtol=0)  # float32 source values
    assert np.allclose(g.roughness_length.numpy(), 0.02, rtol=1e-6) and bool(g.valid_mask.all())
    m = swe.SWE2D.from_netcdf(dem, model_grid=grid, roughness_variable="n", train_dem=False)
    assert torch.equal(m.bed_reference[0, 0], g.bed) and torch.allclose(m.roughness_length, g.roughness_length)
    hole = z.copy()
    hole[3, 4] = -9999.0
    bad = write_classic(tmp_path / "hole.nc", {"x": (("x",), x, {}), "y": (("y",), y, {}),
                                               "z": (("y", "x"), hole, {"_FillValue": np.float32(-9999.0)})},
                        {"x": 9, "y": 7})
    with pytest.raises(ValueError, match="missing values"):
        io_netcdf.load_dem_and_roughness_to_grid(bad, grid)
    # gridded rainfall: 2 records, mm/h, times in hours since a date; conservative remap keeps the total
    rx, ry = (np.arange(4) + 0.5) * 1.0, (np.arange(3) + 0.5) * 1.0
    rain = np.stack([np.full((3, 4), 3.6), np.arange(12.0).reshape(3, 4)]).astype(np.float64)
    rf = write_classic(tmp_path / "rain.nc", {"x": (("x",), rx, {}), "y": (("y",), ry, {}),
                                              "time": (("time",), np.array([1.0, 1.5]),
                                                       {"units": "hours since 2000-01-01"}),
                                              "rainfall": (("time", "y", "x"), rain, {"units": "mm/h"})},
                       {"x": 4, "y": 3, "time": 2})
    target = ModelGrid.from_bounds(0.0, 4.0, 0.0, 3.0, 0.5)
    forcing = io_netcdf.load_rainfall_to_grid(rf, target)
    assert forcing.times.tolist() == [0.0, 1800.0] and tuple(forcing.rates.shape) == (2, 1, 6, 8)
    assert torch.allclose(forcing.rates[0], torch.tensor(1e-6, dtype=torch.float64))
    vol = float(forcing.rates[1].sum()) * 0.25
    assert abs(vol - rain[1].sum() * 1e-3 / 3600) < 1e-15


def test_hdf5_filter_pipeline_pieces():
    import zlib

    data = np.arange(40, dtype="<f8") * 0.37
    raw = data.tobytes()
    shuffled = np.frombuffer(raw, dtype=np.uint8).reshape(-1, 8).T.tobytes()
    packed = zlib.compress(shuffled) + b"\x01\x02\x03\x04"
    filters = [(2, [8]), (1, [6]), (3, [])]  # shuffle, deflate, Fletcher-32, applied in that order on write
    out = hdf5.File._unfilter(packed, filters, 0, 8)
    assert np.array_equal(np.frombuffer(out, dtype="<f8"), data)
    # a set mask bit skips that filter (here Fletcher-32 was not applied)
    assert hdf5.File._unfilter(zlib.compress(shuffled), filters, 0b100, 8) == raw
    with pytest.raises(NotImplementedError, match="filter id 32001"):
        hdf5.File._unfilter(raw, [(32001, [])], 0, 8)

    class Stub:
        buf = memoryview(bytes([1, 2, 0, 0, 0, 0, 0, 0,            # version 1, two filters, reserved
                                2, 0, 0, 0, 1, 0, 1, 0, 8, 0, 0, 0, 0, 0, 0, 0,   # shuffle: 1 value, padded
                                1, 0, 8, 0, 1, 0, 1, 0]) + b"deflate\0" + bytes([4, 0, 0, 0, 0, 0, 0, 0]))

    assert hdf5.File._filter_pipeline(Stub(), 0) == [(2, [8]), (1, [4])]
