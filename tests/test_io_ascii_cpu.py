"""ESRI-ASCII input path (src/io_ascii.py, src/model_loader.py:68-118) on the CPU.

Pinned by closed forms (bilinear interpolation reproduces a plane exactly;
corner and centre origins describe the same raster; the first file row is the
northern one), hand-built NODATA and out-of-domain cases, and the reference's
own CEA C1 example: its printed gauge records give the bed at the four gauge
cells (h + z minus h), which the regridded DEM must reproduce.
"""
import gzip
from pathlib import Path

import numpy as np
import pytest
import torch

import diffswe2d_amd as swe
from diffswe2d_amd import io_ascii
from diffswe2d_amd.grid import ModelGrid

GOLDEN = Path(__file__).resolve().parent / "golden"


def write_asc(path, values, x0=0.0, y0=0.0, res=1.0, origin="corner", nodata=None, sep="\t"):
    """values[0] is the northern row, as in the file."""
    rows, cols = values.shape
    head = [f"ncols {cols}", f"nrows {rows}", f"xll{origin} {x0}", f"yll{origin} {y0}", f"cellsize {res}"]
    if nodata is not None:
        head.append(f"NODATA_value {nodata}")
    body = [sep.join(repr(float(v)) for v in row) for row in values]
    path.write_text("\n".join(head + body) + "\n")
    return path


def plane(x, y):
    return 0.25 + 0.5 * x - 0.125 * y


def test_linear_regrid_reproduces_a_plane(tmp_path):
    res, cols, rows = 0.5, 9, 7
    xc = (np.arange(cols) + 0.5) * res
    yc = (np.arange(rows) + 0.5) * res
    north_first = plane(xc[None, :], yc[::-1, None])
    f = write_asc(tmp_path / "p.asc", north_first, res=res)
    grid = {"xmin": 0.5, "xmax": 4.0, "ymin": 0.25, "ymax": 3.0, "resolution": 0.25}
    out, spec = io_ascii.ascii_to_tensor(f, grid, dtype=torch.float64)
    assert tuple(out.shape) == (11, 14) == spec.shape
    gx, gy = np.meshgrid(spec.x, spec.y, indexing="xy")
    assert np.allclose(out.numpy(), plane(gx, gy), rtol=0, atol=1e-13)


def test_orientation_and_origin_conventions(tmp_path):
    vals = np.arange(12, dtype=float).reshape(3, 4)
    r1 = io_ascii.read_ascii(write_asc(tmp_path / "a.asc", vals, x0=10.0, y0=20.0, res=2.0))
    r2 = io_ascii.read_ascii(write_asc(tmp_path / "b.asc", vals, x0=11.0, y0=21.0, res=2.0, origin="center"))
    assert np.array_equal(r1.values[0], vals[-1]) and np.array_equal(r1.values[-1], vals[0])
    for r in (r1, r2):
        assert (r.xmin, r.xmax, r.ymin, r.ymax) == (10.0, 18.0, 20.0, 26.0)
        assert np.array_equal(r.x, [11.0, 13.0, 15.0, 17.0]) and np.array_equal(r.y, [21.0, 23.0, 25.0])
    assert io_ascii.grids_are_equal(r1, {"xmin": 10, "xmax": 18, "ymin": 20, "ymax": 26, "resolution": 2})
    native, spec = io_ascii.read_ascii_native_tensor(tmp_path / "a.asc", dtype=torch.float64)
    assert np.array_equal(native.numpy(), vals[::-1]) and spec.shape == (3, 4)


def test_nodata_filled_from_nearest_and_outside_policies(tmp_path):
    vals = np.array([[1.0, 2.0, 3.0, 4.0],
                     [5.0, -9999.0, 7.0, 8.0],
                     [9.0, 10.0, 11.0, 12.0]])
    f = write_asc(tmp_path / "n.asc", vals, nodata=-9999)
    r = io_ascii.read_ascii(f)
    assert np.isnan(r.values[1, 1]) and np.isfinite(r.values).sum() == 11
    info = io_ascii.describe_ascii(f)
    assert info["invalid_cell_count"] == 1 and info["minimum"] == 1.0 and info["maximum"] == 12.0
    # the hole at centre (1.5, 1.5) has four valid neighbours at distance 1; nearest-fill takes one of them
    filled = io_ascii.fill_nodata_nearest(r)
    assert filled[1, 1] in (2.0, 5.0, 7.0, 10.0)
    same_grid = {"xmin": 0.0, "xmax": 4.0, "ymin": 0.0, "ymax": 3.0, "resolution": 1.0}
    out = io_ascii.regrid_ascii(r, same_grid, method="nearest")
    assert np.array_equal(out, filled)
    with pytest.raises(ValueError, match="NaN"):
        io_ascii.regrid_ascii(r, same_grid, method="linear", fill_internal_nodata=False)
    # a grid reaching half a cell past the raster's last centres on every side
    wide = {"xmin": -0.5, "xmax": 4.5, "ymin": -0.5, "ymax": 3.5, "resolution": 0.5}
    with pytest.raises(ValueError, match="could not be assigned"):
        io_ascii.regrid_ascii(r, wide, outside_domain="error")
    out_nan = io_ascii.regrid_ascii(r, wide, outside_domain="nan")
    out_near = io_ascii.regrid_ascii(r, wide, outside_domain="nearest")
    holes = np.isnan(out_nan)
    assert holes.any() and np.isfinite(out_near).all()
    assert np.array_equal(out_near[~holes], out_nan[~holes])
    assert out_near[0, 0] == filled[0, 0] and out_near[-1, -1] == filled[-1, -1]


def test_header_and_grid_errors(tmp_path):
    bad = tmp_path / "bad.asc"
    bad.write_text("ncols 2\nnrows 2\nxllcorner 0\nyllcorner 0\n1 2\n3 4\n")
    with pytest.raises(ValueError, match="missing entries"):
        io_ascii.read_ascii(bad)
    bad.write_text("ncols 2\nnrows 2\ncellsize 1\n1 2\n3 4\n")
    with pytest.raises(ValueError, match="xllcorner"):
        io_ascii.read_ascii(bad)
    bad.write_text("ncols 3\nnrows 2\nxllcorner 0\nyllcorner 0\ncellsize 1\n1 2\n3 4\n")
    with pytest.raises(ValueError, match="do not match"):
        io_ascii.read_ascii(bad)
    with pytest.raises(FileNotFoundError):
        io_ascii.read_ascii(tmp_path / "missing.asc")
    with pytest.raises(ValueError, match="not divisible"):
        io_ascii.GridSpec(0.0, 1.05, 0.0, 1.0, 0.1)
    with pytest.raises(KeyError):
        io_ascii.parse_model_grid({"xmin": 0, "xmax": 1})
    with pytest.raises(TypeError):
        io_ascii.parse_model_grid(42)
    with pytest.raises(ValueError, match="divisible"):
        ModelGrid.from_bounds(0.0, 1.05, 0.0, 1.0, 0.1)
    with pytest.raises(ValueError, match="two cells"):
        ModelGrid.from_bounds(0.0, 0.1, 0.0, 1.0, 0.1)
    with pytest.raises(ValueError, match="uniform"):
        ModelGrid.from_cell_centres([0.0, 1.0, 3.0], [0.0, 1.0, 2.0])
    with pytest.raises(ValueError, match="dx == dy"):
        ModelGrid.from_cell_centres([0.0, 1.0, 2.0], [0.0, 2.0, 4.0])


def test_model_grid_forms_agree():
    g = ModelGrid.from_bounds(xmin=0.02, xmax=2.02, ymin=-0.005, ymax=2.535, resolution=0.01)
    assert (g.ny, g.nx) == (254, 200)
    a = io_ascii.parse_model_grid(g)
    b = io_ascii.parse_model_grid({"xmin": 0.02, "xmax": 2.02, "ymin": -0.005, "ymax": 2.535, "resolution": 0.01})
    assert a.shape == b.shape == (254, 200)
    assert np.allclose(a.x, b.x, rtol=0, atol=1e-12) and np.allclose(a.y, b.y, rtol=0, atol=1e-12)
    c = ModelGrid.from_cell_centres(g.x, g.y)
    assert c.dx == pytest.approx(0.01) and c.nx == 200


def cea_files(tmp_path):
    d = np.load(GOLDEN / "cea_c1.npz")
    topo = tmp_path / "topo.asc"
    topo.write_bytes(gzip.decompress((GOLDEN / "cea_topo.asc.gz").read_bytes()))
    rain = tmp_path / "rain.txt"
    np.savetxt(rain, d["rain"])
    return d, topo, rain


def test_cea_dem_reproduces_the_reference_gauge_beds(tmp_path):
    d, topo, rain = cea_files(tmp_path)
    grid = ModelGrid.from_bounds(xmin=0.02, xmax=2.02, ymin=-0.005, ymax=2.535, resolution=0.01)
    model, use_rain, rt, rr = swe.load_dynamic_model(
        dem_filepath=str(topo), rain_filepath=str(rain), model_grid=grid, device="cpu", boundary_left="wall",
        boundary_right="wall", boundary_top="wall", boundary_bottom="transmissive", frictionmodel="manning",
        default_roughness_value=0.012)
    assert use_rain and np.allclose(rr, d["rain"][:, 1] / 3.6e6) and np.array_equal(rt, d["rain"][:, 0])
    assert model.bed.dtype == torch.float64 and tuple(model.bed.shape) == (1, 1, 254, 200)
    assert model.cfg.boundary_bottom == "transmissive" and model.cfg.frictionmodel == "manning"
    assert torch.all(model.roughness_length == 0.012)
    assert model.maximum_dem_correction == 1.0 and not model.dem_correction_parameter.requires_grad
    raster = io_ascii.read_ascii(topo)
    assert raster.shape == (256, 208) and np.isnan(raster.values).any()
    assert torch.isfinite(model.bed).all()
    z_ref = (d["level_u"][0] - d["depth_u"][0]) * 1e-6
    for g, (gx, gy) in enumerate(d["gauge_xy"]):
        ix = int(torch.argmin(torch.abs(grid.x - gx)))
        iy = int(torch.argmin(torch.abs(grid.y - gy)))
        assert abs(float(model.bed[0, 0, iy, ix]) - z_ref[g]) <= 1.5e-6, g


def test_unsupported_dem_formats(tmp_path):
    grid = ModelGrid.from_bounds(0.0, 1.0, 0.0, 1.0, 0.5)
    with pytest.raises(FileNotFoundError):
        swe.load_dynamic_model(dem_filepath=str(tmp_path / "dem.nc"), model_grid=grid, device="cpu")
    with pytest.raises(ValueError, match="Unsupported DEM format"):
        swe.load_dynamic_model(dem_filepath="dem.tif", model_grid=grid, device="cpu")
    with pytest.raises(TypeError, match="model_grid"):
        swe.SWE2D.from_netcdf("dem.nc", model_grid=object())


def test_from_ascii_roughness_raster(tmp_path):
    z = np.array([[0.0, 0.1], [0.2, 0.3]])
    n = np.array([[0.01, 0.02], [0.03, 0.04]])
    fz = write_asc(tmp_path / "z.asc", z)
    fn = write_asc(tmp_path / "n.asc", n)
    grid = {"xmin": 0.0, "xmax": 2.0, "ymin": 0.0, "ymax": 2.0, "resolution": 1.0}
    m = swe.SWE2D.from_ascii(fz, model_grid=grid, roughness_path=fn, dtype=torch.float64)
    assert torch.equal(m.bed_reference[0, 0], torch.tensor(z[::-1].copy()))
    assert torch.equal(m.roughness_length[0, 0], torch.tensor(n[::-1].copy()))
    assert m.dx == m.dy == 1.0 and m.grid_spec.shape == (2, 2) and m.dem_path == str(fz)
    m2 = swe.SWE2D.from_ascii(fz, model_grid=grid)
    assert m2.roughness_length.dtype == torch.float32 and torch.all(m2.roughness_length == 0.03)
