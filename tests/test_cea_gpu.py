"""End to end: the reference's CEA 2008 C1 forward example, rerun through this package.

Examples/forward/Cea2008/Cea2008_C1_new/example_run.py loads a 208 x 256
ESRI-ASCII DEM onto a 200 x 254 model grid (linear regridding, NODATA filled
from the nearest cell), rains 317 mm/h for 45 s by adding rate * 0.1 s of
depth before each 0.1 s chunk (:128-131), and advances 1200 chunks in FP64
with walls on three sides and a transmissive bottom (Manning n = 0.012),
recording every step at four gauges (:92-107, solver.py:781-801), the
bottom-boundary discharge before each chunk (:133-140) and the domain maximum
depth after it (:151-154).  The run on an A100 wrote those records
(gauge_*.txt, bottom_boundary_discharge.txt, diffswe_run.log); the fixture
holds them as printed (tests/golden/make_cea_fixture.py).

Here the same driver runs through ``load_dynamic_model`` -> ``SWE2D.from_ascii``
and ``SWE2D.forward`` on the GPU, and every record is compared with what the
reference printed.  Bars (printed values are rounded to 5e-7, discharge to
5e-9):
  * the same number of steps (80052, after the 759 closing steps under 5e-7 s
    that the reference's compiled run added are set aside);
  * step end times and dt within print rounding + 1e-8;
  * gauge depth and h + z within 1.5e-6 (one unit in the last printed digit),
    and beyond print rounding in at most 1 % of the compared rows;
  * domain maximum depth after each chunk within 1.5e-6, minimum within print
    rounding; bottom discharge before each chunk within print rounding + 1e-10.
Measured on MI355X: time 5.03e-7, dt 5.01e-7, depth 1.07e-6 (42 of 5124 rows
past rounding), maximum depth 8.4e-7, discharge 5.0e-9 (none past rounding).
"""
import gzip
from pathlib import Path

import numpy as np
import pytest
import torch

import diffswe2d_amd as swe

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"
DT_OUT = 0.1


def run_cea_c1(tmp_path, device, n_chunks=1200):
    d = np.load(GOLDEN / "cea_c1.npz")
    topo = tmp_path / "topo.asc"
    topo.write_bytes(gzip.decompress((GOLDEN / "cea_c1_topo.asc.gz").read_bytes()))
    rain = tmp_path / "rain.txt"
    np.savetxt(rain, d["rain"])
    grid = swe.ModelGrid.from_bounds(xmin=0.02, xmax=2.02, ymin=-0.005, ymax=2.535, resolution=0.01,
                                     dtype=torch.float64, device=device)
    model, use_rain, rain_t, rain_r = swe.load_dynamic_model(
        dem_filepath=str(topo), rain_filepath=str(rain), model_grid=grid, device=device, boundary_left="wall",
        boundary_right="wall", boundary_top="wall", boundary_bottom="transmissive", frictionmodel="manning",
        default_roughness_value=0.012)
    assert use_rain
    ny, nx = model.bed.shape[-2:]
    assert (ny, nx) == (254, 200)
    model.track_gauges = True
    for g, (gx, gy) in enumerate(d["gauge_xy"]):
        ix = int(torch.argmin(torch.abs(grid.x - gx)))
        iy = int(torch.argmin(torch.abs(grid.y - gy)))
        model.gauge_indices[f"gauge_{g}"] = (iy, ix)
        model.bed_elevations[f"gauge_{g}"] = float(model.bed[0, 0, iy, ix])
    model.reset_gauge_tracking()
    model.eval()
    U = torch.zeros(1, 3, ny, nx, dtype=torch.float64, device=device)
    hmax = U[:, 0].clone()
    t, discharge, log = 0.0, [], []
    with torch.inference_mode():
        while t < DT_OUT * n_chunks - 1e-6:
            U[:, 0, :, :] += float(np.interp(t, rain_t, rain_r)) * DT_OUT
            discharge.append(float(torch.sum(-U[0, 2, 5, 50:150])) * model.dx)
            U, h = model(U, t_end=DT_OUT, start_time=t)
            hmax = torch.maximum(hmax, h)
            t += DT_OUT
            log.append((float(U[:, 0].max()), float(U[:, 0].min())))
    names = list(model.gauge_indices)
    rec = dict(time=np.array(model.ts_time), dt=np.array(model.ts_dt),
               depth=np.stack([np.array(model.ts_h[n]) for n in names], axis=1),
               level=np.stack([np.array(model.ts_z[n]) for n in names], axis=1))
    return d, rec, np.array(discharge), np.array(log), hmax


def test_cea_c1_matches_reference_records(tmp_path, cuda_device):
    d, rec, q, log, hmax = run_cea_c1(tmp_path, cuda_device)
    n = rec["dt"].size
    rows = d["rows"]
    ref_t, ref_dt = d["time_dt_u"][:, 0] * 1e-6, d["time_dt_u"][:, 1] * 1e-6
    ref_depth, ref_level = d["depth_u"] * 1e-6, d["level_u"] * 1e-6
    print(f"steps {n} (reference {int(d['n_steps'])} + {int(d['n_closing'])} closing steps under 5e-7 s)")
    common = rows[rows < n]
    k = common.size
    err_t = np.abs(rec["time"][common] - ref_t[:k])
    err_dt = np.abs(rec["dt"][common] - ref_dt[:k])
    err_h = np.abs(rec["depth"][common] - ref_depth[:k])
    err_z = np.abs(rec["level"][common] - ref_level[:k])
    for name, e in (("time", err_t), ("dt", err_dt), ("depth", err_h), ("h+z", err_z)):
        off = (e > 5.0e-7 + 1e-12).reshape(k, -1).any(axis=1)
        first = common[int(np.argmax(off))] if off.any() else None
        print(f"  {name}: max |diff| {e.max():.3e}, rows beyond print rounding {int(off.sum())} of {k}, "
              f"first at step {first}")
    err_log = np.abs(log[:, 0] - d["log_u"][:, 2] * 1e-6)
    err_min = np.abs(log[:, 1] - d["log_u"][:, 3] * 1e-6)
    err_q = np.abs(q - d["discharge"][:, 1])
    print(f"  log max depth: max |diff| {err_log.max():.3e} (beyond rounding "
          f"{(err_log > 5.0e-7 + 1e-12).sum()}); min depth {err_min.max():.3e}")
    print(f"  discharge: max |diff| {err_q.max():.3e} (scale {np.abs(d['discharge'][:, 1]).max():.3e}, beyond "
          f"rounding {(err_q > 5.0e-9 + 1e-14).sum()})")
    print(f"  hmax max {float(hmax.max()):.6f}")
    assert n == int(d["n_steps"])
    assert err_t.max() <= 5.0e-7 + 1e-8 and err_dt.max() <= 5.0e-7 + 1e-8
    assert err_h.max() <= 1.5e-6 and err_z.max() <= 1.5e-6
    assert (err_h > 5.0e-7 + 1e-12).any(axis=1).sum() <= 0.01 * k
    assert err_log.max() <= 1.5e-6 and err_min.max() <= 5.0e-7 + 1e-8
    assert err_q.max() <= 5.0e-9 + 1e-10
    assert torch.isfinite(hmax).all() and float(hmax.min()) >= 0.0
    # the run's maximum-depth field (hmax_n_0012.nc): every second row and column, plus whole-field sums
    m = np.load(GOLDEN / "monai.npz")
    h = hmax[0].cpu().numpy()
    e = np.abs(h[::2, ::2] - m["cea_hmax_sub_n"] * 1e-7)
    s_ref, m_ref, wet_ref, _ = m["cea_hmax_stats"]
    print(f"  hmax field: max |diff| {e.max():.3e}; sum rel {abs(h.sum() - s_ref) / s_ref:.2e}; max "
          f"{h.max():.7f} vs {m_ref:.7f}; cells above 1e-6 m {int((h > 1e-6).sum())} vs {int(wet_ref)}")
    assert e.max() <= 2e-6 + 5e-8 and abs(h.sum() - s_ref) <= 1e-6 * s_ref
