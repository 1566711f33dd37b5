"""End to end: the reference's Monai Valley forward examples, rerun through this package.

Examples/forward/Monai_Valley/Monai_Valley_res001 and _res008
(example_run.py) load the NetCDF-4 bathymetry (Monai_Bathy.nc, 394 x 246
points at 0.014 m) onto 550 x 343 cells of 0.01 m or 68 x 42 cells of 0.08 m
(linear remap, nearest extrapolation), impose the laboratory input wave
(MonaiValley_InputWave.txt) as a dynamic water level on the left boundary with
walls elsewhere (Manning n = 0.01), start from a flat surface at 0 m
(:84-87) and run one 30 s ``forward`` in FP64, recording every step at three
(six) gauges and the maximum depth field (:129-160).  The A100 runs wrote
gauge_*.txt (%.6f), hmax.nc and diffswe_run.log; the fixtures hold them
(tests/golden/make_monai_fixture.py).

Here the same driver runs through ``load_dynamic_model`` -> ``SWE2D.from_netcdf``
(this package's own NetCDF-4 reader) and ``SWE2D.forward`` on the GPU.
Bars: the same number of steps; dt within print rounding + 1e-8 and step
end times (sums of dt) within print rounding + 5e-8; gauge depth and h + z
within 1.5e-6 (one unit in the last printed digit), past print rounding in
at most 1 % of the compared rows;
the maximum-depth field within 2e-6 m cell by cell (every second row and
column for the 0.01 m run, all cells at 0.08 m), its sum within 1e-9
relative; the final maximum depth within print rounding + 1e-8.
Measured on MI355X: 0.08 m run, 1786 steps, times 5.20e-7, depth 5.00e-7,
maximum-depth field 2.4e-10 in every cell (sum 3e-11 relative); 0.01 m run,
14555 steps, times
5.08e-7, depth 5.00e-7 (1 of 3712 rows past rounding, by 1e-10), sampled
maximum-depth field 6.4e-8 (of which 5e-8 is the fixture's quantisation), sum
4e-11 relative, the same 176042 cells above 1e-6 m.
"""
from pathlib import Path

import numpy as np
import pytest
import torch

import diffswe2d_amd as swe
from diffswe2d_amd import hdf5

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"
RUNS = {"r8": (5.44, 3.36, 0.08), "r1": (5.5, 3.43, 0.01)}


def run_monai(tag, device):
    d = np.load(GOLDEN / "monai.npz")
    xmax, ymax, res = RUNS[tag]
    grid = swe.ModelGrid.from_bounds(xmin=0.0, xmax=xmax, ymin=0.0, ymax=ymax, resolution=res,
                                     dtype=torch.float64, device=device)
    model, use_rain, _, _ = swe.load_dynamic_model(
        dem_filepath=str(GOLDEN / "monai_bathy.nc"), model_grid=grid, device=device,
        boundary_left="water_level", boundary_right="wall", boundary_top="wall", boundary_bottom="wall",
        frictionmodel="manning", default_roughness_value=0.01, tide_filepath=str(GOLDEN / "monai_input_wave.txt"))
    assert not use_rain and model.use_dynamic_bc
    ny, nx = model.bed.shape[-2:]
    U0 = torch.zeros(1, 3, ny, nx, dtype=torch.float64, device=device)
    U0[:, 0:1] = torch.clamp(0.0 - model.bed, min=0.0)
    model.track_gauges = True
    for g, (gx, gy) in enumerate(d[f"{tag}_gauge_xy"]):
        ix = int(torch.argmin(torch.abs(grid.x - gx)))
        iy = int(torch.argmin(torch.abs(grid.y - gy)))
        model.gauge_indices[f"gauge_{g}"] = (iy, ix)
        model.bed_elevations[f"gauge_{g}"] = float(model.bed[0, 0, iy, ix])
    model.reset_gauge_tracking()
    model.eval()
    with torch.inference_mode():
        U = U0.clone()
        hmax = U0[:, 0].clone()
        U, h = model(U, t_end=30.0, start_time=0.0)
        hmax = torch.maximum(hmax, h)
    names = list(model.gauge_indices)
    rec = dict(time=np.array(model.ts_time), dt=np.array(model.ts_dt),
               depth=np.stack([np.array(model.ts_h[n]) for n in names], axis=1),
               level=np.stack([np.array(model.ts_z[n]) for n in names], axis=1))
    return d, rec, U, hmax[0].cpu().numpy()


def check_records(tag, d, rec):
    n = rec["dt"].size
    rows = d[f"{tag}_rows"]
    k = rows.size
    print(f"{tag}: steps {n} (reference {int(d[f'{tag}_n_steps'])})")
    assert n == int(d[f"{tag}_n_steps"])
    ref_td = d[f"{tag}_time_dt_u"] * 1e-6
    errs = dict(time=np.abs(rec["time"][rows] - ref_td[:, 0]), dt=np.abs(rec["dt"][rows] - ref_td[:, 1]),
                depth=np.abs(rec["depth"][rows] - d[f"{tag}_depth_u"] * 1e-6),
                level=np.abs(rec["level"][rows] - d[f"{tag}_level_u"] * 1e-6))
    for name, e in errs.items():
        off = (e > 5.0e-7 + 1e-12).reshape(k, -1).any(axis=1)
        print(f"  {name}: max |diff| {e.max():.3e}, rows past print rounding {int(off.sum())} of {k}")
    assert errs["time"].max() <= 5.0e-7 + 5e-8 and errs["dt"].max() <= 5.0e-7 + 1e-8
    assert errs["depth"].max() <= 1.5e-6 and errs["level"].max() <= 1.5e-6
    assert (errs["depth"] > 5.0e-7 + 1e-12).any(axis=1).sum() <= 0.01 * k


@pytest.mark.parametrize("tag", ["r8", "r1"])
def test_monai_matches_reference_records(tag, cuda_device):
    d, rec, U, hmax = run_monai(tag, cuda_device)
    check_records(tag, d, rec)
    final_max = float(U[:, 0].max())
    print(f"  final max depth {final_max:.7f} vs {int(d[f'{tag}_final_max_u']) * 1e-6:.6f}")
    assert abs(final_max - int(d[f"{tag}_final_max_u"]) * 1e-6) <= 5.0e-7 + 1e-8
    if tag == "r8":
        with hdf5.File(GOLDEN / "monai_hmax_res008.nc") as f:
            ref = f.dataset("hmax").read()
        assert ref.shape == hmax.shape
        e = np.abs(hmax - ref)
        print(f"  hmax field: max |diff| {e.max():.3e} over {ref.size} cells; sum rel "
              f"{abs(hmax.sum() - ref.sum()) / ref.sum():.2e}")
        assert e.max() <= 2e-6 and abs(hmax.sum() - ref.sum()) <= 1e-9 * ref.sum()
    else:
        sub = d["r1_hmax_sub_n"] * 1e-7
        e = np.abs(hmax[::2, ::2] - sub)
        s_ref, m_ref, wet_ref, sq_ref = d["r1_hmax_stats"]
        print(f"  hmax field: max |diff| {e.max():.3e} on {sub.size} sampled cells; sum rel "
              f"{abs(hmax.sum() - s_ref) / s_ref:.2e}; max {hmax.max():.7f} vs {m_ref:.7f}; cells above 1e-6 m "
              f"{int((hmax > 1e-6).sum())} vs {int(wet_ref)}")
        assert e.max() <= 2e-6 + 5e-8
        assert abs(hmax.sum() - s_ref) <= 1e-9 * s_ref and abs(hmax.max() - m_ref) <= 2e-6
