"""Pack the reference's Monai Valley forward and inversion examples as test fixtures.

Run in the development container only (it reads the reference tree):

    python tests/golden/make_monai_fixture.py [--examples /root/reference/Examples/forward]

Inputs and outputs of Examples/forward/Monai_Valley/Monai_Valley_res001 and
_res008 (driver example_run.py in each): the NetCDF-4 bathymetry
(Monai_Bathy.nc, identical in both), the input wave imposed as a dynamic
water level on the left boundary (MonaiValley_InputWave.txt), and what the
FP64 runs of 30 s on an A100 wrote: per-step gauge records (gauge_*.txt,
%.6f), the maximum depth field (hmax.nc) and the final maximum depth
(diffswe_run.log).  Only data files are read, through this package's own
NetCDF reader; nothing is run.

The inversion example (Examples/inverse/Monai_Valley/
Monai_Velley_BC_inversion_spline_lr002, bc_inversion_train.py) contributes
its laboratory gauge record (MonaiValley_WaveGages_30s_02s.txt, metres every
0.2 s) and the per-epoch history its A100 run wrote
(inverse_results_spline/loss_history.txt); its six-gauge variant
(..._lr002_6simg: simulated gauges in sim_gauges.txt, five of them for
training) contributes the same two files.

Written next to this script:
  monai_bathy.nc, monai_input_wave.txt   the inputs, byte for byte;
  monai_wave_gauges_02s.txt, monai_inversion_loss_history.txt,
  monai_sim_gauges.txt, monai_inversion_6simg_loss_history.txt
                                          the two inversions' observations and
                                          histories, byte for byte;
  monai_hmax_res008.nc                    the 0.08 m run's hmax.nc, byte for byte;
  monai.npz   gauge records (integers in units of 1e-6, as printed; the 0.01 m
              run keeps its first 1000 steps and every 5th after that), every
              second row and column of the 0.01 m run's hmax field (quantised
              to 1e-7 m) with the full field's sum, maximum, count above
              1e-6 m and sum of squares, the final maximum depth of each run.
"""
from __future__ import annotations

import argparse
import re
import shutil
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))

from diffswe2d_amd import hdf5  # noqa: E402


def gauges(folder: Path):
    coords, series = [], []
    for path in sorted(folder.glob("gauge_*.txt"), key=lambda p: int(p.stem.split("_")[1])):
        lines = path.read_text().splitlines()
        m = re.search(r"X=([-\d.]+), Y=([-\d.]+)", lines[0])
        coords.append((float(m.group(1)), float(m.group(2))))
        series.append(np.array([[int(round(float(v) * 1e6)) for v in ln.split()] for ln in lines[2:]],
                               dtype=np.int64))
    for s in series[1:]:
        assert s.shape == series[0].shape and (s[:, :2] == series[0][:, :2]).all()
    return np.array(coords), series


def final_max_depth(folder: Path) -> int:
    m = re.findall(r"Max Depth: ([\d.]+)m", (folder / "diffswe_run.log").read_text())
    return int(round(float(m[-1]) * 1e6))


def hmax_field(path: Path) -> np.ndarray:
    with hdf5.File(path) as f:
        return f.dataset("hmax").read()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--examples", default="/root/reference/Examples/forward")
    ex = Path(ap.parse_args().examples)
    r1, r8 = ex / "Monai_Valley" / "Monai_Valley_res001", ex / "Monai_Valley" / "Monai_Valley_res008"
    shutil.copyfile(r1 / "Monai_Bathy.nc", HERE / "monai_bathy.nc")
    shutil.copyfile(r1 / "MonaiValley_InputWave.txt", HERE / "monai_input_wave.txt")
    shutil.copyfile(r8 / "hmax.nc", HERE / "monai_hmax_res008.nc")
    inv = ex.parent / "inverse" / "Monai_Valley" / "Monai_Velley_BC_inversion_spline_lr002"
    shutil.copyfile(inv / "MonaiValley_WaveGages_30s_02s.txt", HERE / "monai_wave_gauges_02s.txt")
    shutil.copyfile(inv / "inverse_results_spline" / "loss_history.txt", HERE / "monai_inversion_loss_history.txt")
    inv6 = inv.parent / "Monai_Velley_BC_inversion_spline_lr002_6simg"
    shutil.copyfile(inv6 / "sim_gauges.txt", HERE / "monai_sim_gauges.txt")
    shutil.copyfile(inv6 / "inverse_results_spline" / "loss_history.txt",
                    HERE / "monai_inversion_6simg_loss_history.txt")
    out = {}
    for tag, folder, full, stride in (("r1", r1, 1000, 5), ("r8", r8, None, 1)):
        xy, series = gauges(folder)
        n = series[0].shape[0]
        keep = np.arange(n) if full is None else np.unique(
            np.concatenate((np.arange(full), np.arange(full, n, stride), [n - 1])))
        out[f"{tag}_gauge_xy"] = xy
        out[f"{tag}_n_steps"] = np.int64(n)
        out[f"{tag}_rows"] = keep
        out[f"{tag}_time_dt_u"] = series[0][keep, :2]
        out[f"{tag}_depth_u"] = np.stack([s[keep, 2] for s in series], axis=1)
        out[f"{tag}_level_u"] = np.stack([s[keep, 3] for s in series], axis=1)
        out[f"{tag}_final_max_u"] = np.int64(final_max_depth(folder))
    h = hmax_field(r1 / "hmax.nc")
    out["r1_hmax_sub_n"] = np.round(h[::2, ::2] * 1e7).astype(np.int32)
    out["r1_hmax_stats"] = np.array([h.sum(), h.max(), float((h > 1e-6).sum()), float(np.sum(h * h))])
    np.savez_compressed(HERE / "monai.npz", **out)
    print({k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
