"""Pack the reference's CEA 2008 C1 and C3 forward example outputs as test fixtures.

Run in the development container only (it reads the reference tree):

    python tests/golden/make_cea_fixture.py [--examples /root/reference/Examples/forward/Cea2008]

Inputs and outputs of Examples/forward/Cea2008/Cea2008_C1_new and
Cea2008_C3_new (driver example_run.py in each): a 208 x 256 ESRI-ASCII DEM
(topo.asc, identical in both), a rainfall table (C1: cstRain-test1.txt,
317 mm/h for 45 s; C3: cstRain-test3.txt, two 328 mm/h pulses), and what the
FP64 runs on an A100 wrote: per-step gauge records at four cells
(gauge_*.txt: time, dt, depth, h + z, printed %.6f), the bottom-boundary
discharge every 0.1 s (bottom_boundary_discharge.txt, %.8f), the domain
maximum depth every 0.1 s (diffswe_run.log, %.6f) and the maximum depth field
(hmax_n_0012.nc, read with this package's own NetCDF reader).  Only data
files are read; nothing is run.

Written next to this script:
  cea_topo.asc.gz    the DEM, byte for byte, gzip-compressed;
  cea_c1.npz, cea_c3.npz
                     rainfall table, gauge coordinates, discharge columns,
                     the gauge records of the first 2000 steps and every 25th
                     after that (as integers in units of 1e-6, exactly what was
                     printed), the step count, the discharge and log series,
                     every second row and column of the hmax field (quantised
                     to 1e-7 m) with the full field's sum, maximum, count above
                     1e-6 m and sum of squares.

Both runs close most of their 1200 chunks of 0.1 s with one extra step whose
dt prints as 0.000000 (under 5e-7 s, at the chunk's end time): 759 in C1,
767 in C3.  The reference's eager FP64 loop (src/solver.py:738-768) reaches
each chunk end exactly and takes no such step; those runs were made through
torch.compile.  The fixture drops those rows (their count is kept), so the
remaining records compare step for step with an eager-semantics run.
"""
from __future__ import annotations

import argparse
import gzip
import re
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))

from diffswe2d_amd import hdf5  # noqa: E402

FULL_STEPS = 2000
STRIDE = 25
CASES = {"c1": ("Cea2008_C1_new", "cstRain-test1.txt", (50, 150)),
         "c3": ("Cea2008_C3_new", "cstRain-test3.txt", (40, 160))}


def micro(text_rows):
    """%.6f columns -> int64 micro-units, exact."""
    return np.array([[int(round(float(v) * 1e6)) for v in row] for row in text_rows], dtype=np.int64)


def pack(ex: Path, rain_file: str, q_cols) -> dict:
    rain = np.loadtxt(ex / rain_file)
    coords, series = [], []
    for g in range(4):
        lines = (ex / f"gauge_{g}.txt").read_text().splitlines()
        m = re.search(r"X=([-\d.]+), Y=([-\d.]+)", lines[0])
        coords.append((float(m.group(1)), float(m.group(2))))
        series.append(micro(line.split() for line in lines[2:]))
    for s in series[1:]:
        assert s.shape[0] == series[0].shape[0] and (s[:, :2] == series[0][:, :2]).all()
    closing = series[0][:, 1] == 0
    series = [s[~closing] for s in series]
    n_steps = series[0].shape[0]
    keep = np.unique(np.concatenate((np.arange(min(FULL_STEPS, n_steps)),
                                     np.arange(FULL_STEPS, n_steps, STRIDE), [n_steps - 1])))
    q = np.loadtxt(ex / "bottom_boundary_discharge.txt", skiprows=1)
    log = (ex / "diffswe_run.log").read_text()
    steps = re.findall(r"Step:\s+(\d+) \| Sim Clock Time:\s+([\d.]+)s \| Max Depth: ([\d.]+)m \| "
                       r"Min Depth: ([\d.]+)m", log)
    log_u = np.array([[int(a), int(round(float(t) * 100)), int(round(float(mx) * 1e6)),
                       int(round(float(mn) * 1e6))] for a, t, mx, mn in steps], dtype=np.int64)
    with hdf5.File(ex / "hmax_n_0012.nc") as f:
        h = f.dataset("hmax").read()
    return dict(rain=rain, gauge_xy=np.array(coords), discharge_cols=np.array(q_cols), n_steps=np.int64(n_steps),
                n_closing=np.int64(closing.sum()), rows=keep.astype(np.int64), time_dt_u=series[0][keep, :2],
                depth_u=np.stack([s[keep, 2] for s in series], axis=1),
                level_u=np.stack([s[keep, 3] for s in series], axis=1), discharge=q, log_u=log_u,
                hmax_sub_n=np.round(h[::2, ::2] * 1e7).astype(np.int32),
                hmax_stats=np.array([h.sum(), h.max(), float((h > 1e-6).sum()), float(np.sum(h * h))]))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--examples", default="/root/reference/Examples/forward/Cea2008")
    root = Path(ap.parse_args().examples)
    topo = {(root / folder / "topo.asc").read_bytes() for folder, _, _ in CASES.values()}
    assert len(topo) == 1, "the cases no longer share one DEM"
    (HERE / "cea_topo.asc.gz").write_bytes(gzip.compress(topo.pop(), mtime=0))
    for case, (folder, rain_file, q_cols) in CASES.items():
        out = pack(root / folder, rain_file, q_cols)
        np.savez_compressed(HERE / f"cea_{case}.npz", **out)
        print(f"{case}: {int(out['n_steps'])} steps ({int(out['n_closing'])} closing rows dropped), "
              f"{out['rows'].size} gauge rows kept, {out['log_u'].shape[0]} log lines, "
              f"{out['discharge'].shape[0]} discharge rows")


if __name__ == "__main__":
    main()
