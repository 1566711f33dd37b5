"""Pack the reference's CEA 2008 C1 forward example outputs as a test fixture.

Run in the development container only (it reads the reference tree):

    python tests/golden/make_cea_fixture.py [--example /root/reference/Examples/forward/Cea2008/Cea2008_C1_new]

Inputs and outputs of that example (Examples/forward/Cea2008/Cea2008_C1_new,
driver example_run.py): a 208 x 256 ESRI-ASCII DEM (topo.asc), a constant
317 mm/h rainfall for 45 s (cstRain-test1.txt), and what the FP64 run on an
A100 wrote: per-step gauge records at four cells (gauge_*.txt: time, dt,
depth, h + z, printed %.6f), the bottom-boundary discharge every 0.1 s
(bottom_boundary_discharge.txt, %.8f) and the domain maximum depth every
0.1 s (diffswe_run.log, %.6f).  Only data files are read; nothing is run.

Written next to this script:
  cea_c1_topo.asc.gz  the DEM, byte for byte, gzip-compressed;
  cea_c1.npz          rainfall table, gauge coordinates, the gauge records of
                      the first 2000 steps and every 25th after that (as
                      integers in units of 1e-6, exactly what was printed),
                      the step count, the discharge and log series.

The example's run closes 759 of its 1200 chunks of 0.1 s with one extra step
whose dt prints as 0.000000 (under 5e-7 s, at the chunk's end time).  The
reference's eager FP64 loop (src/solver.py:738-768) reaches each chunk end
exactly and takes no such step; that run was made through torch.compile.  The
fixture drops those rows (their count is kept), so the remaining records
compare step for step with an eager-semantics run.
"""
from __future__ import annotations

import argparse
import gzip
import re
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
FULL_STEPS = 2000
STRIDE = 25


def micro(text_rows):
    """%.6f columns -> int64 micro-units, exact."""
    return np.array([[int(round(float(v) * 1e6)) for v in row] for row in text_rows], dtype=np.int64)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--example", default="/root/reference/Examples/forward/Cea2008/Cea2008_C1_new")
    ex = Path(ap.parse_args().example)

    (HERE / "cea_c1_topo.asc.gz").write_bytes(gzip.compress((ex / "topo.asc").read_bytes(), mtime=0))

    rain = np.loadtxt(ex / "cstRain-test1.txt")
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
    time_dt = series[0][keep, :2]
    depth = np.stack([s[keep, 2] for s in series], axis=1)
    level = np.stack([s[keep, 3] for s in series], axis=1)

    q = np.loadtxt(ex / "bottom_boundary_discharge.txt", skiprows=1)
    log = (ex / "diffswe_run.log").read_text()
    steps = re.findall(r"Step:\s+(\d+) \| Sim Clock Time:\s+([\d.]+)s \| Max Depth: ([\d.]+)m \| "
                       r"Min Depth: ([\d.]+)m", log)
    log_max = np.array([[int(a), int(round(float(t) * 100)), int(round(float(mx) * 1e6)),
                         int(round(float(mn) * 1e6))] for a, t, mx, mn in steps], dtype=np.int64)

    np.savez_compressed(HERE / "cea_c1.npz", rain=rain, gauge_xy=np.array(coords), n_steps=np.int64(n_steps),
                        n_closing=np.int64(closing.sum()), rows=keep.astype(np.int64), time_dt_u=time_dt,
                        depth_u=depth, level_u=level, discharge=q, log_u=log_max)
    print(f"{n_steps} steps ({int(closing.sum())} closing rows dropped), {keep.size} gauge rows kept, "
          f"{len(steps)} log lines, {q.shape[0]} discharge rows")


if __name__ == "__main__":
    main()
