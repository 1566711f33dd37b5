"""Write basin512_2000.npz: the 512 x 512 tsunami basin advanced 2000 steps by the oracle.

The basin is ``scenarios.tsunami_basin(n=512)``: the benchmark's generator
(fractal bathymetry 10-4000 m, ~15 % land on the east coast, a 2 m Gaussian
hump of sigma 20 km, Manning n = 0.025, transmissive sides) at 512 x 512 cells
of 100 m.  2000 steps at CFL 0.5 are the benchmark's full run length.

Three oracle runs (oracle/swe_oracle.py, the reference's step restated op for
op; src/solver.py:572-663):

* FP32 with correctly rounded square roots and friction powers (SQRT_MODE =
  "ieee"), the rounding the GPU kernels use: every dt, the wet/dry mask every
  200 steps, the final state;
* FP32 with torch's CPU square root (the reference's own CPU rounding): every
  dt and the final depth;
* FP64: every dt and the final depth.

Run from the repository root: ``python tests/golden/make_basin_fixture.py``
(about 10 minutes on 8 CPU threads).
"""
import sys
import time
from pathlib import Path

import numpy as np
import torch

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import diffswe2d_amd as swe  # noqa: E402
import oracle.swe_oracle as O  # noqa: E402
from diffswe2d_amd import scenarios  # noqa: E402

STEPS = 2000
MASK_EVERY = 200


def run(sc, dtype, sqrt_mode):
    cfg = swe.SWEConfig(boundary_left="transmissive", boundary_right="transmissive",
                        boundary_top="transmissive", boundary_bottom="transmissive")
    bed = torch.tensor(sc["bed"], dtype=dtype)
    rough = torch.tensor(sc["rough"], dtype=dtype)
    U = torch.tensor(sc["U"], dtype=dtype)
    old = O.SQRT_MODE
    O.SQRT_MODE = sqrt_mode
    dts, masks = [], []
    t0 = time.time()
    try:
        with torch.no_grad():
            for k in range(STEPS):
                U, dt = O.step(U, bed, rough, cfg, sc["dx"], sc["dy"], 1e30)
                dts.append(dt)
                if (k + 1) % MASK_EVERY == 0:
                    masks.append((U[0, 0] > cfg.dry_depth).numpy())
    finally:
        O.SQRT_MODE = old
    print(f"{dtype} sqrt={sqrt_mode}: {STEPS} steps in {time.time() - t0:.0f} s", flush=True)
    return np.array(dts), np.packbits(np.stack(masks)), U.numpy()


def main():
    sc = scenarios.tsunami_basin(n=512)
    dt_ieee, mask_ieee, U_ieee = run(sc, torch.float32, "ieee")
    dt_ref, _, U_ref = run(sc, torch.float32, O.SQRT_MODE)
    dt_64, _, U_64 = run(sc, torch.float64, O.SQRT_MODE)
    out = ROOT / "tests" / "golden" / "basin512_2000.npz"
    np.savez_compressed(out, steps=STEPS, mask_every=MASK_EVERY, dt_ieee=dt_ieee.astype(np.float32),
                        mask_ieee=mask_ieee, U_ieee=U_ieee.astype(np.float32),
                        dt_ref=dt_ref.astype(np.float32), h_ref=U_ref[0, 0].astype(np.float32),
                        dt_f64=dt_64, h_f64=U_64[0, 0])
    print(f"wrote {out} ({out.stat().st_size / 1e6:.1f} MB)")


if __name__ == "__main__":
    main()
