"""The tolerance-bounded FP32 arithmetic (``SWE2D.arithmetic = "fast"``, the
library's SWE_ARITH_FAST) against the reference at the parity bars of the
reference-rounding path: FP32 state rel-L2 <= 1e-5 after the run, identical
wet/dry masks, dt series within 1e-6 relative (CFL-limited steps).

The fast path changes only roundings (division from the reciprocal estimate with
one correction, hardware square root, h^(1/3) from log2 / exp2, face velocities
without re-division), so each workload is run twice on the GPU -- reference
rounding and fast -- and both are measured against the same reference run.
"""
import numpy as np
import pytest
import torch

import diffswe2d_amd as swe
from diffswe2d_amd import scenarios

pytestmark = pytest.mark.gpu


def relerr(a, b):
    a = torch.as_tensor(a, dtype=torch.float64).cpu()
    b = torch.as_tensor(b, dtype=torch.float64).cpu()
    return float(torch.linalg.norm(a - b) / max(float(torch.linalg.norm(b)), 1e-300))


def test_fast_dam_break(golden, cuda_device):
    d = golden("dam_break")
    sc = scenarios.dam_break()
    cfg = swe.SWEConfig(boundary_left="wall", boundary_right="wall", boundary_top="wall", boundary_bottom="wall")
    out = {}
    for mode in ("reference", "fast"):
        m = swe.SWE2D(torch.tensor(sc["bed"], dtype=torch.float32), torch.tensor(sc["rough"], dtype=torch.float32),
                      sc["dx"], sc["dy"], config=cfg, train_dem=False).to(cuda_device)
        m.arithmetic = mode
        U = torch.tensor(sc["U"], dtype=torch.float32, device=cuda_device)
        dts = []
        with torch.no_grad():
            for _ in range(500):
                U, dt = m.step(U, 0.0, 1e30)
                dts.append(dt)
        out[mode] = (U, np.array(dts))
    U, dts = out["fast"]
    e = relerr(U, d["U500_f32"])
    e64 = relerr(U, d["U500_f64"])
    dt_err = float(np.max(np.abs(dts - d["dt_f32"]) / d["dt_f32"]))
    e_ref = relerr(out["reference"][0], d["U500_f32"])
    print(f"dam break 500 steps, fast arithmetic: rel-L2 vs reference FP32 {e:.3e} (reference rounding: "
          f"{e_ref:.3e}), vs FP64 {e64:.3e}; dt max rel {dt_err:.3e}")
    assert e <= 1e-5
    assert dt_err <= 1e-6
    h = U[0, 0].cpu()
    assert torch.equal(h, h.T)  # x and y still run the same arithmetic


def test_fast_full_runup(golden, cuda_device):
    from test_runup_full_gpu import capped_steps, run

    d = golden("runup_full_f32")
    ie = golden("runup_full_ieee_f32")
    mask_every = int(d["mask_every"])
    res = {mode: run(torch.float32, cuda_device, d["levels"], d["gauges"], mask_every, arithmetic=mode)
           for mode in ("reference", "fast")}
    cfl = ~capped_steps(d["dt"], d["chunk_steps"], 0.05)
    lines, rows = [], []
    for mode, (U, dts, gauge, masks) in res.items():
        for name, ref in (("reference", d), ("IEEE oracle", ie)):
            ref_masks = np.unpackbits(ref["mask"])[: np.stack(masks).size].reshape(np.stack(masks).shape).astype(bool)
            mism = int(sum(int((a != b).sum()) for a, b in zip(masks, ref_masks)))
            same_steps = len(dts) == len(ref["dt"])
            dt_err = float(np.max(np.abs(dts - ref["dt"])[cfl] / ref["dt"][cfl])) if same_steps else float("inf")
            e = relerr(U, ref["U"])
            lines.append(f"{mode} rounding vs {name}: steps {len(dts)}/{len(ref['dt'])}, rel-L2 {e:.3e}, "
                         f"CFL dt max rel {dt_err:.3e}, mask mismatches {mism}")
            rows.append((mode, name, same_steps, e, dt_err, mism))
    print("run-up 544x344, 450 chunks:\n  " + "\n  ".join(lines), flush=True)
    fast_vs_ref = [r for r in rows if r[0] == "fast" and r[1] == "reference"][0]
    assert fast_vs_ref[2]
    assert fast_vs_ref[3] <= 1e-5
