"""Roughness-length ("smart") friction on the GPU against the oracle.

Any ``frictionmodel`` other than "manning" selects the BG_Flood roughness-length
coefficient (reference src/solver.py:654-661, src/friction.py:41-59, 69-99):
Cf from h / z0, shallow branch 1 / (0.46 h / z0) below e, deep branch
1 / (2.5 (ln r - 1 + 1.359 / r))^2 above.  The oracle's version of it is pinned
bit for bit to the reference's own ``apply_roughness_length_friction`` by the
``friction.npz`` golden vectors in the CPU suite (tests/test_oracle.py; the
vectors were written on the development host, whose vectorised FP64 log can
differ in the last bit from another host's).

Bars: steps FP32 rel-L2 <= 1e-6 / FP64 <= 1e-12 per step with identical masks
and dt; a device-driven forward run over a moving shoreline: FP32 rel-L2 <= 1e-5,
dt series 1e-6, identical masks; the step adjoint equals oracle autograd to
1e-10 (FP64) and 1e-4 (FP32).
"""
import numpy as np
import pytest
import torch

import diffswe2d_amd as swe
from diffswe2d_amd import scenarios
from oracle import swe_oracle as O

pytestmark = pytest.mark.gpu

DT = {"f32": torch.float32, "f64": torch.float64}
CASES = [
    ("wall", "wall", "wall", "wall"),
    ("water_level", "wall", "transmissive", "water_level"),
    ("transmissive", "absorbing", "periodic", "periodic"),
]
LEVELS = [0.0, 0.45, 0.2]
FRICTION = "roughness_length"


def relerr(a, b):
    a = torch.as_tensor(a, dtype=torch.float64).cpu()
    b = torch.as_tensor(b, dtype=torch.float64).cpu()
    return float(torch.linalg.norm(a - b) / max(float(torch.linalg.norm(b)), 1e-300))


def z0_field(shape, seed=21):
    """Roughness lengths from 0.1 mm to 0.2 m: h / z0 falls on both sides of e."""
    return np.random.default_rng(seed).uniform(1e-4, 0.2, shape)


@pytest.fixture
def ieee_oracle():
    old = O.SQRT_MODE
    O.SQRT_MODE = "ieee"
    yield O
    O.SQRT_MODE = old


def make(d, case, dtype, dev):
    bc = CASES[case]
    cfg = swe.SWEConfig(boundary_left=bc[0], boundary_right=bc[1], boundary_top=bc[2], boundary_bottom=bc[3],
                        apply_friction=True, frictionmodel=FRICTION, constant_value=0.05)
    bed = torch.tensor(d["bed"], dtype=dtype)
    z0 = torch.tensor(z0_field(d["bed"].shape), dtype=dtype)
    m = swe.SWE2D(bed, z0, 0.1, 0.1, config=cfg, train_dem=False).to(dev)
    return m, cfg, bed, z0


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("name", ["f32", "f64"])
def test_steps_with_roughness_length_friction(golden, cuda_device, case, name):
    d = golden("steps")
    dtype = DT[name]
    m, cfg, bed, z0 = make(d, case, dtype, cuda_device)
    lvl = torch.tensor(LEVELS[case], dtype=dtype)
    U = torch.tensor(d["U0"], dtype=dtype)
    Ug = U.to(cuda_device)
    tol = 1e-6 if name == "f32" else 1e-12
    worst = 0.0
    for k in range(12):
        U, dt_ref = O.step(U, bed, z0, cfg, 0.1, 0.1, 1e30, lvl)
        Ug, dt = m.step(Ug, 0.0, 1e30, boundary_level=lvl.to(cuda_device))
        e = relerr(Ug, U)
        worst = max(worst, e)
        assert e <= tol, f"step {k}: rel-L2 {e:.3e}"
        assert abs(dt - dt_ref) <= tol * dt_ref
        assert torch.equal((Ug[:, 0] > cfg.dry_depth).cpu(), U[:, 0] > cfg.dry_depth)
    print(f"case {case} {name}: 12 steps with roughness-length friction, worst rel-L2 {worst:.3e}")


@pytest.mark.parametrize("name", ["f32", "f64"])
def test_forward_run_with_roughness_length_friction(cuda_device, ieee_oracle, name):
    """Reduced coastal run-up (68 x 43, dx 0.08) with a roughness-length field: 2 s, moving shoreline."""
    dtype = DT[name]
    sc = scenarios.coastal_runup(nx=68, ny=43, dx=0.08)
    cfg = swe.SWEConfig(boundary_left="water_level", boundary_right="wall", boundary_top="wall",
                        boundary_bottom="wall", frictionmodel=FRICTION)
    bed = torch.tensor(sc["bed"], dtype=dtype)
    z0 = torch.tensor(z0_field(sc["bed"].shape, seed=22) * 0.05, dtype=dtype)  # 5 um .. 1 cm
    m = swe.SWE2D(bed, z0, 0.08, 0.08, config=cfg, train_dem=False).to(cuda_device)
    U0 = torch.tensor(sc["U"], dtype=dtype)
    level = 0.02
    dts = []
    with torch.no_grad():
        Uo, _ = O.forward(U0, bed, z0, cfg, 0.08, 0.08, 2.0, boundary_level=level, dt_log=dts)
        Ug, _ = m(U0.to(cuda_device), t_end=2.0, boundary_level=level)
    dts_g = m.last_dt_series.cpu().double().numpy()
    dts = np.array(dts)
    assert len(dts_g) == len(dts)
    dt_err = float(np.max(np.abs(dts_g - dts) / dts))
    e = relerr(Ug, Uo)
    mism = int(((Ug[:, 0] > cfg.dry_depth).cpu() != (Uo[:, 0] > cfg.dry_depth)).sum())
    flooded = int(((U0[0, 0] == 0) & (Uo[0, 0] > cfg.dry_depth)).sum())
    print(f"{name}: {len(dts)} steps, rel-L2 vs IEEE-sqrt oracle {e:.3e}, dt max rel {dt_err:.3e}, "
          f"mask mismatches {mism}, cells wetted {flooded}")
    assert flooded > 0
    assert e <= (1e-5 if name == "f32" else 1e-12)
    assert dt_err <= (1e-6 if name == "f32" else 1e-12)
    assert mism == 0


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_step_adjoint_with_roughness_length_friction(golden, cuda_device, ieee_oracle, case, dtype):
    d = golden("steps")
    m, cfg, bed, z0 = make(d, case, dtype, cuda_device)
    lvl = torch.tensor(LEVELS[case], dtype=dtype)
    U = torch.tensor(d["U0"], dtype=dtype)
    lam = torch.randn(U.shape, dtype=dtype, generator=torch.Generator().manual_seed(30 + case))
    Ur, lr = U.clone().requires_grad_(True), lvl.clone().requires_grad_(True)
    out, _ = O.step(Ur, bed, z0, cfg, 0.1, 0.1, 1e30, lr)
    gU_ref, gl_ref = torch.autograd.grad(out, (Ur, lr), lam, allow_unused=True)
    Ug = U.to(cuda_device).requires_grad_(True)
    lg = lvl.to(cuda_device).requires_grad_(True)
    outg, _ = m.step(Ug, 0.0, 1e30, boundary_level=lg)
    gU, gl = torch.autograd.grad(outg, (Ug, lg), lam.to(cuda_device), allow_unused=True)
    e = relerr(gU, gU_ref)
    tol = 1e-10 if dtype == torch.float64 else 1e-4
    print(f"case {case} {dtype}: step adjoint with roughness-length friction, rel err {e:.2e}")
    assert e <= tol
    if gl_ref is not None and gl is not None:
        assert abs(float(gl) - float(gl_ref)) <= tol * max(1.0, abs(float(gl_ref))) * 10
