"""Hand-written adjoint on the GPU against oracle autograd, central FD and the
reference's own autograd gradient (golden vectors).

Bars: FP64 adjoint equals oracle autograd to 1e-10 relative; the spline
inversion gradient matches the reference's autograd gradient to 1e-4 relative.
"""
import numpy as np
import pytest
import torch

import diffswe2d_amd as swe
from diffswe2d_amd import inversion, scenarios
from oracle import swe_oracle as O

pytestmark = pytest.mark.gpu

STEP_CASES = [
    dict(bc=("wall", "wall", "wall", "wall"), friction=True, level=0.0),
    dict(bc=("water_level", "wall", "transmissive", "water_level"), friction=True, level=0.45),
    dict(bc=("periodic", "periodic", "absorbing", "constant"), friction=False, level=0.3),
    dict(bc=("transmissive", "absorbing", "periodic", "periodic"), friction=True, level=0.2),
    dict(bc=("absorbing", "water_level", "wall", "absorbing"), friction=True, level=0.35),
]


def relerr(a, b):
    a = torch.as_tensor(a).double().cpu()
    b = torch.as_tensor(b).double().cpu()
    return float(torch.linalg.norm(a - b) / max(float(torch.linalg.norm(b)), 1e-300))


def setup(golden, case, dtype, dev):
    d = golden("steps")
    c = STEP_CASES[case]
    cfg = swe.SWEConfig(boundary_left=c["bc"][0], boundary_right=c["bc"][1], boundary_top=c["bc"][2],
                        boundary_bottom=c["bc"][3], apply_friction=c["friction"], constant_value=0.05)
    bed = torch.tensor(d["bed"], dtype=dtype)
    rough = torch.tensor(d["rough"], dtype=dtype)
    m = swe.SWE2D(bed, rough, 0.1, 0.1, config=cfg, train_dem=False).to(dev)
    U = torch.tensor(d["U0"], dtype=dtype)
    return m, cfg, bed, rough, U, torch.tensor(c["level"], dtype=dtype)


@pytest.mark.parametrize("case", range(len(STEP_CASES)))
def test_operator_adjoint_fp64(golden, cuda_device, case):
    m, cfg, bed, rough, U, lvl = setup(golden, case, torch.float64, cuda_device)
    lam = torch.randn(U.shape, dtype=torch.float64, generator=torch.Generator().manual_seed(case))
    Ur, lr = U.clone().requires_grad_(True), lvl.clone().requires_grad_(True)
    gU_ref, gl_ref = torch.autograd.grad(O.spatial_operator(Ur, bed, cfg, 0.1, 0.1, lr), (Ur, lr), lam,
                                         allow_unused=True)
    Ug = U.to(cuda_device).requires_grad_(True)
    lg = lvl.to(cuda_device).requires_grad_(True)
    rhs = m.spatial_operator(Ug, boundary_level=lg)
    gU, gl = torch.autograd.grad(rhs, (Ug, lg), lam.to(cuda_device), allow_unused=True)
    e = relerr(gU, gU_ref)
    print(f"case {case}: operator adjoint rel err {e:.2e}")
    assert e <= 1e-10
    if gl_ref is not None:
        assert abs(float(gl) - float(gl_ref)) <= 1e-9 * max(1.0, abs(float(gl_ref)))


@pytest.fixture
def ieee_oracle():
    """The oracle with correctly rounded sqrt: its forward values match the GPU's bit for bit."""
    old = O.SQRT_MODE
    O.SQRT_MODE = "ieee"
    yield O
    O.SQRT_MODE = old


@pytest.mark.parametrize("case", range(len(STEP_CASES)))
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_step_adjoint(golden, cuda_device, ieee_oracle, case, dtype):
    """Against oracle autograd in the same precision (FP32 backward rounding allowed 1e-4)."""
    m, cfg, bed, rough, U, lvl = setup(golden, case, dtype, cuda_device)
    lam = torch.randn(U.shape, dtype=dtype, generator=torch.Generator().manual_seed(10 + case))
    Ur, lr = U.clone().requires_grad_(True), lvl.clone().requires_grad_(True)
    out, dt = O.step(Ur, bed, rough, cfg, 0.1, 0.1, 1e30, lr)
    gU_ref, gl_ref = torch.autograd.grad(out, (Ur, lr), lam, allow_unused=True)
    Ug = U.to(cuda_device).requires_grad_(True)
    lg = lvl.to(cuda_device).requires_grad_(True)
    outg, dtg = m.step(Ug, 0.0, 1e30, boundary_level=lg)
    gU, gl = torch.autograd.grad(outg, (Ug, lg), lam.to(cuda_device), allow_unused=True)
    e = relerr(gU, gU_ref)
    tol = 1e-10 if dtype == torch.float64 else 1e-4
    print(f"case {case} {dtype}: step adjoint rel err vs oracle autograd {e:.2e}")
    assert e <= tol
    if gl_ref is not None and gl is not None:
        assert abs(float(gl) - float(gl_ref)) <= tol * max(1.0, abs(float(gl_ref))) * 10


@pytest.mark.parametrize("bc", [("water_level", "wall", "periodic", "periodic"),
                                ("transmissive", "absorbing", "wall", "water_level")])
def test_step_adjoint_long_sides_fp64(cuda_device, ieee_oracle, bc):
    """A grid whose long side exceeds one block's ghost fold (2100 > 2048): the four
    per-side launches and the partial-sum level reduction against oracle autograd."""
    ny, nx = 2100, 12
    g = torch.Generator().manual_seed(5)
    yy = torch.arange(ny, dtype=torch.float64)[:, None]
    xx = torch.arange(nx, dtype=torch.float64)[None, :]
    bed = (-0.5 + 0.01 * xx + 0.05 * torch.sin(0.01 * yy)).expand(ny, nx).contiguous()[None, None]
    rough = torch.full((1, 1, ny, nx), 0.02, dtype=torch.float64)
    h = (0.3 - bed[0, 0]).clamp_min(0) + 0.01 * torch.rand(ny, nx, generator=g, dtype=torch.float64)
    U = torch.stack([h, 0.02 * torch.rand(ny, nx, generator=g, dtype=torch.float64) * h,
                     0.02 * torch.rand(ny, nx, generator=g, dtype=torch.float64) * h])[None]
    cfg = swe.SWEConfig(boundary_left=bc[0], boundary_right=bc[1], boundary_top=bc[2], boundary_bottom=bc[3])
    m = swe.SWE2D(bed, rough, 0.1, 0.1, config=cfg, train_dem=False).to(cuda_device)
    lvl = torch.tensor(0.35, dtype=torch.float64)
    lam = torch.randn(U.shape, dtype=torch.float64, generator=g)
    Ur, lr = U.clone().requires_grad_(True), lvl.clone().requires_grad_(True)
    out, _ = O.step(Ur, bed, rough, cfg, 0.1, 0.1, 1e30, lr)
    gU_ref, gl_ref = torch.autograd.grad(out, (Ur, lr), lam)
    Ug, lg = U.to(cuda_device).requires_grad_(True), lvl.to(cuda_device).requires_grad_(True)
    outg, _ = m.step(Ug, 0.0, 1e30, boundary_level=lg)
    gU, gl = torch.autograd.grad(outg, (Ug, lg), lam.to(cuda_device))
    e = relerr(gU, gU_ref)
    el = abs(float(gl) - float(gl_ref)) / max(abs(float(gl_ref)), 1e-300)
    print(f"{bc}: step adjoint rel err {e:.2e}, level grad {float(gl):.6e} vs {float(gl_ref):.6e}")
    assert e <= 1e-10 and el <= 1e-9


@pytest.mark.parametrize("mode", ["replay", "verify", "tape", "mixed"])
def test_forward_adjoint_with_checkpoints_fp64(cuda_device, ieee_oracle, mode):
    """A 0.3 s forward call (many steps, K = 5) against oracle autograd through the same steps.

    The initial surface carries a smooth ripple: a lake exactly at rest has
    bitwise ties (eta_L == eta_R, zero slopes) where the derivative is a
    subgradient chosen by last-bit rounding; the single-step tests cover the
    kernels themselves.
    """
    sc = scenarios.coastal_runup(nx=40, ny=26, dx=0.136)
    cfg = swe.SWEConfig(boundary_left="water_level", boundary_right="wall", boundary_top="wall",
                        boundary_bottom="wall")
    bed = torch.tensor(sc["bed"])
    rough = torch.tensor(sc["rough"])
    m = swe.SWE2D(bed, rough, 0.136, 0.136, config=cfg, train_dem=False).to(cuda_device)
    m.adjoint_checkpoint_every = 5
    # replay: checkpoints only, one library call per segment; verify: per-step replay
    # with dt read-back; tape: the forward records every segment's tape (no replay);
    # mixed: the tape budget covers the first two segments only
    m.adjoint_verify_replay = mode == "verify"
    state_bytes = 3 * 26 * 40 * 8
    m.adjoint_tape = {"replay": False, "verify": False, "tape": True, "mixed": 2 * 16 * state_bytes}[mode]
    U0 = torch.tensor(sc["U"])
    yy, xx = torch.meshgrid(torch.arange(26, dtype=torch.float64), torch.arange(40, dtype=torch.float64),
                            indexing="ij")
    ripple = 2e-3 * torch.sin(0.7 * xx + 0.3) * torch.cos(0.45 * yy + 0.2)
    wet = U0[0, 0] > 0
    U0[0, 0] = torch.where(wet, U0[0, 0] + ripple.abs() + 1e-3, U0[0, 0])
    U0[0, 1] = torch.where(wet, 1e-3 * torch.cos(0.3 * yy) * U0[0, 0], U0[0, 1])
    lvl = torch.tensor(0.015, dtype=torch.float64)
    lam = torch.randn(U0.shape, dtype=torch.float64, generator=torch.Generator().manual_seed(3))
    Ur, lr = U0.clone().requires_grad_(True), lvl.clone().requires_grad_(True)
    out, _ = O.forward(Ur, bed, rough, cfg, 0.136, 0.136, 0.3, boundary_level=lr)
    gU_ref, gl_ref = torch.autograd.grad(out, (Ur, lr), lam)
    Ug = U0.to(cuda_device).requires_grad_(True)
    lg = lvl.to(cuda_device).requires_grad_(True)
    outg, _ = m(Ug, t_end=0.3, boundary_level=lg)
    assert relerr(outg.detach(), out.detach()) < 1e-12
    nsteps = len(m.last_dt_series)
    kinds = [kind for kind, _, _ in m.last_adjoint_segments]
    expect = {"replay": {"ckpt"}, "verify": {"ckpt"}, "tape": {"tape"}, "mixed": {"tape", "ckpt"}}[mode]
    assert set(kinds) == expect, kinds
    assert sum(n for _, _, n in m.last_adjoint_segments) == nsteps
    gU, gl = torch.autograd.grad(outg, (Ug, lg), lam.to(cuda_device))
    e = relerr(gU, gU_ref)
    el = abs(float(gl) - float(gl_ref)) / abs(float(gl_ref))
    print(f"{nsteps} steps: forward adjoint rel err dU0 {e:.2e}, dlevel {el:.2e}")
    assert nsteps > 10
    assert e <= 1e-9 and el <= 1e-9


def spline_run(golden, dtype, dev):
    d = golden("gradient")
    cfg = swe.SWEConfig(boundary_left="water_level", boundary_right="wall", boundary_top="wall",
                        boundary_bottom="wall")
    m = swe.SWE2D(torch.tensor(d["bed"], dtype=dtype), torch.tensor(d["rough"], dtype=dtype), 0.08, 0.08,
                  config=cfg, train_dem=False).to(dev)
    U0 = torch.tensor(d["U0"], dtype=dtype, device=dev)
    gauges = [tuple(int(v) for v in g) for g in d["gauges"]]
    name = "f64" if dtype == torch.float64 else "f32"
    prob = inversion.GaugeProblem(m, U0, gauges, chunk=0.2, n_chunks=10)
    basis = torch.tensor(d[f"basis_{name}"], dtype=dtype, device=dev)
    target = torch.tensor(d[f"target_{name}"], dtype=dtype, device=dev)
    return d, prob, basis, target, name


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_spline_gradient_matches_reference_autograd(golden, cuda_device, dtype):
    """Config-4 gradient pin: 64 spline controls, 6 gauges, 10 x 0.2 s chunks on the coarse run-up grid."""
    d, prob, basis, target, name = spline_run(golden, dtype, cuda_device)
    ctrl = torch.tensor(d["controls"], dtype=dtype, device=cuda_device, requires_grad=True)
    obs = prob.simulate(basis @ ctrl)
    loss = prob.misfit(obs, target)
    loss.backward()
    loss = loss.detach()
    g = ctrl.grad.double().cpu()
    ref = torch.tensor(d[f"grad_{name}"])
    ref64 = torch.tensor(d["grad_f64"])
    e = relerr(g, ref)
    e64 = relerr(g, ref64)
    print(f"{name}: loss {float(loss):.6e} vs ref {float(d[f'loss_{name}']):.6e}; grad rel err vs reference "
          f"autograd {e:.2e}, vs FP64 reference {e64:.2e} (ref FP32 vs FP64 {relerr(d['grad_f32'], ref64):.2e})")
    assert abs(float(loss) - float(d[f"loss_{name}"])) <= 1e-4 * float(d[f"loss_{name}"])
    assert e <= 1e-4


def oracle_gauge_loss(d, basis, target, controls, dts_per_chunk):
    """Oracle (FP64, IEEE sqrt) gauge loss with every step's dt frozen to the given series."""
    cfg = swe.SWEConfig(boundary_left="water_level", boundary_right="wall", boundary_top="wall",
                        boundary_bottom="wall")
    bed, rough = torch.tensor(d["bed"]), torch.tensor(d["rough"])
    U = torch.tensor(d["U0"])
    series = basis @ controls
    obs = []
    for k, dts in enumerate(dts_per_chunk):
        for dt in dts:
            U, _ = O.step(U, bed, rough, cfg, 0.08, 0.08, 1e30, series[k], fixed_dt=dt)
        obs.append(torch.stack([U[0, 0, iy, ix] + bed[0, 0, iy, ix] for iy, ix in d["gauges"]]))
    return float(((torch.stack(obs) - target) ** 2).mean(dim=0).mean())


def test_spline_gradient_central_differences(golden, cuda_device, ieee_oracle):
    """Adjoint gradient vs central differences of the dt-frozen map on 8 seeded controls.

    The reference detaches dt (solver.py:625-626), so its gradient -- and the
    adjoint's -- is the derivative with the dt series held fixed; the finite
    differences therefore freeze each step's dt at the unperturbed run's value.
    """
    d, prob, basis, target, _ = spline_run(golden, torch.float64, cuda_device)
    ctrl0 = torch.tensor(d["controls"], dtype=torch.float64, device=cuda_device)
    ctrl = ctrl0.clone().requires_grad_(True)
    dts_per_chunk = []
    s = prob.U0
    series = basis @ ctrl
    obs = []
    for k in range(prob.n_chunks):
        s, _ = prob.model(s, t_end=prob.chunk, start_time=k * prob.chunk, boundary_level=series[k])
        dts_per_chunk.append(list(prob.model.last_dt_series))
        obs.append(s[0, 0, prob._iy, prob._ix] + prob._bed_g)
    loss = prob.misfit(torch.stack(obs), target)
    loss.backward()
    # 64 controls over 10 samples: only the controls whose basis functions reach a sample move the loss
    active = np.flatnonzero(basis.abs().sum(0).cpu().numpy() > 0)
    rng = np.random.default_rng(4)
    idx = rng.choice(active, size=min(8, len(active)), replace=False)
    h = 1e-6
    basis_c, target_c, c0 = basis.cpu(), target.cpu(), ctrl0.cpu()
    for i in idx:
        cp, cm = c0.clone(), c0.clone()
        cp[i] += h
        cm[i] -= h
        fd = (oracle_gauge_loss(d, basis_c, target_c, cp, dts_per_chunk)
              - oracle_gauge_loss(d, basis_c, target_c, cm, dts_per_chunk)) / (2 * h)
        ad = float(ctrl.grad[i])
        print(f"control {i}: adjoint {ad:.8e}, central FD (dt frozen) {fd:.8e}, rel {abs(ad - fd) / max(abs(fd), 1e-300):.2e}")
        assert abs(ad - fd) <= 1e-4 * max(abs(fd), 1e-3 * float(ctrl.grad.abs().max()))
