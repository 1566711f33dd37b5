"""Rainfall record and infiltration in the step kernels against the oracle.

The oracle's step takes the two rain fields explicitly; the reference samples
its RainfallForcing at the step's model time (stage 1) and at time + dt/2
(stage 2) (solver.py:609-641), which the test does with the reference-shaped
``RainfallForcing.at`` on the CPU.
"""
import pytest
import torch

import diffswe2d_amd as swe
from oracle import swe_oracle as O

pytestmark = pytest.mark.gpu


def relerr(a, b):
    a, b = torch.as_tensor(a).double().cpu(), torch.as_tensor(b).double().cpu()
    return float(torch.linalg.norm(a - b) / max(float(torch.linalg.norm(b)), 1e-300))


def setup(dtype, dev, batch=2, ny=40, nx=56):
    g = torch.Generator().manual_seed(11)
    yy = torch.arange(ny, dtype=torch.float64)[:, None]
    xx = torch.arange(nx, dtype=torch.float64)[None, :]
    bed = (-0.4 + 0.01 * xx + 0.03 * torch.sin(0.2 * yy)).expand(ny, nx)[None, None].contiguous()
    rough = torch.full((1, 1, ny, nx), 0.02, dtype=torch.float64)
    h = (0.1 - bed[0, 0]).clamp_min(0) + 0.02 * torch.rand(batch, ny, nx, generator=g, dtype=torch.float64)
    U = torch.stack([h, 0.05 * h * torch.rand(batch, ny, nx, generator=g, dtype=torch.float64),
                     -0.03 * h * torch.rand(batch, ny, nx, generator=g, dtype=torch.float64)], 1)
    times = torch.tensor([0.0, 0.3, 0.5, 1.2], dtype=torch.float64)
    rates = 1e-3 * torch.rand(4, 1, ny, nx, generator=g, dtype=torch.float64)
    infil = 4e-4 * torch.rand(batch, 1, ny, nx, generator=g, dtype=torch.float64) - 1e-4  # some < 0 (clamped)
    cfg = swe.SWEConfig(boundary_left="water_level", boundary_right="wall", boundary_top="wall",
                        boundary_bottom="transmissive")
    rf = swe.RainfallForcing(times.to(dtype), rates.to(dtype))  # CPU copy for the oracle
    m = swe.SWE2D(bed.to(dtype), rough.to(dtype), 0.1, 0.1, config=cfg, train_dem=False,
                  rainfall=swe.RainfallForcing(times.to(dtype), rates.to(dtype))).to(dev)
    return m, cfg, bed.to(dtype), rough.to(dtype), U.to(dtype), rf, infil.to(dtype)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("t", [0.29, 0.5, 1.19, 3.0])
def test_step_with_rain_and_infiltration(cuda_device, dtype, t):
    m, cfg, bed, rough, U, rf, infil = setup(dtype, cuda_device)
    tt = torch.tensor(t, dtype=dtype)
    level = torch.tensor(0.12, dtype=dtype)
    with torch.no_grad():
        _, dt_ref = O.step(U, bed, rough, cfg, 0.1, 0.1, 1e30, level, infiltration=infil)
        tf = float(tt)
        out_ref, dt2 = O.step(U, bed, rough, cfg, 0.1, 0.1, 1e30, level, rain_start=rf.at(tf),
                              rain_half=rf.at(tf + 0.5 * float(dt_ref)), infiltration=infil)
        out, dt = m.step(U.to(cuda_device), tt.to(cuda_device), 1e30, infiltration=infil.to(cuda_device),
                         boundary_level=level.to(cuda_device))
    e = relerr(out, out_ref)
    print(f"{dtype} t={t}: dt {dt:.9e} vs {float(dt2):.9e}; rel err {e:.2e}")
    assert abs(dt - float(dt2)) <= 1e-6 * float(dt2)
    assert e <= (1e-6 if dtype == torch.float32 else 1e-13)


def test_forward_with_rain_device_loop_equals_host_loop(cuda_device):
    m, cfg, bed, rough, U, rf, infil = setup(torch.float32, cuda_device)
    Ud = U.to(cuda_device)
    inf = infil.to(cuda_device)
    with torch.no_grad():
        a, ha = m(Ud, t_end=0.9, start_time=0.1, infiltration=inf, boundary_level=0.12)
        b, hb = m(Ud, t_end=0.9, start_time=0.1, infiltration=inf, boundary_level=0.12, callback=lambda *x: None)
    assert torch.equal(a, b) and torch.equal(ha, hb)
    # the rain made a difference
    m2 = swe.SWE2D(bed, rough, 0.1, 0.1, config=cfg, train_dem=False).to(cuda_device)
    with torch.no_grad():
        c, _ = m2(Ud, t_end=0.9, start_time=0.1, infiltration=inf, boundary_level=0.12)
    assert float((a[:, 0] - c[:, 0]).abs().max()) > 1e-5


def test_step_adjoint_with_rain_fp64(cuda_device):
    m, cfg, bed, rough, U, rf, infil = setup(torch.float64, cuda_device, batch=1)
    lam = torch.randn(U.shape, dtype=torch.float64, generator=torch.Generator().manual_seed(2))
    level = torch.tensor(0.12, dtype=torch.float64)
    t = 0.4
    with torch.no_grad():
        _, dt_ref = O.step(U, bed, rough, cfg, 0.1, 0.1, 1e30, level, infiltration=infil)
    Ur, lr = U.clone().requires_grad_(True), level.clone().requires_grad_(True)
    out, _ = O.step(Ur, bed, rough, cfg, 0.1, 0.1, 1e30, lr, rain_start=rf.at(t),
                    rain_half=rf.at(t + 0.5 * float(dt_ref)), infiltration=infil)
    gU_ref, gl_ref = torch.autograd.grad(out, (Ur, lr), lam)
    Ug, lg = U.to(cuda_device).requires_grad_(True), level.to(cuda_device).requires_grad_(True)
    outg, _ = m.step(Ug, torch.tensor(t, dtype=torch.float64), 1e30, infiltration=infil.to(cuda_device),
                     boundary_level=lg)
    assert relerr(outg.detach(), out.detach()) <= 1e-13
    gU, gl = torch.autograd.grad(outg, (Ug, lg), lam.to(cuda_device))
    e = relerr(gU, gU_ref)
    print(f"adjoint with rain: rel err {e:.2e}; level {float(gl):.6e} vs {float(gl_ref):.6e}")
    assert e <= 1e-10
    assert abs(float(gl) - float(gl_ref)) <= 1e-9 * max(1.0, abs(float(gl_ref)))


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_operator_with_per_member_sources(cuda_device, dtype):
    """spatial_operator with rain and infiltration fields that differ between members: the call
    without autograd gives the values of the call with autograd and of the oracle
    (src/solver.py:520-525 adds them to the depth tendency per member)."""
    m, cfg, bed, rough, U, rf, infil = setup(dtype, cuda_device)
    rain = 1e-3 * torch.rand(U.shape[0], 1, *U.shape[-2:], dtype=dtype, generator=torch.Generator().manual_seed(4))
    level = torch.tensor(0.12, dtype=dtype)
    with torch.no_grad():
        rhs = m.spatial_operator(U.to(cuda_device), rainfall_rate=rain.to(cuda_device),
                                 infiltration=infil.to(cuda_device), boundary_level=level.to(cuda_device))
        old, O.SQRT_MODE = O.SQRT_MODE, "ieee"  # correctly rounded square roots, as the GPU's
        try:
            ref = O.spatial_operator(U, bed, cfg, 0.1, 0.1, level, rain=rain, infiltration=infil)
        finally:
            O.SQRT_MODE = old
    Ug = U.to(cuda_device).requires_grad_(True)
    rhs_g = m.spatial_operator(Ug, rainfall_rate=rain.to(cuda_device), infiltration=infil.to(cuda_device),
                               boundary_level=level.to(cuda_device))
    assert torch.equal(rhs, rhs_g.detach())
    e = relerr(rhs, ref)
    print(f"{dtype}: per-member sources, rel err vs oracle {e:.2e}")
    assert e <= (1e-6 if dtype == torch.float32 else 1e-13)


def test_step_refuses_infiltration_gradient(cuda_device):
    """step / forward do not differentiate the infiltration field: asking for it raises."""
    m, cfg, bed, rough, U, rf, infil = setup(torch.float64, cuda_device)
    i = infil.to(cuda_device).requires_grad_(True)
    with pytest.raises(NotImplementedError, match="infiltration"):
        m.step(U.to(cuda_device), 0.0, 1e30, infiltration=i, boundary_level=0.12)
    with pytest.raises(NotImplementedError, match="infiltration"):
        m(U.to(cuda_device), t_end=0.1, infiltration=i, boundary_level=0.12)
    with torch.no_grad():
        m.step(U.to(cuda_device), 0.0, 1e30, infiltration=i, boundary_level=0.12)


def test_gauge_problem_with_rainfall_runs_chunk_by_chunk(cuda_device):
    """An inversion problem on a model with a rainfall record takes the per-chunk loop (the fused
    chunked run does not carry the record) and observes the same values as explicit forward calls."""
    from diffswe2d_amd import inversion

    m, cfg, bed, rough, U, rf, infil = setup(torch.float64, cuda_device, batch=1)
    U0 = U.to(cuda_device)
    prob = inversion.GaugeProblem(m, U0, [(5, 7), (20, 30)], chunk=0.1, n_chunks=3)
    series = torch.tensor([0.12, 0.13, 0.11], dtype=torch.float64, device=cuda_device)
    with torch.no_grad():
        obs = prob.simulate(series)
        s, ref = U0, []
        for k in range(3):
            s, _ = m(s, t_end=0.1, start_time=0.1 * k, boundary_level=series[k])
            ref.append(torch.stack([s[0, 0, 5, 7] + m.bed[0, 0, 5, 7], s[0, 0, 20, 30] + m.bed[0, 0, 20, 30]]))
    assert torch.equal(obs, torch.stack(ref))
