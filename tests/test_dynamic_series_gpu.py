"""A dynamic boundary series (SWE2D.set_dynamic_boundary; levels interpolated at every
step's start time, src/solver.py:391-415) evaluated on the device.

* Without gradients the device-driven run equals the per-step host loop (forced
  by a callback, the reference's loop structure) bit for bit: state, hmax, every
  dt and the gauge records; FP32 and FP64; the series starts after the run
  starts and ends before it ends (both ends extrapolate), and a one-sample
  series.
* With gradients the device-driven run (taped, and checkpointed with replay)
  gives the per-step loop's gradients with respect to U0, the series' water
  levels and a trainable bed.
* A backward through the device run's hmax raises instead of dropping that part
  of the gradient; through the per-step loop (``differentiable_hmax``) the hmax
  gradient equals the oracle's autograd.
"""
import pytest
import torch

import diffswe2d_amd as swe
from diffswe2d_amd import scenarios

pytestmark = pytest.mark.gpu


def relerr(a, b):
    a = torch.as_tensor(a, dtype=torch.float64).cpu()
    b = torch.as_tensor(b, dtype=torch.float64).cpu()
    return float(torch.linalg.norm(a - b) / max(float(torch.linalg.norm(b)), 1e-300))


def model(dev, dtype, train_dem=False):
    sc = scenarios.coastal_runup(nx=68, ny=43, dx=0.08)
    cfg = swe.SWEConfig(boundary_left="water_level", boundary_right="wall", boundary_top="absorbing",
                        boundary_bottom="wall")
    m = swe.SWE2D(torch.tensor(sc["bed"], dtype=dtype), torch.tensor(sc["rough"], dtype=dtype), 0.08, 0.08,
                  config=cfg, train_dem=train_dem).to(dev)
    U0 = torch.tensor(sc["U"], dtype=dtype, device=dev)
    return m, U0, sc


SERIES = {
    "inside": ([0.3, 0.5, 0.9, 1.2], [0.01, 0.03, -0.01, 0.02]),  # run 0.2..1.5 s: both ends extrapolate
    "one": ([0.7], [0.025]),
}


@pytest.mark.parametrize("case", sorted(SERIES))
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_device_series_equals_host_loop(cuda_device, case, dtype):
    m, U0, sc = model(cuda_device, dtype)
    times, wls = (torch.tensor(v, dtype=torch.float64) for v in SERIES[case])
    m.set_dynamic_boundary(times, wls)
    m.track_gauges = True
    m.gauge_indices = {f"g{i}": g for i, g in enumerate(sc["gauges"])}
    runs = {}
    for name, cb in (("device", None), ("host", lambda *a: None)):
        m.reset_gauge_tracking()
        with torch.no_grad():
            U, h = m(U0, t_end=1.3, start_time=0.2, callback=cb)
        runs[name] = (U, h, list(m.ts_dt), list(m.ts_h["g0"]), list(m.ts_z["g3"]), list(m.ts_time))
    d, hst = runs["device"], runs["host"]
    print(f"{case} {dtype}: {len(d[2])} steps device, {len(hst[2])} host; rel-L2 {relerr(d[0], hst[0]):.1e}")
    assert torch.equal(d[0], hst[0])
    assert torch.equal(d[1], hst[1])
    assert d[2] == hst[2] and d[3] == hst[3] and d[4] == hst[4] and d[5] == hst[5]


@pytest.mark.parametrize("tape", [True, False])
def test_device_series_gradients_equal_host_loop(cuda_device, tape):
    m, U0, _ = model(cuda_device, torch.float64, train_dem=True)
    m.adjoint_tape = tape
    m.adjoint_checkpoint_every = 7
    times = torch.tensor([0.0, 0.1, 0.25, 0.4], dtype=torch.float64, device=cuda_device)
    wls = torch.tensor([0.010, 0.022, 0.005, 0.018], dtype=torch.float64, device=cuda_device)
    lam = torch.randn(U0.shape, dtype=torch.float64, generator=torch.Generator().manual_seed(3)).to(cuda_device)
    grads = {}
    for name in ("device", "host"):
        m.differentiable_hmax = name == "host"
        w = wls.clone().requires_grad_(True)
        m.set_dynamic_boundary(times, w)
        Ug = U0.clone().requires_grad_(True)
        out, _ = m(Ug, t_end=0.45, start_time=0.02)
        g = torch.autograd.grad(out, (Ug, w, m.dem_correction_parameter), lam)
        grads[name] = (out.detach(), *g)
        if name == "device":
            segs = {k for k, _, _ in m.last_adjoint_segments}
            assert segs == ({"tape"} if tape else {"ckpt"}), segs
    d, h = grads["device"], grads["host"]
    e = [relerr(a, b) for a, b in zip(d, h)]
    print(f"tape={tape}: state {e[0]:.1e}; gradients U0 {e[1]:.1e}, series {e[2]:.1e}, bed {e[3]:.1e}")
    assert torch.equal(d[0], h[0])
    assert max(e[1:]) <= 1e-12
    assert float(d[2].abs().sum()) > 0  # the series reaches the loss


def test_device_run_hmax_gradient_raises(cuda_device):
    m, U0, _ = model(cuda_device, torch.float64)
    Ug = U0.clone().requires_grad_(True)
    out, hmax = m(Ug, t_end=0.2, boundary_level=0.02)
    with pytest.raises(NotImplementedError, match="differentiable_hmax"):
        hmax.sum().backward()
    # the per-step loop's hmax carries its gradient, as the reference's does
    m.differentiable_hmax = True
    Ug = U0.clone().requires_grad_(True)
    out, hmax = m(Ug, t_end=0.2, boundary_level=0.02)
    (g,) = torch.autograd.grad(hmax.sum(), Ug)
    assert float(g.abs().sum()) > 0


def test_per_step_hmax_gradient_matches_oracle_autograd(cuda_device):
    """d(sum hmax)/dU0 through the per-step loop (``differentiable_hmax``) against the oracle's
    autograd of the same run (src/solver.py:765: hmax = maximum(hmax, h) at every step), FP64.

    The oracle takes correctly rounded square roots, so both runs meet the same ties
    (dry cells tie at every step; torch splits a tie's gradient 0.5 / 0.5 on both sides)."""
    import oracle.swe_oracle as O

    m, U0, sc = model(cuda_device, torch.float64)
    m.differentiable_hmax = True
    Ug = U0.clone().requires_grad_(True)
    _, hmax = m(Ug, t_end=0.2, boundary_level=0.02)
    (g,) = torch.autograd.grad(hmax.sum(), Ug)
    old = O.SQRT_MODE
    O.SQRT_MODE = "ieee"
    try:
        Uc = U0.detach().cpu().clone().requires_grad_(True)
        bed = m.bed.detach().cpu()
        rough = m.roughness_length.detach().cpu()
        _, hc = O.forward(Uc, bed, rough, m.cfg, 0.08, 0.08, 0.2, boundary_level=0.02)
        (gc,) = torch.autograd.grad(hc.sum(), Uc)
    finally:
        O.SQRT_MODE = old
    e_h = relerr(hmax.detach(), hc.detach())
    e_g = relerr(g, gc)
    print(f"hmax vs oracle {e_h:.2e}; d(sum hmax)/dU0 vs oracle autograd {e_g:.2e}", flush=True)
    assert e_h <= 1e-13
    assert e_g <= 1e-10
    assert float(gc.abs().sum()) > 0
