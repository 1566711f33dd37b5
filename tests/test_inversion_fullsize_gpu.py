"""The boundary-control inversion's gradient at its full size, on the GPU.

The workload is the coastal run-up at 544 x 344, dx = 0.01 m: 64 cubic-spline
controls of the west water level, 6 gauges sampled at the ends of 450 chunks of
0.05 s (about 10,800 FP64 steps).  The loss is the mean over gauges of the mean
squared misfit against a truth run.

The reference detaches dt (src/solver.py:625-626), so its autograd gradient,
and the hand-written adjoint's, is the derivative of the map with every step's
dt held fixed.  The reference's autograd cannot hold this many steps, and the
CPU oracle would take hours, so the gradient is checked here against central
differences of that dt-frozen map, run on the GPU: each perturbed run takes
the unperturbed run's dt series step for step through the library's fixed-dt
steps (swe_config.fixed_dt: dt = max_dt, not min(CFL dt, max_dt)), chunk by
chunk with each chunk's level.  The differences are taken on 8 seeded controls.

The dt-frozen map is piecewise smooth: the limiter switches branches and cells
cross the dry threshold along the moving shoreline throughout the record, so
finite differences of it scatter with the step size.  Measured at this size
(controls 3, 11, 12): central differences with steps from 1e-8 to 1e-4 and
least-squares slopes over 9 to 17 points spread up to 3e-3 relative around
each other, and the adjoint lies inside that spread every time.  The bar,
1e-3 at a step of 1e-6 (worst control 4.9e-4), is that roughness, not the
adjoint's accuracy, which the reduced-horizon tests pin to 3.5e-5 against
central differences and to 1.8e-8 against the reference's own autograd.
"""
import ctypes

import numpy as np
import pytest
import torch

from diffswe2d_amd import _lib, scenarios, warmstart

pytestmark = pytest.mark.gpu


def frozen_loss(setup, series, dts_per_chunk):
    """Gauge misfit of the run that takes the given dt series (FP64, fixed-dt steps)."""
    prob, m = setup.problem, setup.model
    lib = _lib.load()
    U = prob.U0.detach().clone()
    dev = U.device
    bed, rough = m._fields(U)
    grid = _lib.make_grid(1, U.shape[-2], U.shape[-1], 1, 1, m.dx, m.dy)
    cfg = _lib.make_config(m.cfg)
    cfg.fixed_dt = 1
    ws = torch.zeros(lib.swe_workspace_bytes(ctypes.byref(grid), 8), dtype=torch.uint8, device=dev)
    stream = _lib.stream_ptr(dev)
    _lib.check(lib.swe_workspace_init(_lib.ptr(ws), ws.numel(), stream), "workspace_init")
    out = torch.empty_like(U)
    parity = 0
    obs = []
    with torch.no_grad():
        for k, dts in enumerate(dts_per_chunk):
            lvl = series[k].reshape(1).contiguous()
            for dt in dts:
                _lib.check(lib.swe_step_f64(ctypes.byref(grid), ctypes.byref(cfg), _lib.ptr(U), _lib.ptr(bed),
                                            _lib.ptr(rough), _lib.ptr(lvl), 0, float(dt), parity, _lib.ptr(out),
                                            _lib.ptr(ws), 0, stream), "fixed-dt step")
                parity ^= 1
                U, out = out, U
            obs.append(U[0, 0, prob._iy, prob._ix] + prob._bed_g)
        last = _lib.read_control(ws).last_dt
    assert last == float(dts_per_chunk[-1][-1])
    return float(prob.misfit(torch.stack(obs), setup.target))


def test_full_size_spline_gradient_central_differences(cuda_device):
    setup = warmstart.InversionSetup(cuda_device, dtype=torch.float64)
    prob, basis = setup.problem, setup.basis
    rng = np.random.default_rng(11)
    c0_np, _ = scenarios.random_waveform_controls(rng, basis.shape[1])
    c0 = torch.tensor(c0_np, dtype=torch.float64, device=cuda_device)

    # adjoint gradient through the device-driven chunked run (taped reverse)
    ctrl = c0.clone().requires_grad_(True)
    loss = prob.misfit(prob.simulate(basis @ ctrl), setup.target)
    loss.backward()
    grad = ctrl.grad.detach().cpu().numpy()
    steps = prob.stats["steps"]

    # the same run chunk by chunk, for each chunk's dt series
    dts_per_chunk = []
    with torch.no_grad():
        s = prob.U0
        series0 = basis @ c0
        for k in range(prob.n_chunks):
            s, _ = setup.model(s, t_end=prob.chunk, start_time=k * prob.chunk, boundary_level=series0[k])
            dts_per_chunk.append([float(v) for v in setup.model.last_dt_series])
    assert sum(len(d) for d in dts_per_chunk) == steps

    base = frozen_loss(setup, series0, dts_per_chunk)
    loss = float(loss.detach())
    print(f"run-up inversion 544x344 FP64: {steps} steps; loss {loss:.10e}, frozen-dt map "
          f"{base:.10e}", flush=True)
    assert abs(base - loss) <= 1e-12 * loss

    idx = np.sort(np.random.default_rng(12).choice(basis.shape[1], size=8, replace=False))
    eps = 1e-6
    worst = 0.0
    for i in idx:
        e = torch.zeros_like(c0)
        e[i] = eps
        lp = frozen_loss(setup, basis @ (c0 + e), dts_per_chunk)
        lm = frozen_loss(setup, basis @ (c0 - e), dts_per_chunk)
        fd = (lp - lm) / (2 * eps)
        rel = abs(grad[i] - fd) / max(abs(fd), 1e-300)
        worst = max(worst, rel)
        print(f"control {i:2d}: adjoint {grad[i]: .8e}, central FD (dt frozen) {fd: .8e}, rel {rel:.2e}", flush=True)
    assert worst <= 1e-3

    # the FP32 adjoint at the same point: its distance from the FP64 gradient at full size
    s32 = warmstart.InversionSetup(cuda_device, dtype=torch.float32)
    c32 = c0.float().clone().requires_grad_(True)
    loss32 = s32.problem.misfit(s32.problem.simulate(s32.basis @ c32), s32.target)
    loss32.backward()
    g32 = c32.grad.detach().double().cpu().numpy()
    e32 = float(np.linalg.norm(g32 - grad) / np.linalg.norm(grad))
    print(f"FP32 adjoint gradient vs FP64: rel-L2 {e32:.2e} ({s32.problem.stats['steps']} FP32 steps; loss "
          f"{float(loss32.detach()):.8e})", flush=True)
    assert e32 <= 0.05
