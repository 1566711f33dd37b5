"""The whole chunked record as one device-driven run against the reference's chunk-by-chunk loop.

The chunk-by-chunk loop (``GaugeProblem(fused=False)``) calls ``SWE2D.forward``
once per chunk with that chunk's level, as the reference's inversion driver
does (bc_inversion_train.py:332-349, 401-410); it is itself checked against the
oracle and the reference's autograd in test_forward_gpu / test_adjoint_gpu.
The fused run must reproduce its observations and final state bit for bit and
its gradients to rounding.
"""
import pytest
import torch

import diffswe2d_amd as swe
from diffswe2d_amd import chunked, inversion, scenarios

pytestmark = pytest.mark.gpu


def problem(dev, dtype, fused, n_chunks=24):
    sc = scenarios.coastal_runup(nx=68, ny=43, dx=0.08)
    cfg = swe.SWEConfig(boundary_left="water_level", boundary_right="wall", boundary_top="wall",
                        boundary_bottom="wall")
    m = swe.SWE2D(torch.tensor(sc["bed"], dtype=dtype), torch.tensor(sc["rough"], dtype=dtype), 0.08, 0.08,
                  config=cfg, train_dem=False).to(dev)
    U0 = torch.tensor(sc["U"], dtype=dtype, device=dev)
    gauges = scenarios.physical_gauges(scenarios.coastal_runup()["gauges"], 0.01, 0.08)
    prob = inversion.GaugeProblem(m, U0, gauges, chunk=0.05, n_chunks=n_chunks, fused=fused)
    g = torch.Generator().manual_seed(7)
    series = (0.02 * torch.rand(n_chunks, generator=g, dtype=torch.float64)).to(dtype=dtype, device=dev)
    return m, prob, series


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_fused_observations_equal_chunk_loop(cuda_device, dtype):
    _, pf, series = problem(cuda_device, dtype, True)
    _, pl, _ = problem(cuda_device, dtype, False)
    with torch.no_grad():
        of = pf.simulate(series)
        ol = pl.simulate(series)
    print(f"{dtype}: {pf.stats['steps']} steps fused, {pl.stats['steps']} chunk by chunk")
    assert pf.stats["steps"] == pl.stats["steps"]
    assert torch.equal(of, ol)


def test_fused_final_state_and_dt_series(cuda_device):
    m, pf, series = problem(cuda_device, torch.float32, True)
    with torch.no_grad():
        obs, U = chunked.gauge_run(m, pf.U0, series, 0.05, pf.gauges)
        dts = list(m.last_dt_series)
        s = pf.U0
        loop_dts = []
        for k in range(series.numel()):
            s, _ = m(s, t_end=0.05, boundary_level=series[k])
            loop_dts += [float(v) for v in m.last_dt_series]
    assert torch.equal(U, s)
    assert dts == loop_dts


@pytest.mark.parametrize("mode", ["tape", "replay", "mixed"])
def test_fused_gradient_equals_chunk_loop(cuda_device, mode):
    dtype = torch.float64
    m, pf, series = problem(cuda_device, dtype, True)
    ml, pl, _ = problem(cuda_device, dtype, False)
    m.adjoint_checkpoint_every = 5
    state = 3 * 43 * 68 * 8
    m.adjoint_tape = {"tape": True, "replay": False, "mixed": 5 * 16 * state}[mode]  # mixed: 5 taped segments
    target = torch.zeros(series.numel(), len(pf.gauges), dtype=dtype, device=cuda_device) + 0.004

    def grads(prob, model):
        s = series.clone().requires_grad_(True)
        U0 = prob.U0.clone().requires_grad_(True)
        prob.U0 = U0
        loss = prob.misfit(prob.simulate(s), target)
        gs, gu = torch.autograd.grad(loss, (s, U0))
        return float(loss.detach()), gs, gu

    lf, gsf, guf = grads(pf, m)
    kinds = {k for k, _, _ in m.last_adjoint_segments}
    ll, gsl, gul = grads(pl, ml)
    es = float((gsf - gsl).norm() / gsl.norm())
    eu = float((guf - gul).norm() / gul.norm())
    print(f"{mode}: segments {sorted(kinds)}; loss {lf:.10e} vs {ll:.10e}; series grad rel diff {es:.2e}, "
          f"U0 grad rel diff {eu:.2e}")
    assert kinds == {"tape": {"tape"}, "replay": {"ckpt"}, "mixed": {"tape", "ckpt"}}[mode]
    assert lf == ll
    assert es <= 1e-12 and eu <= 1e-12


def test_gauge_run_rejects_bad_input(cuda_device):
    m, pf, series = problem(cuda_device, torch.float32, True)
    with pytest.raises(ValueError):
        chunked.gauge_run(m, pf.U0.expand(2, -1, -1, -1).contiguous(), series, 0.05, pf.gauges)
    with pytest.raises(IndexError):
        chunked.gauge_run(m, pf.U0, series, 0.05, [(43, 0)])
    with pytest.raises(RuntimeError):
        chunked.gauge_run(m, pf.U0, series, 0.05, pf.gauges, max_steps=3)
