"""Learned warm start on a reduced inversion: the trained surrogate's start beats the prior.

Reduced sizes (coarse 136 x 86 basin, 120 chunks of 0.05 s, 16 controls,
48 training scenarios) keep the test to a few seconds; bench-scale numbers come
from diffswe2d_amd.warmstart at full size.
"""
import pytest
import torch

from diffswe2d_amd import inversion, warmstart
from diffswe2d_amd.surrogate import Surrogate, default_specs, train

pytestmark = pytest.mark.gpu


def test_surrogate_warm_start_beats_prior(cuda_device):
    n_ctrl, n_chunks = 16, 120
    grids = ((43, 68, 0.08), (86, 136, 0.04))
    series, gauges, ctrl, stats = warmstart.generate_pairs(48, cuda_device, seed=11, n_controls=n_ctrl,
                                                          n_chunks=n_chunks, grids=grids, log=print)
    assert torch.isfinite(gauges).all()
    net = Surrogate(default_specs(channels=64, dilations=(1, 2, 4, 8, 16, 32), n_cond=1), device=cuda_device, seed=0)
    hist = train(net, series, gauges, epochs=80, batch=16, lr=3e-3, cond=stats["cond"])
    print(f"surrogate: {net.n_params()} params, loss {hist[0][2]:.3e} -> {hist[-1][2]:.3e}")
    setup = warmstart.InversionSetup(cuda_device, n_controls=n_ctrl, n_chunks=n_chunks, nx=136, ny=86, dx=0.04)
    m_prior = setup.misfit(setup.prior_raw)
    raw_ws = warmstart.surrogate_start(net, setup.basis, setup.target, setup.prior_raw, iters=200, dx=setup.dx)
    m_ws = setup.misfit(raw_ws)
    print(f"plain-solver misfit: prior {m_prior:.4e}, surrogate start {m_ws:.4e}")
    assert m_ws < 0.7 * m_prior
    # several starts (prior + nearest ensemble pairs on the same grid) refined as one batch
    same = stats["cond"][:, 0] == warmstart.resolution_condition(setup.dx)
    idx, mis = warmstart.nearest_pairs(gauges, setup.target, k=4, mask=same)
    assert len(idx) == 4 and bool(same[idx].all())
    starts = torch.cat((setup.prior_raw[None],
                        inversion.raw_from_controls(torch.as_tensor(ctrl, device=cuda_device)[idx].float(), -0.02, 0.05)))
    info = {}
    raw_multi = warmstart.surrogate_start(net, setup.basis, setup.target, starts, iters=200, dx=setup.dx, info=info)
    assert len(info["surrogate_misfit_after"]) == 5
    assert min(info["surrogate_misfit_after"]) <= info["surrogate_misfit_after"][0]
    m_multi = setup.misfit(raw_multi)
    print(f"multi-start surrogate misfits {info['surrogate_misfit_after']}; plain-solver misfit {m_multi:.4e}")
    assert m_multi < 0.7 * m_prior
    # and the adjoint improves on it further
    raw, hist = inversion.adam_invert(setup.problem, setup.basis, setup.target, raw_ws, epochs=3, lr=0.02)
    assert hist[-1][2] <= m_ws * 1.01


def test_untrained_surrogate_falls_back_to_the_prior(cuda_device):
    """A surrogate with random weights proposes a start the plain solver rejects: the
    inversion starts from the prior, as if there were no surrogate."""
    n_ctrl, n_chunks = 16, 60
    setup = warmstart.InversionSetup(cuda_device, n_controls=n_ctrl, n_chunks=n_chunks, nx=136, ny=86, dx=0.04)
    net = Surrogate(default_specs(channels=32, dilations=(1, 2, 4), n_cond=1), device=cuda_device, seed=3)
    m_prior = setup.misfit(setup.prior_raw)
    raw_ws = warmstart.surrogate_start(net, setup.basis, setup.target, setup.prior_raw, iters=50, dx=setup.dx)
    start, fallback, m_ws = warmstart.judge_start(setup, raw_ws, m_prior, accept=0.5)
    print(f"untrained surrogate start: plain-solver misfit {m_ws:.4e} against the prior's {m_prior:.4e}; "
          f"fallback {fallback}")
    assert fallback
    assert torch.equal(start, setup.prior_raw)
    # the inversion from the fallback start is the inversion from the prior
    _, hist_a = inversion.adam_invert(setup.problem, setup.basis, setup.target, start, epochs=2, lr=0.02)
    _, hist_b = inversion.adam_invert(setup.problem, setup.basis, setup.target, setup.prior_raw, epochs=2, lr=0.02)
    assert [h[2] for h in hist_a] == [h[2] for h in hist_b]
    # a start is kept only when it beats the prior by the factor: accept = 0 always falls back
    _, fb0, _ = warmstart.judge_start(setup, setup.prior_raw, m_prior, accept=0.0)
    assert fb0
