"""Surrogate-warm-started boundary-control inversion.

Pipeline (no reference counterpart beyond the plain inversion it accelerates,
Examples/inverse/.../bc_inversion_train.py):

1. ``generate_pairs``: an ensemble of seeded scenarios (uneven grids, own dt,
   csrc/swe_ensemble.hip) produces (boundary series, 6 gauge series) pairs.
2. ``surrogate.train``: the BF16 MFMA 1-D conv surrogate learns series -> gauges.
3. ``surrogate_start``: the controls that make the surrogate match the observed
   gauges (Adam through the surrogate's own backward kernels).
4. The plain solver judges that start: if its gauge misfit is not better than
   the prior's by the factor ``accept``, the prior start is used instead.
5. The adjoint inversion (``inversion.adam_invert``) finishes from the start.
"""
from __future__ import annotations

import time

import numpy as np
import torch

from . import inversion, scenarios
from .config import SWEConfig
from .ensemble import Ensemble
from .solver import SWE2D
from .surrogate import Surrogate, train


def runup_config():
    return SWEConfig(boundary_left="water_level", boundary_right="wall", boundary_top="wall", boundary_bottom="wall")


def generate_pairs(n, device, seed=5, n_controls=64, n_chunks=450, sample_dt=0.05, grids=scenarios.RUNUP_GRIDS,
                   log=None):
    """Run the ensemble; returns (series [n, n_chunks], gauges [n, 6, n_chunks], controls, stats)."""
    sc, ctrl = scenarios.ensemble_scenarios(n=n, seed=seed, n_controls=n_controls, n_chunks=n_chunks,
                                            sample_dt=sample_dt, grids=grids)
    t0 = time.perf_counter()
    ens = Ensemble(runup_config(), sc, device)
    steps = ens.run()
    sec = time.perf_counter() - t0
    res = ens.results()
    gauges = torch.stack([r["gauges"].T for r in res])  # [n, 6, n_chunks]
    series = torch.tensor(np.stack([s["levels"] for s in sc]), dtype=torch.float32, device=device)
    cells = sum(s["bed"].size * r["steps"] for s, r in zip(sc, res))
    stats = dict(scenarios=n, seconds=sec, global_steps=steps, cell_updates=cells,
                 gcell_per_s=cells / sec / 1e9, per_scenario_steps=[r["steps"] for r in res])
    if log:
        log(f"ensemble: {n} scenarios, {cells / 1e9:.2f} G cell-updates in {sec:.2f} s "
            f"({stats['gcell_per_s']:.2f} Gcell/s)")
    return series, gauges, ctrl, stats


def surrogate_start(net, basis, target, raw0, bounds=(-0.02, 0.05), iters=300, lr=0.05):
    """Controls (raw) minimising the surrogate's gauge misfit; gradients from the MFMA backward."""
    raw = raw0.detach().clone().float().requires_grad_(True)
    opt = torch.optim.Adam([raw], lr=lr)
    b = basis.float()
    tgt = target.float()
    flags = [p.requires_grad for p in net.weights]
    for p in net.weights:
        p.requires_grad_(False)  # only the input gradient is needed
    try:
        for _ in range(iters):
            opt.zero_grad()
            series = b @ inversion.bounded_controls(raw, *bounds)
            pred = net(series[None])[0].T  # [n_chunks, 6]
            loss = ((pred - tgt) / net.out_scale).square().mean()
            loss.backward()
            opt.step()
    finally:
        for p, f in zip(net.weights, flags):
            p.requires_grad_(f)
    return raw.detach().to(raw0.dtype)


class InversionSetup:
    """Coastal run-up domain, truth waveform and gauge observations of the inversion."""

    def __init__(self, device, n_controls=64, n_chunks=450, sample_dt=0.05, seed=4, nx=544, ny=344, dx=0.01,
                 dtype=torch.float32):
        p = scenarios.inversion_problem(seed=seed, nx=nx, ny=ny, dx=dx, n_controls=n_controls, n_chunks=n_chunks,
                                        sample_dt=sample_dt)
        base = p["base"]
        self.device, self.dtype = device, dtype
        self.model = SWE2D(torch.tensor(base["bed"], dtype=dtype), torch.tensor(base["rough"], dtype=dtype), dx, dx,
                           config=runup_config(), train_dem=False).to(device)
        U0 = torch.tensor(base["U"], dtype=dtype, device=device)
        self.problem = inversion.GaugeProblem(self.model, U0, base["gauges"], chunk=sample_dt, n_chunks=n_chunks)
        times = torch.tensor(p["times"], dtype=torch.float64)
        self.basis = inversion.clamped_bspline_basis(times, n_controls).to(device=device, dtype=dtype)
        truth = torch.tensor(p["truth_levels"], dtype=dtype, device=device)
        with torch.no_grad():
            self.target = self.problem.simulate(truth)
        self.truth = truth
        self.prior_raw = inversion.inverse_bounded(0.0, -0.02, 0.05, torch.zeros(n_controls, dtype=dtype,
                                                                                  device=device))

    def misfit(self, raw):
        with torch.no_grad():
            series = self.basis @ inversion.bounded_controls(raw, -0.02, 0.05)
            return float(self.problem.misfit(self.problem.simulate(series), self.target))
