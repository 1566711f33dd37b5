"""Surrogate-warm-started boundary-control inversion.

Pipeline (no reference counterpart beyond the plain inversion it accelerates,
Examples/inverse/.../bc_inversion_train.py):

1. ``generate_pairs``: an ensemble of seeded scenarios (uneven grids, own dt,
   csrc/swe_ensemble.hip) produces (boundary series, 6 gauge series) pairs,
   sampled at the same physical gauge points at every resolution.  The pairs
   only train the surrogate, so by default the ensemble runs the library's
   tolerance-bounded fast FP32 arithmetic (gauge series within ~1e-6 of the
   reference rounding's; the surrogate's own error is ~1e-2).  The inversions
   themselves and the plain-solver judgement use the reference rounding.
2. ``surrogate.train``: the BF16 MFMA 1-D conv surrogate learns series -> gauges,
   conditioned on the resolution (log2 of dx / 1 cm) of the run.
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

from . import autotune, inversion, scenarios
from .config import SWEConfig
from .ensemble import Ensemble
from .solver import SWE2D
from .surrogate import Surrogate, default_specs, train


def runup_config():
    return SWEConfig(boundary_left="water_level", boundary_right="wall", boundary_top="wall", boundary_bottom="wall")


def resolution_condition(dx):
    """Surrogate condition channel of a run on a grid of spacing dx (metres)."""
    return float(np.log2(dx / 0.01))


def generate_pairs(n, device, seed=5, n_controls=64, n_chunks=450, sample_dt=0.05, grids=scenarios.RUNUP_GRIDS,
                   log=None, tune=True, arithmetic="fast"):
    """Run the ensemble; returns (series [n, n_chunks], gauges [n, 6, n_chunks], controls, stats).

    arithmetic: the ensemble's FP32 arithmetic ("fast" or "reference", see ``Ensemble``).

    With ``tune`` the rows per strip of each grid size are tuned first unless
    the table already holds them (``autotune.tune_ensemble_rows``; its time is
    stats["tune_seconds"], outside stats["seconds"]).  stats["cond"] holds each
    pair's resolution condition [n, 1]."""
    sc, ctrl = scenarios.ensemble_scenarios(n=n, seed=seed, n_controls=n_controls, n_chunks=n_chunks,
                                            sample_dt=sample_dt, grids=grids)
    tune_s = 0.0
    if tune:
        if not autotune.lookup_ensemble(device, autotune.ensemble_mix(sc), arithmetic):
            t = time.perf_counter()
            autotune.tune_ensemble_rows(runup_config(), sc, device, arithmetic=arithmetic)
            tune_s = time.perf_counter() - t
    t0 = time.perf_counter()
    ens = Ensemble(runup_config(), sc, device, arithmetic=arithmetic)
    steps = ens.run(log=log)
    sec = time.perf_counter() - t0
    res = ens.results()
    gauges = torch.stack([r["gauges"].T for r in res])  # [n, 6, n_chunks]
    series = torch.tensor(np.stack([s["levels"] for s in sc]), dtype=torch.float32, device=device)
    cells = sum(s["bed"].size * r["steps"] for s, r in zip(sc, res))
    cond = torch.tensor([[resolution_condition(s["dx"])] for s in sc], dtype=torch.float32, device=device)
    stats = dict(scenarios=n, arithmetic=arithmetic, seconds=sec, tune_seconds=tune_s, global_steps=steps,
                 cell_updates=cells,
                 gcell_per_s=cells / sec / 1e9, per_scenario_steps=[r["steps"] for r in res], cond=cond)
    if log:
        log(f"ensemble: {n} scenarios, {cells / 1e9:.2f} G cell-updates in {sec:.2f} s "
            f"({stats['gcell_per_s']:.2f} Gcell/s)")
    return series, gauges, ctrl, stats


def surrogate_start(net, basis, target, raw0, bounds=(-0.02, 0.05), iters=300, lr=0.05, dx=None, info=None):
    """Controls (raw) minimising the surrogate's gauge misfit; gradients from the MFMA backward.

    raw0 is one start [n_controls] or several [k, n_controls]; several starts
    are refined together as one batch through the surrogate and the one with
    the lowest surrogate misfit is returned.  dx: grid spacing of the inverted
    model (needed when the surrogate is conditioned).  ``info`` (a dict, if
    given) receives the per-start surrogate misfits before and after."""
    starts = raw0.detach().float().reshape(-1, raw0.shape[-1]).clone().requires_grad_(True)
    k = starts.shape[0]
    cond = None
    if net.n_cond:
        if dx is None:
            raise ValueError("a conditioned surrogate needs the grid spacing dx")
        cond = torch.full((k, net.n_cond), resolution_condition(dx), dtype=torch.float32, device=starts.device)
    opt = torch.optim.Adam([starts], lr=lr)
    bT = basis.float().T.contiguous()  # [n_controls, n_chunks]
    tgt = target.float().T[None]  # [1, 6, n_chunks]

    def per_start():
        pred = net(inversion.bounded_controls(starts, *bounds) @ bT, cond)  # [k, 6, n_chunks]
        return ((pred - tgt) / net.out_scale).square().mean(dim=(1, 2))

    flags = [p.requires_grad for p in net.weights]
    for p in net.weights:
        p.requires_grad_(False)  # only the input gradient is needed
    try:
        with torch.no_grad():
            first = per_start()
        for _ in range(iters):
            opt.zero_grad()
            per_start().sum().backward()  # the starts are independent: the sum's gradient is each one's own
            opt.step()
        with torch.no_grad():
            last = per_start()
    finally:
        for p, f in zip(net.weights, flags):
            p.requires_grad_(f)
    if info is not None:
        info["surrogate_misfit_before"] = first.tolist()
        info["surrogate_misfit_after"] = last.tolist()
    return starts[int(torch.argmin(last))].detach().to(raw0.dtype)


def nearest_pairs(gauges, target, k=8, mask=None):
    """Indices of the k ensemble pairs whose gauge series [n, 6, n_chunks] best match target [n_chunks, 6]."""
    mis = (gauges - target.T[None].to(gauges.dtype)).square().mean(dim=(1, 2))
    if mask is not None:
        mis = torch.where(mask, mis, torch.full_like(mis, float("inf")))
    order = torch.argsort(mis)
    return order[:min(k, int(torch.isfinite(mis).sum()))], mis


def judge_start(setup, raw_ws, m_prior, accept=0.5):
    """The plain solver's verdict on a surrogate start.

    Runs the start through the solver (``setup.misfit``) and keeps it only when
    its gauge misfit is below ``accept`` times the prior's (``m_prior``, the
    misfit of the prior start); otherwise the prior is the start.  Returns
    (start raw controls, fallback taken, the surrogate start's misfit)."""
    m_ws = setup.misfit(raw_ws)
    fallback = not (m_ws < accept * m_prior)
    return (setup.prior_raw if fallback else raw_ws), fallback, m_ws


class InversionSetup:
    """Coastal run-up domain, truth waveform and gauge observations of the inversion."""

    def __init__(self, device, n_controls=64, n_chunks=450, sample_dt=0.05, seed=4, nx=544, ny=344, dx=0.01,
                 dtype=torch.float32):
        p = scenarios.inversion_problem(seed=seed, nx=nx, ny=ny, dx=dx, n_controls=n_controls, n_chunks=n_chunks,
                                        sample_dt=sample_dt)
        base = p["base"]
        self.device, self.dtype, self.dx = device, dtype, dx
        self.model = SWE2D(torch.tensor(base["bed"], dtype=dtype), torch.tensor(base["rough"], dtype=dtype), dx, dx,
                           config=runup_config(), train_dem=False).to(device)
        U0 = torch.tensor(base["U"], dtype=dtype, device=device)
        if dtype == torch.float32:  # strip height for this grid (the step is latency-bound at this size)
            autotune.tune_rows(self.model, U0, candidates=(1, 2, 4, 8, 16, 32), save=False)
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


def run(device, n_scenarios=256, n_controls=64, n_chunks=450, grids=scenarios.RUNUP_GRIDS, epochs_plain=60,
        surrogate_epochs=40, channels=160, accept=0.5, log=print, seed=5, n_starts=8, truth_seed=4,
        ensemble_arithmetic="fast"):
    """Full pipeline; returns a report with the time-to-misfit of both inversions.

    The plain inversion starts from the prior (level 0) and runs ``epochs_plain``
    adjoint epochs; its final misfit is the bar.  The warm-started inversion is
    timed from the surrogate start (surrogate inversion + one plain-solver
    judgement) until it reaches that bar; ensemble and training time are
    reported separately (a one-off cost shared by every later inversion).
    """
    report = {}
    setup = InversionSetup(device, n_controls=n_controls, n_chunks=n_chunks, seed=truth_seed)
    t0 = time.perf_counter()
    raw_p, hist_p = inversion.adam_invert(setup.problem, setup.basis, setup.target, setup.prior_raw,
                                          epochs=epochs_plain, lr=0.02,
                                          log=(lambda m: log("plain " + m)) if log else None)
    t_plain = time.perf_counter() - t0
    bar = min(h[2] for h in hist_p)
    report["plain"] = dict(seconds=t_plain, epochs=len(hist_p), best_misfit=bar, history=hist_p,
                           steps_per_forward=setup.problem.stats.get("steps"))
    series, gauges, ctrl, stats = generate_pairs(n_scenarios, device, seed=seed, n_controls=n_controls,
                                                 n_chunks=n_chunks, grids=grids, log=log,
                                                 arithmetic=ensemble_arithmetic)
    report["ensemble"] = {k: v for k, v in stats.items() if k not in ("per_scenario_steps", "cond")}
    net = Surrogate(default_specs(channels=channels, n_cond=1), device=device, seed=0)
    t1 = time.perf_counter()
    hist_s = train(net, series, gauges, epochs=surrogate_epochs, batch=32, lr=2e-3, cond=stats["cond"],
                   log=(lambda m: log(m)) if log else None)
    report["surrogate"] = dict(params=net.n_params(), seconds=time.perf_counter() - t1, loss=hist_s[-1][2])
    t2 = time.perf_counter()
    # starts: the prior and the ensemble pairs whose gauges already resemble the
    # observations (same resolution), all refined through the surrogate at once
    same_grid = (stats["cond"][:, 0] == resolution_condition(setup.dx))
    idx, pair_mis = nearest_pairs(gauges, setup.target, k=n_starts, mask=same_grid)
    ctrl_t = torch.as_tensor(ctrl, dtype=setup.dtype, device=device)[idx]
    starts = torch.cat((setup.prior_raw[None], inversion.raw_from_controls(ctrl_t, -0.02, 0.05)))
    info = {}
    raw_ws = surrogate_start(net, setup.basis, setup.target, starts, dx=setup.dx, info=info)
    m_prior = hist_p[0][2]
    start, fallback, m_ws = judge_start(setup, raw_ws, m_prior, accept)
    t_start = time.perf_counter() - t2
    raw_w, hist_w = inversion.adam_invert(setup.problem, setup.basis, setup.target, start, epochs=epochs_plain,
                                          lr=0.01 if not fallback else 0.02, target_misfit=bar,
                                          log=(lambda m: log("warm " + m)) if log else None)
    t_warm = t_start + hist_w[-1][1]
    reached = min(h[2] for h in hist_w) <= bar
    # time at which the plain run first reached the warm run's misfit levels, for the curve
    report["warm"] = dict(seconds_to_bar=t_warm if reached else None, start_seconds=t_start,
                          nearest_pair_gauge_misfit=float(pair_mis[idx[0]]) if len(idx) else None, **info,
                          start_misfit=m_ws, prior_misfit=m_prior, fallback=fallback, epochs=len(hist_w),
                          history=hist_w, reached=reached)
    report["ratio_warm_over_plain"] = (t_warm / t_plain) if reached else None
    return report


if __name__ == "__main__":
    import argparse
    import json

    ap = argparse.ArgumentParser()
    ap.add_argument("--scenarios", type=int, default=256)
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--surrogate-epochs", type=int, default=40)
    ap.add_argument("--truth-seed", type=int, default=4, help="seed of the truth waveform being inverted")
    ap.add_argument("--out", default="warmstart.json")
    a = ap.parse_args()
    rep = run(torch.device("cuda:0"), n_scenarios=a.scenarios, epochs_plain=a.epochs,
              surrogate_epochs=a.surrogate_epochs, truth_seed=a.truth_seed)
    with open(a.out, "w") as f:
        json.dump(rep, f, indent=1)
    print(json.dumps({k: v for k, v in rep.items() if k not in ("plain", "warm")}))
    print(json.dumps({k: {kk: vv for kk, vv in rep[k].items() if kk != "history"} for k in ("plain", "warm")}))
