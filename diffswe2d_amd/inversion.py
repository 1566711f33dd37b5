"""Boundary-control inversion on top of the HIP solver.

Restates the reference's spline inversion driver
(Examples/inverse/Monai_Valley/.../bc_inversion_train.py) for the time-step
path:

* ``clamped_bspline_basis``   clamped B-spline basis on the sample times by the
                               Cox-de Boor recursion (:189-231); dense series =
                               basis @ controls (:238-242)
* ``bounded_controls``        sigmoid bounds on the raw controls (:173-182)
* ``GaugeProblem``            piecewise-constant boundary level per output
                               chunk (:332-349), observation h + z at the gauge
                               cells after each chunk (:401-410), loss = mean
                               over gauges of each gauge's MSE plus the dense-
                               series smoothness terms (:421-452); the chunk
                               loop runs on the device (``chunked.gauge_run``)
                               or, with ``fused=False``, chunk by chunk as the
                               reference does
* ``adam_invert``             Adam with gradient-norm clipping (:632-720)

Gradients come from the hand-written adjoint (``diffswe2d_amd.adjoint``).
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np
import torch


def clamped_bspline_basis(sample_times: torch.Tensor, n_controls: int, degree: int = 3) -> torch.Tensor:
    """[len(sample_times), n_controls] clamped B-spline basis (a partition of unity).

    Knots: (degree+1)-fold at both ends, interior knots evenly spaced; basis
    functions are right-open except the last sample, which belongs to the last
    basis function.
    """
    if degree < 1 or n_controls < degree + 1:
        raise ValueError("Require degree >= 1 and number_of_controls >= degree + 1.")
    if sample_times.ndim != 1 or sample_times.numel() < 2:
        raise ValueError("sample_times must be a one-dimensional tensor with at least two values.")
    t = sample_times
    a, b = t[0], t[-1]
    n_inner = n_controls - degree - 1
    inner = (torch.linspace(a, b, n_inner + 2, dtype=t.dtype, device=t.device)[1:-1]
             if n_inner > 0 else t.new_empty(0))
    knots = torch.cat((a.repeat(degree + 1), inner, b.repeat(degree + 1)))
    x = t[:, None]
    eps = torch.finfo(t.dtype).eps
    # order 0: indicator of each knot span
    B = ((x >= knots[:-1]) & (x < knots[1:])).to(t.dtype)
    for p in range(1, degree + 1):
        m = knots.numel() - p - 1
        lo, hi = knots[:m], knots[p:p + m]
        lo1, hi1 = knots[1:1 + m], knots[p + 1:p + 1 + m]
        w_left = torch.where(hi - lo > 0, (x - lo) / (hi - lo).clamp_min(eps), 0.0)
        w_right = torch.where(hi1 - lo1 > 0, (hi1 - x) / (hi1 - lo1).clamp_min(eps), 0.0)
        B = w_left * B[:, :m] + w_right * B[:, 1:m + 1]
    B[-1].zero_()
    B[-1, -1] = 1.0
    tol = 1e-10 if t.dtype == torch.float64 else 1e-5  # the reference builds the basis in float64
    if not torch.allclose(B.sum(1), torch.ones_like(B[:, 0]), atol=tol, rtol=tol):
        raise RuntimeError("B-spline basis does not form a partition of unity.")
    return B


def bounded_controls(raw: torch.Tensor, lo: float, hi: float) -> torch.Tensor:
    return lo + (hi - lo) * torch.sigmoid(raw)


def inverse_bounded(value: float, lo: float, hi: float, like: torch.Tensor) -> torch.Tensor:
    frac = torch.full_like(like, (value - lo) / (hi - lo))
    eps = torch.finfo(like.dtype).eps
    return torch.logit(frac.clamp(eps, 1.0 - eps))


def raw_from_controls(values: torch.Tensor, lo: float, hi: float) -> torch.Tensor:
    """Raw parameters whose bounded_controls equal values (clamped into the open interval)."""
    eps = torch.finfo(values.dtype).eps
    return torch.logit(((values - lo) / (hi - lo)).clamp(eps, 1.0 - eps))


def smoothness(series: torch.Tensor):
    d1 = series[1:] - series[:-1]
    d2 = series[2:] - 2.0 * series[1:-1] + series[:-2]
    return d1.square().mean(), d2.square().mean(), series.square().mean()


@dataclass
class GaugeProblem:
    """Chunked forward run with a per-chunk water level and gauge observations."""

    model: object
    U0: torch.Tensor
    gauges: list
    chunk: float
    n_chunks: int
    start: float = 0.0
    observe: str = "water_level"
    weights: tuple = (1.0e-3, 1.0e-2, 1.0e-6)  # first, second derivative, amplitude
    stats: dict = field(default_factory=dict)
    fused: bool = True  # the whole record as one device-driven run (diffswe2d_amd.chunked)

    def __post_init__(self):
        bed = self.model.bed.detach().to(device=self.U0.device, dtype=self.U0.dtype)
        self._iy = torch.tensor([g[0] for g in self.gauges], device=self.U0.device)
        self._ix = torch.tensor([g[1] for g in self.gauges], device=self.U0.device)
        self._bed_g = bed[0, 0, self._iy, self._ix]

    def _gauge_bed(self):
        """The bed at the gauges: the current SWE2D.bed, differentiable when the DEM correction trains."""
        if torch.is_grad_enabled() and self.model._bed_trainable():
            bed = self.model.bed.to(device=self.U0.device, dtype=self.U0.dtype)
            return bed[0, 0, self._iy, self._ix]
        return self._bed_g

    def simulate(self, series: torch.Tensor) -> torch.Tensor:
        """Observations [n_chunks, n_gauges] for the dense boundary series."""
        if series.numel() < self.n_chunks:
            raise ValueError("series shorter than the number of chunks")
        bed_g = self._gauge_bed()
        if (self.fused and self.U0.shape[0] == 1 and not self.model.use_dynamic_bc
                and self.model.rainfall is None):
            from .chunked import gauge_run

            obs, _ = gauge_run(self.model, self.U0, series[:self.n_chunks], self.chunk, self.gauges)
            self.stats["steps"] = len(self.model.last_dt_series)
            return obs if self.observe == "water_level" else obs - bed_g
        s = self.U0
        obs = []
        steps = 0
        for k in range(self.n_chunks):
            s, _ = self.model(s, t_end=self.chunk, start_time=self.start + k * self.chunk,
                              boundary_level=series[k])
            steps += len(getattr(self.model, "last_dt_series", []))
            depth = s[0, 0, self._iy, self._ix]
            obs.append(depth + bed_g if self.observe == "water_level" else depth)
        self.stats["steps"] = steps
        return torch.stack(obs)

    @staticmethod
    def misfit(obs: torch.Tensor, target: torch.Tensor) -> torch.Tensor:
        """Mean over gauges of each gauge's mean squared error, accumulated in FP64 on the device."""
        d = obs.double() - target.double()
        return (d * d).mean(dim=0).mean()

    def objective(self, obs, target, series):
        first, second, amp = smoothness(series)
        w1, w2, w3 = self.weights
        return self.misfit(obs, target) + w1 * first + w2 * second + w3 * amp


def adam_invert(problem: GaugeProblem, basis: torch.Tensor, target: torch.Tensor, raw0: torch.Tensor,
                bounds=(-0.02, 0.05), lr=0.02, epochs=50, clip=1.0, target_misfit=None, log=None,
                time_budget=None):
    """Adam on raw spline controls; returns (raw, history) with (epoch, seconds, misfit) rows."""
    raw = raw0.detach().clone().requires_grad_(True)
    opt = torch.optim.Adam([raw], lr=lr)
    hist = []
    t0 = time.perf_counter()
    for ep in range(epochs):
        opt.zero_grad()
        controls = bounded_controls(raw, *bounds)
        series = basis @ controls
        obs = problem.simulate(series)
        misfit = problem.misfit(obs, target)
        total = problem.objective(obs, target, series)
        total.backward()
        torch.nn.utils.clip_grad_norm_([raw], clip)
        opt.step()
        torch.cuda.synchronize() if raw.is_cuda else None
        el = time.perf_counter() - t0
        m = float(misfit.detach())
        hist.append((ep, el, m))
        if log:
            log(f"epoch {ep}: misfit {m:.4e} ({el:.2f} s)")
        if target_misfit is not None and m <= target_misfit:
            break
        if time_budget is not None and el > time_budget:
            break
    return raw.detach(), hist


HISTORY_COLUMNS = ("epoch", "total_loss", "train_loss", "test_loss", "first_derivative", "second_derivative",
                   "amplitude", "gradient_norm", "learning_rate", "elapsed_seconds")


def spline_inversion(problem: GaugeProblem, basis: torch.Tensor, target: torch.Tensor, raw0: torch.Tensor,
                     train, test=(), bounds=(-0.02, 0.05), lr=0.02, max_epochs=500, clip=1.0, min_lr=1.0e-5,
                     lr_factor=0.5, lr_patience=3, min_delta=1.0e-8, stop_patience=10, log=None):
    """The spline inversion's training loop (bc_inversion_train.py:625-766).

    Each epoch: dense series = basis @ bounded controls; one simulation of all
    gauges; objective = mean over the ``train`` gauge columns of each gauge's
    MSE plus the weighted smoothness terms; backward; the best state is kept
    by training loss (improvement beyond ``min_delta``); gradient-norm
    clipping; Adam step; ReduceLROnPlateau on the training loss; early stop
    after ``stop_patience`` epochs without improvement.  ``test`` columns are
    reported, never differentiated.

    Returns (best raw controls, best epoch, history); each history row is a
    dict with ``HISTORY_COLUMNS``, the columns of the reference's
    loss_history.txt.
    """
    train, test = list(train), list(test)
    raw = torch.nn.Parameter(raw0.detach().clone())
    opt = torch.optim.Adam([raw], lr=lr, weight_decay=0.0)
    sched = torch.optim.lr_scheduler.ReduceLROnPlateau(opt, mode="min", factor=lr_factor, patience=lr_patience,
                                                       threshold=min_delta, threshold_mode="abs", min_lr=min_lr)
    w1, w2, w3 = problem.weights
    best, best_epoch, best_raw, waited, history = float("inf"), 0, None, 0, []
    for index in range(max_epochs):
        epoch, t0 = index + 1, time.perf_counter()
        opt.zero_grad(set_to_none=True)
        series = basis @ bounded_controls(raw, *bounds)
        obs = problem.simulate(series)
        train_loss = problem.misfit(obs[:, train], target[:, train])
        first, second, amp = smoothness(series)
        total = train_loss + w1 * first + w2 * second + w3 * amp
        test_loss = problem.misfit(obs[:, test], target[:, test]) if test else total.new_zeros(())
        total.backward()
        if not torch.isfinite(total) or raw.grad is None or not torch.isfinite(raw.grad).all():
            raise FloatingPointError("Non-finite loss or spline-control gradient.")
        train_value = float(train_loss.detach())
        grad_norm = float(torch.linalg.vector_norm(raw.grad))
        if grad_norm == 0.0:
            raise RuntimeError("All spline-control gradients are zero.")
        if train_value < best - min_delta:
            best, best_epoch, best_raw, waited = train_value, epoch, raw.detach().clone(), 0
        else:
            waited += 1
        torch.nn.utils.clip_grad_norm_([raw], clip, error_if_nonfinite=True)
        opt.step()
        sched.step(train_value)
        if raw.is_cuda:
            torch.cuda.synchronize(raw.device)
        row = dict(epoch=epoch, total_loss=float(total.detach()), train_loss=train_value,
                   test_loss=float(test_loss.detach()), first_derivative=float(first.detach()),
                   second_derivative=float(second.detach()), amplitude=float(amp.detach()),
                   gradient_norm=grad_norm, learning_rate=opt.param_groups[0]["lr"],
                   elapsed_seconds=time.perf_counter() - t0)
        history.append(row)
        if log:
            log(f"epoch {epoch:03d}: train {train_value:.6e} test {row['test_loss']:.6e} total "
                f"{row['total_loss']:.6e} grad {grad_norm:.3e} lr {row['learning_rate']:.3e} "
                f"({row['elapsed_seconds']:.3f} s)")
        if waited >= stop_patience:
            break
    if best_raw is None:
        raise RuntimeError("Training did not produce a valid best state.")
    return best_raw, best_epoch, history


MONAI_GAUGES = ((4.521, 1.196), (4.521, 1.696), (4.521, 2.196), (1.000, 1.696), (2.000, 1.696), (3.000, 1.696))


def monai_valley_setup(bathymetry, observations, device, n_gauges=3, dtype=torch.float64):
    """The set-up of the reference's Monai Valley boundary inversion (bc_inversion_train.py:84-131,
    253-321, 576-631): the NetCDF bathymetry on 68 x 42 cells of 0.08 m, a water-level left
    boundary and walls elsewhere (Manning n = 0.01), still water at 0 m, the first ``n_gauges``
    gauge cells, the observation columns interpolated to the 149 chunk ends of 0.2 s, and a
    clamped cubic B-spline over 20 controls started at a level of 0 within (-0.02, 0.05).

    Returns a dict with model, problem (GaugeProblem), target [149, n_gauges], basis, raw0, bounds."""
    from . import load_dynamic_model
    from .grid import ModelGrid

    chunk, n_chunks, n_controls, bounds = 0.2, 149, 20, (-0.02, 0.05)
    grid = ModelGrid.from_bounds(xmin=0.0, xmax=5.44, ymin=0.0, ymax=3.36, resolution=0.08, dtype=dtype,
                                 device=device)
    model, _, _, _ = load_dynamic_model(dem_filepath=str(bathymetry), model_grid=grid, device=device,
                                        boundary_left="water_level", boundary_right="wall", boundary_top="wall",
                                        boundary_bottom="wall", frictionmodel="manning", roughness_var_name=None,
                                        default_roughness_value=0.01)
    model = model.to(device=device, dtype=dtype)
    model.dem_correction_parameter.requires_grad_(False)
    model.clear_dynamic_boundary()
    model.train()
    cells = []
    for g, (gx, gy) in enumerate(MONAI_GAUGES[:n_gauges]):
        ix = int(torch.argmin(torch.abs(grid.x - gx)))
        iy = int(torch.argmin(torch.abs(grid.y - gy)))
        model.gauge_indices[f"gauge_{g}"] = (iy, ix)
        cells.append((iy, ix))
    data = np.loadtxt(observations, skiprows=1)
    ends = np.arange(1, n_chunks + 1, dtype=np.float64) * chunk
    target = torch.tensor(np.stack([np.interp(ends, data[:, 0], data[:, c]) for c in range(1, n_gauges + 1)],
                                   axis=1), dtype=dtype, device=device)
    depth = torch.clamp(0.0 - model.bed.detach(), min=0.0)
    U0 = torch.cat((depth, torch.zeros_like(depth), torch.zeros_like(depth)), dim=1)
    times = torch.arange(n_chunks, dtype=dtype, device=device) * chunk
    basis = clamped_bspline_basis(times, n_controls, 3)
    raw0 = inverse_bounded(0.0, *bounds, torch.zeros(n_controls, dtype=dtype, device=device))
    return dict(model=model, problem=GaugeProblem(model, U0, cells, chunk, n_chunks), cells=cells, U0=U0,
                target=target, basis=basis, raw0=raw0, bounds=bounds)
