"""Reverse mode for the HIP time step, behind ``torch.autograd.Function``.

The reference obtains gradients by running torch autograd through every op of
``SWE2D.step`` (src/solver.py:572-663) with the CFL dt detached (:625-626), and
through whole ``forward`` calls driven by the inversion script
(Examples/inverse/.../bc_inversion_train.py:332-412).  Here the same gradients
come from hand-written adjoint kernels (csrc/swe_adjoint.hip):

* ``differentiable_step``     one step; inputs U and the boundary level.
* ``differentiable_forward``  a duration-driven forward call.  The forward
  keeps a checkpoint of the state every K steps (``model.adjoint_checkpoint_every``,
  default 64) plus the dt series; the backward replays each K-step segment
  (recording k1, k2 per step) and sweeps it in reverse.
* ``operator_adjoint``        L^T for ``spatial_operator``.

Gradients reach U and the water-level boundary value.  The bed (DEM
correction) and roughness are treated as constants.
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib

_ADJ_MSG = "swe_adjoint_last_error_message"


def _adj_ws(model, U):
    B, _, ny, nx = U.shape
    bed, rough = model._fields(U)
    key = ("adj", U.device, U.dtype, B, ny, nx, bed.shape[0], rough.shape[0])
    ws = model._ws.get(key)
    if ws is None:
        lib = _lib.load()
        grid = _lib.make_grid(B, ny, nx, bed.shape[0], rough.shape[0], model.dx, model.dy)
        nbytes = lib.swe_adjoint_workspace_bytes(grid, U.element_size())
        buf = torch.zeros(nbytes, dtype=torch.uint8, device=U.device)
        ws = dict(buf=buf, grid=grid)
        model._ws[key] = ws
    return ws


def _fn(lib, name, dtype):
    return getattr(lib, f"{name}_{'f32' if dtype == torch.float32 else 'f64'}")


def record_step(model, U, level, max_dt, read_dt=True):
    """One forward step that also returns k1, k2 and dt (for replay).

    With read_dt False the dt is not read back (no host synchronisation) and
    max_dt is returned: a replay with the recorded dt reproduces it exactly,
    because the recorded dt never exceeds the CFL value it is the minimum of."""
    lib = _lib.load()
    bed, rough = model._fields(U)
    ws = model._workspace(U, bed, rough)
    lvl, stride = model._level_arg(level, U)
    out, k1, k2 = torch.empty_like(U), torch.empty_like(U), torch.empty_like(U)
    _lib.check(_fn(lib, "swe_step_record", U.dtype)(
        ws["grid"], ws["cfg"], _lib.ptr(U), _lib.ptr(bed), _lib.ptr(rough), _lib.ptr(lvl), stride, float(max_dt),
        ws["parity"], _lib.ptr(out), _lib.ptr(k1), _lib.ptr(k2), _lib.ptr(ws["buf"]), int(model.rows_hint),
        _lib.stream_ptr(U.device)), "step_record")
    ws["parity"] ^= 1
    dt = _lib.read_control(ws["buf"]).last_dt if read_dt else float(max_dt)
    return out, k1, k2, dt


def step_adjoint(model, U, k1, k2, level, dt, lam_out, lam_level):
    """lam_U of one step; adds the level gradient into lam_level [B]."""
    lib = _lib.load()
    bed, rough = model._fields(U)
    ws = model._workspace(U, bed, rough)
    aws = _adj_ws(model, U)
    lvl, stride = model._level_arg(level, U)
    lam_U = torch.empty_like(U)
    _lib.check(_fn(lib, "swe_step_adjoint", U.dtype)(
        ws["grid"], ws["cfg"], _lib.ptr(U), _lib.ptr(k1), _lib.ptr(k2), _lib.ptr(bed), _lib.ptr(rough),
        _lib.ptr(lvl), stride, float(dt), _lib.ptr(lam_out.contiguous()), _lib.ptr(lam_U), _lib.ptr(lam_level),
        _lib.ptr(aws["buf"]), int(model.rows_hint), _lib.stream_ptr(U.device)), "step_adjoint", _ADJ_MSG)
    return lam_U


def _level_grad(level, lam_level):
    """Reduce per-member level gradients to the level tensor's shape."""
    if level.numel() == 1:
        return lam_level.sum().reshape(level.shape).to(level.dtype)
    return lam_level.reshape(level.shape).to(level.dtype)


class _StepFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, U, level, model, max_dt):
        U_ = U.detach().contiguous()
        lvl = level.detach()
        out, k1, k2, dt = record_step(model, U_, lvl, max_dt)
        if not (dt == dt and abs(dt) != float("inf")):
            raise FloatingPointError("The calculated timestep is non-finite.")
        if dt <= 0.0:
            raise FloatingPointError(f"Non-positive timestep: {dt}.")
        ctx.model, ctx.dt = model, dt
        ctx.save_for_backward(U_, k1, k2, lvl)
        dt_t = torch.tensor(dt, dtype=torch.float64)
        ctx.mark_non_differentiable(dt_t)
        return out, dt_t

    @staticmethod
    def backward(ctx, g_out, _g_dt):
        U, k1, k2, lvl = ctx.saved_tensors
        lam_level = torch.zeros(U.shape[0], dtype=U.dtype, device=U.device)
        lam_U = step_adjoint(ctx.model, U, k1, k2, lvl, ctx.dt, g_out.contiguous(), lam_level)
        g_level = _level_grad(lvl, lam_level) if ctx.needs_input_grad[1] else None
        return lam_U, g_level, None, None


def differentiable_step(model, U, level, max_dt):
    """SWE2D.step with gradients to U and the level; returns (U_new, float dt)."""
    out, dt = _StepFn.apply(U, level, model, max_dt)
    return out, float(dt)


def _checkpointed_run(model, U0, level, duration, max_steps, cap, K):
    """Device-driven forward (no per-step host sync) keeping the state every K steps.

    Returns (U, hmax, dts, checkpoints, done, elapsed)."""
    lib = _lib.load()
    bed, rough = model._fields(U0)
    ws = model._workspace(U0, bed, rough)
    lvl, stride = model._level_arg(level, U0)
    stream = _lib.stream_ptr(U0.device)
    Ua, Ub = U0.clone(), torch.empty_like(U0)
    hmax = U0[:, 0].clone()
    cap_arg = -1.0 if cap is None else float(cap)
    _lib.check(lib.swe_run_begin(_lib.ptr(ws["buf"]), float(duration), cap_arg, U0.element_size(), stream),
               "run_begin")
    fn = _fn(lib, "swe_advance", U0.dtype)
    capacity = max(K, 64)
    dts = torch.empty(capacity, dtype=U0.dtype, device=U0.device)
    checkpoints = []
    launched, steps = 0, 0
    ctl = None
    while launched < max_steps:
        cur = Ua if steps % 2 == 0 else Ub
        checkpoints.append(cur.clone())
        n = min(K, max_steps - launched)
        if launched + n > capacity:
            capacity = 2 * (launched + n)
            dts = torch.cat((dts, torch.empty(capacity - dts.numel(), dtype=dts.dtype, device=dts.device)))
        _lib.check(fn(ws["grid"], ws["cfg"], _lib.ptr(bed), _lib.ptr(rough), _lib.ptr(lvl), stride, _lib.ptr(Ua),
                      _lib.ptr(Ub), launched, n, _lib.ptr(hmax), _lib.ptr(dts), capacity, None, 0, None, None, 0,
                      _lib.ptr(ws["buf"]), model._rows(U0), stream), "advance")
        launched += n
        ctl = _lib.read_control(ws["buf"])
        steps = ctl.steps
        if ctl.done:
            break
    if ctl.status != 0:
        if ctl.status == 1:
            raise FloatingPointError("The calculated timestep is non-finite.")
        raise FloatingPointError(f"Non-positive timestep: {ctl.bad_dt}.")
    U = Ua if steps % 2 == 0 else Ub
    dt_list = dts[:steps].cpu().tolist()
    # a checkpoint taken after the last step is not needed
    checkpoints = checkpoints[:(steps + K - 1) // K] if steps > 0 else checkpoints[:1]
    return U, hmax, dt_list, checkpoints, bool(ctl.done), ctl.elapsed[steps & 1]


class _ForwardFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, U0, level, model, duration, max_steps, cap):
        U0_ = U0.detach().contiguous()
        lvl = level.detach()
        K = int(getattr(model, "adjoint_checkpoint_every", 64))
        U, hmax, dts, checkpoints, done, elapsed = _checkpointed_run(model, U0_, lvl, duration, max_steps, cap, K)
        if not done:
            raise RuntimeError(
                f"max_steps={max_steps} reached after {elapsed:.8f} s; requested duration={duration:.8f} s.")
        ctx.model, ctx.dts, ctx.K = model, dts, K
        ctx.save_for_backward(lvl, *checkpoints)
        ctx.mark_non_differentiable(hmax)
        model.last_dt_series = dts
        return U, hmax

    @staticmethod
    def backward(ctx, g_U, _g_hmax):
        lvl, *checkpoints = ctx.saved_tensors
        model, dts, K = ctx.model, ctx.dts, ctx.K
        lam = g_U.contiguous().clone()
        lam_level = torch.zeros(lam.shape[0], dtype=lam.dtype, device=lam.device)
        n = len(dts)
        if n == 0:
            g_level = _level_grad(lvl, lam_level) if ctx.needs_input_grad[1] else None
            return lam, g_level, None, None, None, None
        if getattr(model, "adjoint_verify_replay", False):
            lam = _replay_checked(model, checkpoints, lvl, dts, K, lam, lam_level)
        else:
            lam = _replay_segments(model, checkpoints, lvl, dts, K, lam, lam_level)
        g_level = _level_grad(lvl, lam_level) if ctx.needs_input_grad[1] else None
        return lam, g_level, None, None, None, None


def _replay_segments(model, checkpoints, lvl, dts, K, lam, lam_level):
    """Each segment: one library call replays it onto a tape and sweeps it in reverse."""
    lib = _lib.load()
    U0 = checkpoints[0]
    bed, rough = model._fields(U0)
    ws = model._workspace(U0, bed, rough)
    aws = _adj_ws(model, U0)
    l, stride = model._level_arg(lvl, U0)
    n = len(dts)
    seg_len = min(K, n)
    tape = torch.empty(lib.swe_segment_tape_bytes(ws["grid"], seg_len, U0.element_size()) // U0.element_size(),
                       dtype=U0.dtype, device=U0.device)
    lam_tmp = torch.empty_like(lam)
    host_dts = (ctypes.c_double * n)(*dts)
    fn = _fn(lib, "swe_segment_adjoint", U0.dtype)
    stream = _lib.stream_ptr(U0.device)
    rows = int(model.rows_hint)
    for seg in range((n + K - 1) // K - 1, -1, -1):
        first, last = seg * K, min(n, (seg + 1) * K)
        _lib.check(fn(ws["grid"], ws["cfg"], _lib.ptr(checkpoints[seg]), _lib.ptr(bed), _lib.ptr(rough), _lib.ptr(l),
                      stride, ctypes.addressof(host_dts) + 8 * first, last - first, ws["parity"], _lib.ptr(lam),
                      _lib.ptr(lam_tmp), _lib.ptr(lam_level), _lib.ptr(tape), _lib.ptr(ws["buf"]),
                      _lib.ptr(aws["buf"]), rows, stream), "segment_adjoint", _ADJ_MSG)
        ws["parity"] ^= (last - first) & 1
    return lam


def _replay_checked(model, checkpoints, lvl, dts, K, lam, lam_level):
    """Step-by-step replay that reads every dt back and checks it against the recorded one."""
    n = len(dts)
    for seg in range((n + K - 1) // K - 1, -1, -1):
        first, last = seg * K, min(n, (seg + 1) * K)
        U = checkpoints[seg]
        tape = []
        for k in range(first, last):
            U_next, k1, k2, dt = record_step(model, U, lvl, dts[k], read_dt=True)
            if dt != dts[k]:
                raise RuntimeError(f"adjoint replay diverged at step {k}: dt {dt} != {dts[k]}")
            tape.append((U, k1, k2))
            U = U_next
        for k in range(last - 1, first - 1, -1):
            Uk, k1, k2 = tape[k - first]
            lam = step_adjoint(model, Uk, k1, k2, lvl, dts[k], lam, lam_level)
        del tape
    return lam


def differentiable_forward(model, U0, duration, absolute_start, max_steps, boundary_level, cap):
    """SWE2D.forward with gradients to U0 and an explicit boundary level."""
    if boundary_level is None and model.use_dynamic_bc:
        raise NotImplementedError("gradients through a dynamic boundary series are not supported by the adjoint")
    level = model._prepare_boundary_level(
        model.cfg.constant_level if boundary_level is None else boundary_level, U0)
    model._check_state(U0, "U0")
    return _ForwardFn.apply(U0, level, model, float(duration), int(max_steps), cap)


class _OperatorFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, U, level, model):
        U_ = U.detach().contiguous()
        lvl = level.detach()
        rhs = model.spatial_operator(U_, boundary_level=lvl)
        ctx.model = model
        ctx.save_for_backward(U_, lvl)
        return rhs

    @staticmethod
    def backward(ctx, g):
        U, lvl = ctx.saved_tensors
        model = ctx.model
        lib = _lib.load()
        bed, rough = model._fields(U)
        ws = model._workspace(U, bed, rough)
        aws = _adj_ws(model, U)
        l, stride = model._level_arg(lvl, U)
        lam_U = torch.empty_like(U)
        lam_level = torch.zeros(U.shape[0], dtype=U.dtype, device=U.device)
        _lib.check(_fn(lib, "swe_operator_adjoint", U.dtype)(
            ws["grid"], ws["cfg"], _lib.ptr(U), _lib.ptr(bed), _lib.ptr(l), stride, _lib.ptr(g.contiguous()),
            _lib.ptr(lam_U), _lib.ptr(lam_level), _lib.ptr(aws["buf"]), int(model.rows_hint),
            _lib.stream_ptr(U.device)), "operator_adjoint", _ADJ_MSG)
        g_level = _level_grad(lvl, lam_level) if ctx.needs_input_grad[1] else None
        return lam_U, g_level, None


def operator_adjoint(model, U, level):
    """spatial_operator(U, boundary_level=level) with gradients to U and level."""
    return _OperatorFn.apply(U, level, model)
