"""Reverse mode for the HIP time step, behind ``torch.autograd.Function``.

The reference obtains gradients by running torch autograd through every op of
``SWE2D.step`` (src/solver.py:572-663) with the CFL dt detached (:625-626), and
through whole ``forward`` calls driven by the inversion script
(Examples/inverse/.../bc_inversion_train.py:332-412).  Here the same gradients
come from hand-written adjoint kernels (csrc/swe_adjoint.hip):

* ``differentiable_step``     one step; inputs U, the boundary level and the bed.
* ``differentiable_forward``  a duration-driven forward call.  The forward
  records the tape of each K-step segment (states, k1, k2) while it fits in
  memory and otherwise keeps a checkpoint of the segment's first state
  (``model.adjoint_checkpoint_every``, default 64); the backward sweeps recorded
  segments in reverse and replays checkpointed ones first.
* ``operator_adjoint``        L^T for ``spatial_operator``.

Gradients reach U, the water-level boundary value (a scalar, one per member, a
dynamic series through ``chunked.dynamic_forward`` or the per-step loop) and
the bed (``SWE2D.bed``, and through it a trainable DEM correction).  Roughness
is a buffer in the reference and is not differentiated; the source terms of
``step`` / ``forward`` (rain, infiltration) are not differentiated either
(``spatial_operator`` returns their gradients).  The hmax of a device-driven
forward carries no gradient: a backward through it raises, and
``model.differentiable_hmax = True`` runs the per-step loop instead.
"""
from __future__ import annotations

import contextlib
import ctypes

import torch

from . import _lib
from .solver import sync_parity

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
        with torch.inference_mode(False):
            buf = torch.zeros(nbytes, dtype=torch.uint8, device=U.device)
        ws = dict(buf=buf, grid=grid)
        model._ws[key] = ws
    return ws


def _fn(lib, name, dtype):
    return getattr(lib, f"{name}_{'f32' if dtype == torch.float32 else 'f64'}")


def record_step(model, U, level, max_dt, read_dt=True, forcing=None):
    """One forward step that also returns k1, k2 and dt (for replay).

    With read_dt False the dt is not read back (no host synchronisation) and
    max_dt is returned: a replay with the recorded dt reproduces it exactly,
    because the recorded dt never exceeds the CFL value it is the minimum of.
    forcing: None or (infiltration, model time) of a step with source terms."""
    lib = _lib.load()
    bed, rough = model._fields(U)
    ws = model._workspace(U, bed, rough)
    lvl, stride = model._level_arg(level, U)
    out, k1, k2 = torch.empty_like(U), torch.empty_like(U), torch.empty_like(U)
    if forcing is not None:
        fz, keep = model._forcing(U, *forcing)
        _lib.check(_fn(lib, "swe_step_forced", U.dtype)(
            ws["grid"], ws["cfg"], ctypes.byref(fz), _lib.ptr(U), _lib.ptr(bed), _lib.ptr(rough), _lib.ptr(lvl),
            stride, float(max_dt), ws["parity"], _lib.ptr(out), _lib.ptr(k1), _lib.ptr(k2), _lib.ptr(ws["buf"]),
            int(model.rows_hint), _lib.stream_ptr(U.device)), "step_forced")
    else:
        _lib.check(_fn(lib, "swe_step_record", U.dtype)(
            ws["grid"], ws["cfg"], _lib.ptr(U), _lib.ptr(bed), _lib.ptr(rough), _lib.ptr(lvl), stride,
            float(max_dt), ws["parity"], _lib.ptr(out), _lib.ptr(k1), _lib.ptr(k2), _lib.ptr(ws["buf"]),
            int(model.rows_hint), _lib.stream_ptr(U.device)), "step_record")
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


def _bed_input(model):
    """The differentiable bed (SWE2D.bed, src/solver.py:134-147) when the DEM correction is trainable, else None."""
    p = getattr(model, "dem_correction_parameter", None)
    return model.bed if p is not None and p.requires_grad else None


def _bed_region(model, U):
    """The adjoint workspace's bed adjoint [B, 1, ny, nx] (accumulated by adjoint calls made with grad_bed set)."""
    lib = _lib.load()
    aws = _adj_ws(model, U)
    B, _, ny, nx = U.shape
    off = lib.swe_adjoint_bed_grad_offset(aws["grid"], U.element_size())
    n = B * ny * nx * U.element_size()
    return aws["buf"][off:off + n].view(U.dtype).view(B, 1, ny, nx)


@contextlib.contextmanager
def _bed_gradient(model, U, wanted):
    """Clear the bed adjoint and let the adjoint calls inside accumulate into it."""
    if not wanted:
        yield None
        return
    region = _bed_region(model, U)
    region.zero_()
    model._grad_bed = True
    try:
        yield region
    finally:
        model._grad_bed = False


def _bed_grad(region, bed_meta):
    """The bed adjoint reduced to the bed's shape (summed over members sharing one bed)."""
    shape, dtype, device = bed_meta
    g = region.sum(dim=0, keepdim=True) if shape[0] == 1 and region.shape[0] > 1 else region
    return g.reshape(shape).to(device=device, dtype=dtype).clone()


def _meta(bed):
    return None if bed is None else (tuple(bed.shape), bed.dtype, bed.device)


def _level_grad(level, lam_level):
    """Reduce per-member level gradients to the level tensor's shape."""
    if level.numel() == 1:
        return lam_level.sum().reshape(level.shape).to(level.dtype)
    return lam_level.reshape(level.shape).to(level.dtype)


class _StepFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, U, level, bed, model, max_dt, forcing):
        U_ = U.detach().contiguous()
        lvl = level.detach()
        ctx.bed_meta = _meta(bed)
        # source terms enter k1 and k2 only; their adjoint with respect to U is zero
        out, k1, k2, dt = record_step(model, U_, lvl, max_dt, forcing=forcing)
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
        want_bed = ctx.needs_input_grad[2]
        with _bed_gradient(ctx.model, U, want_bed) as region:
            lam_U = step_adjoint(ctx.model, U, k1, k2, lvl, ctx.dt, g_out.contiguous(), lam_level)
        g_level = _level_grad(lvl, lam_level) if ctx.needs_input_grad[1] else None
        g_bed = _bed_grad(region, ctx.bed_meta) if want_bed else None
        return lam_U, g_level, g_bed, None, None, None


def differentiable_step(model, U, level, max_dt, forcing=None):
    """SWE2D.step with gradients to U, the level and (a trainable DEM correction) the bed; returns (U_new, dt)."""
    out, dt = _StepFn.apply(U, level, _bed_input(model), model, max_dt, forcing)
    return out, float(dt)


def tape_budget(model, U0):
    """Bytes the forward may spend on segment tapes (k1, k2 and every state).

    ``model.adjoint_tape``: "auto" (default) spends up to half of the device's
    free memory, True spends whatever it needs, False keeps checkpoints only,
    a number is a budget in bytes."""
    mode = getattr(model, "adjoint_tape", "auto")
    if mode is False or mode == "off":
        return 0
    if mode is True or mode == "on":
        return float("inf")
    if isinstance(mode, (int, float)):
        return float(mode)
    if mode != "auto":
        raise ValueError(f"adjoint_tape must be 'auto', True or False, got {mode!r}")
    free, _ = torch.cuda.mem_get_info(U0.device)
    return 0.5 * free


def _checkpointed_run(model, U0, level, duration, max_steps, cap, K, budget=0):
    """Device-driven forward (no per-step host sync) in segments of K steps.

    A segment records its full tape (states, k1, k2; 3K + 1 states) while the
    tapes fit in ``budget`` bytes; otherwise it keeps only its starting state,
    and the backward replays it.  Returns (U, hmax, dts, segments, done,
    elapsed); a segment is (kind, first, n_steps, data) with kind "tape"
    (data = (U slots, k1 slots, k2 slots)) or "ckpt" (data = starting state)."""
    lib = _lib.load()
    bed, rough = model._fields(U0)
    ws = model._workspace(U0, bed, rough)
    lvl, stride = model._level_arg(level, U0)
    stream = _lib.stream_ptr(U0.device)
    rows = model._rows(U0)
    cur = U0.clone()
    Ua, Ub = None, None
    hmax = U0[:, 0].clone()
    cap_arg = -1.0 if cap is None else float(cap)
    _lib.check(lib.swe_run_begin(_lib.ptr(ws["buf"]), float(duration), cap_arg, U0.element_size(), stream),
               "run_begin")
    advance = _fn(lib, "swe_advance", U0.dtype)
    advance_tape = _fn(lib, "swe_advance_tape", U0.dtype)
    state_bytes = U0.numel() * U0.element_size()
    capacity = max(K, 64)
    dts = torch.empty(capacity, dtype=U0.dtype, device=U0.device)
    segments = []
    used = 0
    launched, steps = 0, 0
    ctl = None
    while launched < max_steps:
        n = min(K, max_steps - launched)
        if launched + n > capacity:
            capacity = 2 * (launched + n)
            dts = torch.cat((dts, torch.empty(capacity - dts.numel(), dtype=dts.dtype, device=dts.device)))
        need = (3 * n + 1) * state_bytes
        if used + need <= budget:
            tape = torch.empty((3 * n + 1, *U0.shape), dtype=U0.dtype, device=U0.device)
            tape[0].copy_(cur)
            _lib.check(advance_tape(ws["grid"], ws["cfg"], _lib.ptr(bed), _lib.ptr(rough), _lib.ptr(lvl), stride,
                                    _lib.ptr(tape[0]), _lib.ptr(tape[n + 1]), _lib.ptr(tape[2 * n + 1]), launched,
                                    launched, n, _lib.ptr(hmax), _lib.ptr(dts), capacity, _lib.ptr(ws["buf"]), rows,
                                    stream), "advance_tape")
            segments.append(["tape", launched, n, tape])
            used += need
        else:
            if Ua is None:
                Ua, Ub = torch.empty_like(U0), torch.empty_like(U0)
            src = Ua if launched % 2 == 0 else Ub
            src.copy_(cur)
            segments.append(["ckpt", launched, n, cur.clone()])
            _lib.check(advance(ws["grid"], ws["cfg"], _lib.ptr(bed), _lib.ptr(rough), _lib.ptr(lvl), stride,
                               _lib.ptr(Ua), _lib.ptr(Ub), launched, n, _lib.ptr(hmax), _lib.ptr(dts), capacity, None,
                               0, None, None, 0, _lib.ptr(ws["buf"]), rows, stream), "advance")
        ctl = _lib.read_control(ws["buf"])
        steps = ctl.steps
        kind, first, _, data = segments[-1]
        done_here = max(0, min(steps, launched + n) - first)
        cur = data[done_here] if kind == "tape" else (Ua if steps % 2 == 0 else Ub)
        launched += n
        if ctl.done:
            break
    sync_parity(ws, ctl)
    if ctl.status != 0:
        if ctl.status == 1:
            raise FloatingPointError("The calculated timestep is non-finite.")
        raise FloatingPointError(f"Non-positive timestep: {ctl.bad_dt}.")
    U = cur.clone()
    dt_list = dts[:steps].cpu().tolist()
    kept = []
    for kind, first, n, data in segments:
        m = min(n, steps - first)
        if m <= 0:
            continue
        if kind == "tape":
            data = (data[:m], data[n + 1:n + 1 + m], data[2 * n + 1:2 * n + 1 + m])
        kept.append((kind, first, m, data))
    return U, hmax, dt_list, kept, bool(ctl.done), ctl.elapsed[steps & 1]


class _ForwardFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, U0, level, bed, model, duration, max_steps, cap):
        U0_ = U0.detach().contiguous()
        lvl = level.detach()
        ctx.bed_meta = _meta(bed)
        K = int(getattr(model, "adjoint_checkpoint_every", 64))
        if K < 1:
            raise ValueError("adjoint_checkpoint_every must be at least 1")
        budget = tape_budget(model, U0_)
        U, hmax, dts, segments, done, elapsed = _checkpointed_run(model, U0_, lvl, duration, max_steps, cap, K,
                                                                  budget)
        if not done:
            raise RuntimeError(
                f"max_steps={max_steps} reached after {elapsed:.8f} s; requested duration={duration:.8f} s.")
        ctx.model, ctx.dts, ctx.segments = model, dts, segments
        ctx.U_meta = (U0_.shape, U0_.dtype, U0_.device)
        ctx.save_for_backward(lvl)
        model.last_dt_series = dts
        model.last_adjoint_segments = [(kind, first, n) for kind, first, n, _ in segments]
        return U, hmax

    @staticmethod
    def backward(ctx, g_U, g_hmax):
        if g_hmax is not None and bool((g_hmax != 0).any()):
            raise NotImplementedError(
                "gradients through hmax of a device-driven run are not supported; set "
                "model.differentiable_hmax = True to run the per-step loop, whose hmax is differentiable")
        lvl, = ctx.saved_tensors
        model, dts, segments = ctx.model, ctx.dts, ctx.segments
        shape, dtype, device = ctx.U_meta
        lam = g_U.contiguous().clone() if g_U is not None else torch.zeros(shape, dtype=dtype, device=device)
        lam_level = torch.zeros(lam.shape[0], dtype=lam.dtype, device=lam.device)
        want_bed = ctx.needs_input_grad[2]
        region = None
        if segments:
            with _bed_gradient(model, lam, want_bed) as region:
                if getattr(model, "adjoint_verify_replay", False):
                    lam = _replay_checked(model, segments, lvl, dts, lam, lam_level)
                else:
                    lam = _reverse_segments(model, segments, lvl, dts, lam, lam_level)
        ctx.segments = None
        g_level = _level_grad(lvl, lam_level) if ctx.needs_input_grad[1] else None
        g_bed = None
        if want_bed:
            meta = ctx.bed_meta
            g_bed = _bed_grad(region, meta) if region is not None else torch.zeros(meta[0], dtype=meta[1],
                                                                                  device=meta[2])
        return lam, g_level, g_bed, None, None, None, None


def _reverse_segments(model, segments, lvl, dts, lam, lam_level):
    """Last segment first: a taped segment is swept in reverse at once; a
    checkpointed one is replayed onto a scratch tape and swept, in one call."""
    lib = _lib.load()
    U0 = segments[0][3][0][0] if segments[0][0] == "tape" else segments[0][3]  # a state of the run
    bed, rough = model._fields(U0)
    ws = model._workspace(U0, bed, rough)
    aws = _adj_ws(model, U0)
    l, stride = model._level_arg(lvl, U0)
    host_dts = (ctypes.c_double * len(dts))(*dts)
    stream = _lib.stream_ptr(U0.device)
    rows = int(model.rows_hint)
    lam_tmp = torch.empty_like(lam)
    scratch = None
    for kind, first, n, data in reversed(segments):
        dts_ptr = ctypes.addressof(host_dts) + 8 * first
        if kind == "tape":
            tU, tk1, tk2 = data
            _lib.check(_fn(lib, "swe_segment_reverse", U0.dtype)(
                ws["grid"], ws["cfg"], _lib.ptr(tU), _lib.ptr(tk1), _lib.ptr(tk2), _lib.ptr(bed), _lib.ptr(rough),
                _lib.ptr(l), stride, dts_ptr, n, _lib.ptr(lam), _lib.ptr(lam_tmp), _lib.ptr(lam_level),
                _lib.ptr(aws["buf"]), rows, stream), "segment_reverse", _ADJ_MSG)
        else:
            need = lib.swe_segment_tape_bytes(ws["grid"], n, U0.element_size()) // U0.element_size()
            if scratch is None or scratch.numel() < need:
                scratch = torch.empty(need, dtype=U0.dtype, device=U0.device)
            _lib.check(_fn(lib, "swe_segment_adjoint", U0.dtype)(
                ws["grid"], ws["cfg"], _lib.ptr(data), _lib.ptr(bed), _lib.ptr(rough), _lib.ptr(l), stride, dts_ptr,
                n, ws["parity"], _lib.ptr(lam), _lib.ptr(lam_tmp), _lib.ptr(lam_level), _lib.ptr(scratch),
                _lib.ptr(ws["buf"]), _lib.ptr(aws["buf"]), rows, stream), "segment_adjoint", _ADJ_MSG)
            ws["parity"] ^= n & 1
    return lam


def _replay_checked(model, segments, lvl, dts, lam, lam_level):
    """Step-by-step replay from each segment's first state, reading every dt back
    and checking it against the recorded one (taped segments are replayed too)."""
    for kind, first, n, data in reversed(segments):
        U = data[0][0] if kind == "tape" else data
        tape = []
        for k in range(first, first + n):
            U_next, k1, k2, dt = record_step(model, U, lvl, dts[k], read_dt=True)
            if dt != dts[k]:
                raise RuntimeError(f"adjoint replay diverged at step {k}: dt {dt} != {dts[k]}")
            tape.append((U, k1, k2))
            U = U_next
        for k in range(first + n - 1, first - 1, -1):
            Uk, k1, k2 = tape[k - first]
            lam = step_adjoint(model, Uk, k1, k2, lvl, dts[k], lam, lam_level)
        del tape
    return lam


def differentiable_forward(model, U0, duration, absolute_start, max_steps, boundary_level, cap):
    """SWE2D.forward with gradients to U0, an explicit boundary level and (a trainable DEM correction) the bed."""
    if boundary_level is None and model.use_dynamic_bc:
        raise NotImplementedError("gradients through a dynamic boundary series are not supported by the adjoint")
    level = model._prepare_boundary_level(
        model.cfg.constant_level if boundary_level is None else boundary_level, U0)
    model._check_state(U0, "U0")
    return _ForwardFn.apply(U0, level, _bed_input(model), model, float(duration), int(max_steps), cap)


class _OperatorFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, U, level, bed, model):
        U_ = U.detach().contiguous()
        lvl = level.detach()
        ctx.bed_meta = _meta(bed)
        with torch.no_grad():
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
        want_bed = ctx.needs_input_grad[2]
        with _bed_gradient(model, U, want_bed) as region:
            cfg = _lib.make_config(model.cfg)
            cfg.grad_bed = int(want_bed)
            _lib.check(_fn(lib, "swe_operator_adjoint", U.dtype)(
                ws["grid"], cfg, _lib.ptr(U), _lib.ptr(bed), _lib.ptr(l), stride, _lib.ptr(g.contiguous()),
                _lib.ptr(lam_U), _lib.ptr(lam_level), _lib.ptr(aws["buf"]), int(model.rows_hint),
                _lib.stream_ptr(U.device)), "operator_adjoint", _ADJ_MSG)
        g_level = _level_grad(lvl, lam_level) if ctx.needs_input_grad[1] else None
        g_bed = _bed_grad(region, ctx.bed_meta) if want_bed else None
        return lam_U, g_level, g_bed, None


def operator_adjoint(model, U, level):
    """spatial_operator(U, boundary_level=level) with gradients to U, the level and (a trainable DEM) the bed."""
    return _OperatorFn.apply(U, level, _bed_input(model), model)
