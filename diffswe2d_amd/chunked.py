"""Whole chunked record in one device-driven run, with its reverse mode.

The reference's inversion driver advances the model chunk by chunk
(Examples/inverse/.../bc_inversion_train.py:332-349: ``forward(U,
t_end=chunk, boundary_level=series[c])`` per chunk) and observes h + z at the
gauges after every chunk (:401-410).  Driven from the host that costs one
launch sequence and two synchronisations per chunk; here the chunk clock lives
on the device (csrc/swe_step.hip ``chunk_clock_kernel``): the run only returns
to the host every few hundred steps.  The result equals the chunk-by-chunk
loop bit for bit (same steps, same dt, same levels).

``gauge_run(model, U0, series, chunk, gauges)`` returns (obs [n_chunks,
n_gauges] of h + z, final state); with gradients enabled it records tapes (or
checkpoints, as ``SWE2D.forward`` does) and the backward sweeps the run in
reverse, adding the observation adjoints at chunk ends and collecting the
level adjoint of every chunk (csrc/swe_adjoint.hip, ``*_chunked_*``).
"""
from __future__ import annotations

import ctypes

import torch

from . import _lib
from .adjoint import _ADJ_MSG, _adj_ws, _bed_grad, _bed_gradient, _bed_input, _fn, _meta, tape_budget
from .solver import sync_parity

_TAPE_SEGMENT = 512     # steps per recorded segment when the tape fits
_PLAIN_SEGMENT = 1024   # steps per launch batch without gradients


def _run(model, U0, levels, chunk, gauges, max_steps, K, budget):
    """Device-driven chunked run; returns (obs, U, dts, chunks, segments)."""
    lib = _lib.load()
    n_chunks = levels.numel()
    bed, rough = model._fields(U0)
    ws = model._workspace(U0, bed, rough)
    stream = _lib.stream_ptr(U0.device)
    rows = model._rows(U0)
    esize = U0.element_size()
    _lib.check(lib.swe_run_begin_chunked(_lib.ptr(ws["buf"]), float(chunk), n_chunks, _lib.ptr(levels), esize,
                                         stream), "run_begin_chunked")
    advance = _fn(lib, "swe_advance_chunked", U0.dtype)
    n_g = gauges.shape[0]
    obs = torch.full((n_chunks, max(n_g, 1)), float("nan"), dtype=U0.dtype, device=U0.device)
    capacity = 1024
    dts = torch.empty(capacity, dtype=U0.dtype, device=U0.device)
    chunks = torch.empty(capacity, dtype=torch.int32, device=U0.device)
    state_bytes = U0.numel() * esize
    segments, used = [], 0
    cur = U0.clone()
    Ua = Ub = None
    launched, ctl = 0, None
    while launched < max_steps:
        kind, n = "ckpt", min(K, max_steps - launched)
        if budget is None:
            kind, n = "plain", min(_PLAIN_SEGMENT, max_steps - launched)
        else:
            for m in (min(_TAPE_SEGMENT, max_steps - launched), n):
                if used + (3 * m + 1) * state_bytes <= budget:
                    kind, n = "tape", m
                    break
        if launched + n > capacity:
            capacity = max(2 * capacity, launched + n)
            dts = torch.cat((dts, torch.empty(capacity - dts.numel(), dtype=dts.dtype, device=dts.device)))
            chunks = torch.cat((chunks, torch.empty(capacity - chunks.numel(), dtype=chunks.dtype,
                                                    device=chunks.device)))
        if kind == "tape":
            tape = torch.empty((3 * n + 1, *U0.shape), dtype=U0.dtype, device=U0.device)
            tape[0].copy_(cur)
            _lib.check(advance(ws["grid"], ws["cfg"], _lib.ptr(bed), _lib.ptr(rough), _lib.ptr(tape[0]), None,
                               _lib.ptr(tape[n + 1]), _lib.ptr(tape[2 * n + 1]), launched, launched, n,
                               _lib.ptr(dts), _lib.ptr(chunks), capacity, _lib.ptr(gauges), n_g, _lib.ptr(obs),
                               _lib.ptr(ws["buf"]), rows, stream), "advance_chunked")
            segments.append(["tape", launched, n, tape])
            used += (3 * n + 1) * state_bytes
        else:
            if Ua is None:
                Ua, Ub = torch.empty_like(U0), torch.empty_like(U0)
            src = Ua if launched % 2 == 0 else Ub
            src.copy_(cur)
            if kind == "ckpt":
                segments.append(["ckpt", launched, n, cur.clone()])
            _lib.check(advance(ws["grid"], ws["cfg"], _lib.ptr(bed), _lib.ptr(rough), _lib.ptr(Ua), _lib.ptr(Ub),
                               None, None, 0, launched, n, _lib.ptr(dts), _lib.ptr(chunks), capacity,
                               _lib.ptr(gauges), n_g, _lib.ptr(obs), _lib.ptr(ws["buf"]), rows, stream),
                       "advance_chunked")
        ctl = _lib.read_control(ws["buf"])
        steps = ctl.steps
        if kind == "tape":
            cur = segments[-1][3][max(0, min(steps, launched + n) - launched)]
        else:
            cur = Ua if steps % 2 == 0 else Ub
        launched += n
        if ctl.done:
            break
    sync_parity(ws, ctl)
    if ctl.status != 0:
        if ctl.status == 1:
            raise FloatingPointError("The calculated timestep is non-finite.")
        raise FloatingPointError(f"Non-positive timestep: {ctl.bad_dt}.")
    if not ctl.done:
        raise RuntimeError(f"max_steps={max_steps} reached in chunk {ctl.chunk} of {n_chunks}.")
    steps = ctl.steps
    kept = []
    for kind, first, n, data in segments:
        m = min(n, steps - first)
        if m <= 0:
            continue
        if kind == "tape":
            data = (data[:m], data[n + 1:n + 1 + m], data[2 * n + 1:2 * n + 1 + m])
        kept.append((kind, first, m, data))
    return obs[:, :n_g], cur.clone(), dts[:steps], chunks[:steps], kept


class _ChunkedFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, U0, levels, bed, model, chunk, gauges, max_steps):
        U0_ = U0.detach().contiguous()
        lv = levels.detach().contiguous()
        ctx.bed_meta = _meta(bed)
        K = int(getattr(model, "adjoint_checkpoint_every", 64))
        if K < 1:
            raise ValueError("adjoint_checkpoint_every must be at least 1")
        obs, U, dts, chunks, segments = _run(model, U0_, lv, chunk, gauges, max_steps, K, tape_budget(model, U0_))
        ctx.model, ctx.segments, ctx.gauges = model, segments, gauges
        ctx.dts = dts.cpu().tolist()
        ctx.chunks = chunks.cpu().tolist()
        ctx.save_for_backward(lv)
        model.last_dt_series = ctx.dts
        model.last_adjoint_segments = [(kind, first, n) for kind, first, n, _ in segments]
        return obs, U

    @staticmethod
    def backward(ctx, g_obs, g_U):
        lv, = ctx.saved_tensors
        model, segments, gauges = ctx.model, ctx.segments, ctx.gauges
        lib = _lib.load()
        some = segments[0][3]
        U0 = some[0][0] if segments[0][0] == "tape" else some
        dtype, dev = U0.dtype, U0.device
        want_bed = ctx.needs_input_grad[2]
        with _bed_gradient(model, U0, want_bed) as region:
            bed, rough = model._fields(U0)
            ws = model._workspace(U0, bed, rough)
            aws = _adj_ws(model, U0)
            stream = _lib.stream_ptr(dev)
            rows = int(model.rows_hint)
            lam = g_U.contiguous().clone() if g_U is not None else torch.zeros_like(U0)
            lam_tmp = torch.empty_like(lam)
            level_grad = torch.zeros(lv.numel(), dtype=dtype, device=dev)
            n_g = gauges.shape[0]
            g_obs = (g_obs.contiguous() if g_obs is not None
                     else torch.zeros(lv.numel(), max(n_g, 1), dtype=dtype, device=dev))
            n = len(ctx.dts)
            chunk = ctx.chunks
            ends = [1 if k == n - 1 or chunk[k + 1] != chunk[k] else 0 for k in range(n)]
            host_dts = (ctypes.c_double * n)(*ctx.dts)
            host_chunk = (ctypes.c_int32 * n)(*chunk)
            host_ends = (ctypes.c_int32 * n)(*ends)
            scratch = None
            for kind, first, m, data in reversed(segments):
                dptr = ctypes.addressof(host_dts) + 8 * first
                cptr = ctypes.addressof(host_chunk) + 4 * first
                eptr = ctypes.addressof(host_ends) + 4 * first
                if kind == "tape":
                    tU, tk1, tk2 = data
                    _lib.check(_fn(lib, "swe_segment_reverse_chunked", dtype)(
                        ws["grid"], ws["cfg"], _lib.ptr(tU), _lib.ptr(tk1), _lib.ptr(tk2), _lib.ptr(bed),
                        _lib.ptr(rough), _lib.ptr(lv), dptr, cptr, eptr, m, _lib.ptr(gauges), n_g, _lib.ptr(g_obs),
                        _lib.ptr(lam), _lib.ptr(lam_tmp), _lib.ptr(level_grad), _lib.ptr(aws["buf"]), rows, stream),
                        "segment_reverse_chunked", _ADJ_MSG)
                else:
                    need = lib.swe_segment_tape_bytes(ws["grid"], m, U0.element_size()) // U0.element_size()
                    if scratch is None or scratch.numel() < need:
                        scratch = torch.empty(need, dtype=dtype, device=dev)
                    _lib.check(_fn(lib, "swe_segment_adjoint_chunked", dtype)(
                        ws["grid"], ws["cfg"], _lib.ptr(data), _lib.ptr(bed), _lib.ptr(rough), _lib.ptr(lv), dptr,
                        cptr, eptr, m, ws["parity"], _lib.ptr(gauges), n_g, _lib.ptr(g_obs), _lib.ptr(lam),
                        _lib.ptr(lam_tmp), _lib.ptr(level_grad), _lib.ptr(scratch), _lib.ptr(ws["buf"]),
                        _lib.ptr(aws["buf"]), rows, stream), "segment_adjoint_chunked", _ADJ_MSG)
                    ws["parity"] ^= m & 1
        ctx.segments = None
        g_bed = None
        if want_bed:
            g_bed = _bed_grad(region, ctx.bed_meta)
            if g_obs is not None and n_g > 0:
                # observations are h + z at the gauge cells: their adjoint reaches the bed directly
                gsum = g_obs[:, :n_g].sum(dim=0).to(g_bed.dtype)
                gi = gauges.long().to(g_bed.device)
                g_bed[0, 0].index_put_((gi[:, 0], gi[:, 1]), gsum.to(g_bed.device), accumulate=True)
        return lam, level_grad, g_bed, None, None, None, None


def gauge_run(model, U0, series, chunk, gauges, max_steps=10_000_000):
    """Run len(series) chunks of `chunk` seconds with level series[c] in chunk c.

    Returns (obs [n_chunks, n_gauges] = h + z at the gauge cells of member 0 at
    each chunk end, final state).  U0 must be [1, 3, ny, nx]; gradients flow to
    U0 and series when either requires them."""
    if U0.dim() != 4 or U0.shape[0] != 1:
        raise ValueError("gauge_run takes a single member U0 of shape [1, 3, ny, nx]")
    if model.use_dynamic_bc:
        raise NotImplementedError("gauge_run uses its explicit per-chunk levels, not a dynamic boundary series")
    if model.rainfall is not None:
        raise NotImplementedError("the chunked device run does not carry the rainfall record; "
                                  "run the chunks through SWE2D.forward")
    model._check_state(U0, "U0")
    ny, nx = U0.shape[-2:]
    g = torch.as_tensor(gauges, dtype=torch.int32).reshape(-1, 2)
    if g.numel() and not (((g[:, 0] >= 0) & (g[:, 0] < ny) & (g[:, 1] >= 0) & (g[:, 1] < nx)).all()):
        raise IndexError(f"a gauge lies outside the {ny}x{nx} grid")
    if g.shape[0] > 1024:
        raise ValueError("at most 1024 gauges")
    g = g.to(U0.device).contiguous()
    levels = series.reshape(-1).to(dtype=U0.dtype)
    if not (chunk > 0):
        raise ValueError("chunk duration must be positive")
    if torch.is_grad_enabled() and (U0.requires_grad or levels.requires_grad or model._bed_trainable()):
        return _ChunkedFn.apply(U0, levels, _bed_input(model), model, float(chunk), g, int(max_steps))
    with torch.no_grad():
        obs, U, dts, _, _ = _run(model, U0.detach().contiguous(), levels.detach().contiguous(), float(chunk), g,
                                 int(max_steps), 0, None)
    model.last_dt_series = dts.cpu().tolist()
    return obs, U


# ---------------------------------------------------------------------------
# dynamic boundary series on the device
# ---------------------------------------------------------------------------
def dynamic_series(model, U):
    """The model's dynamic series (bc_times, bc_wls) as contiguous device arrays of U's dtype
    (src/solver.py:402-403: ``times/levels.to(dtype=reference.dtype, device=reference.device)``)."""
    times = model.bc_times.detach().to(dtype=U.dtype, device=U.device).contiguous()
    levels = model.bc_wls.detach().to(dtype=U.dtype, device=U.device).contiguous()
    return times, levels


def dynamic_arg(times, levels, start, rec=None):
    d = _lib.SweDynamicLevel()
    d.times, d.levels, d.n, d.start_time = times.data_ptr(), levels.data_ptr(), times.numel(), float(start)
    if rec is not None:
        d.step_level, d.step_index, d.step_weight = (rec[0].data_ptr(), rec[1].data_ptr(), rec[2].data_ptr())
    return d


def _dyn_run(model, U0, times, levels, duration, start, max_steps, cap, K, budget):
    """Device-driven run with the dynamic series, in segments of K steps that record
    their tape while it fits in ``budget`` (else keep their first state); also
    records each step's level, interval index and weight.  Returns (U, hmax, dts,
    (level, index, weight) per step, segments)."""
    lib = _lib.load()
    bed, rough = model._fields(U0)
    ws = model._workspace(U0, bed, rough)
    stream = _lib.stream_ptr(U0.device)
    rows = model._rows(U0)
    _lib.check(lib.swe_run_begin(_lib.ptr(ws["buf"]), float(duration), -1.0 if cap is None else float(cap),
                                 U0.element_size(), stream), "run_begin")
    advance = _fn(lib, "swe_advance_dynamic", U0.dtype)
    capacity = max(K, 64)

    def grow(t, dtype):
        return torch.empty(capacity, dtype=dtype, device=U0.device) if t is None else torch.cat(
            (t, torch.empty(capacity - t.numel(), dtype=t.dtype, device=t.device)))

    dts = grow(None, U0.dtype)
    rec = [grow(None, U0.dtype), grow(None, torch.int32), grow(None, U0.dtype)]
    hmax = U0[:, 0].clone()
    state_bytes = U0.numel() * U0.element_size()
    segments, used = [], 0
    cur = U0.clone()
    Ua = Ub = None
    launched, ctl = 0, None
    while launched < max_steps:
        n = min(K, max_steps - launched)
        if launched + n > capacity:
            capacity = 2 * (launched + n)
            dts = grow(dts, None)
            rec = [grow(r, None) for r in rec]
        dyn = dynamic_arg(times, levels, start, rec)
        need = (3 * n + 1) * state_bytes
        if used + need <= budget:
            tape = torch.empty((3 * n + 1, *U0.shape), dtype=U0.dtype, device=U0.device)
            tape[0].copy_(cur)
            _lib.check(advance(ws["grid"], ws["cfg"], ctypes.byref(dyn), None, _lib.ptr(bed), _lib.ptr(rough),
                               _lib.ptr(tape[0]), None, _lib.ptr(tape[n + 1]), _lib.ptr(tape[2 * n + 1]), launched,
                               launched, n, _lib.ptr(hmax), _lib.ptr(dts), capacity, None, 0, None, None,
                               _lib.ptr(ws["buf"]), rows, stream), "advance_dynamic")
            segments.append(["tape", launched, n, tape])
            used += need
        else:
            if Ua is None:
                Ua, Ub = torch.empty_like(U0), torch.empty_like(U0)
            (Ua if launched % 2 == 0 else Ub).copy_(cur)
            segments.append(["ckpt", launched, n, cur.clone()])
            _lib.check(advance(ws["grid"], ws["cfg"], ctypes.byref(dyn), None, _lib.ptr(bed), _lib.ptr(rough),
                               _lib.ptr(Ua), _lib.ptr(Ub), None, None, 0, launched, n, _lib.ptr(hmax), _lib.ptr(dts),
                               capacity, None, 0, None, None, _lib.ptr(ws["buf"]), rows, stream), "advance_dynamic")
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
    if not ctl.done:
        elapsed = ctl.elapsed[ctl.steps & 1]
        raise RuntimeError(
            f"max_steps={max_steps} reached after {elapsed:.8f} s; requested duration={duration:.8f} s.")
    steps = ctl.steps
    kept = []
    for kind, first, n, data in segments:
        m = min(n, steps - first)
        if m <= 0:
            continue
        if kind == "tape":
            data = (data[:m], data[n + 1:n + 1 + m], data[2 * n + 1:2 * n + 1 + m])
        kept.append((kind, first, m, data))
    return cur.clone(), hmax, dts[:steps], [r[:steps] for r in rec], kept


class _DynamicFn(torch.autograd.Function):
    """SWE2D.forward with a dynamic boundary series, device-driven, differentiable
    with respect to U0, the series' water levels and a trainable bed."""

    @staticmethod
    def forward(ctx, U0, wls, bed, model, duration, start, max_steps, cap):
        U0_ = U0.detach().contiguous()
        ctx.bed_meta = _meta(bed)
        K = int(getattr(model, "adjoint_checkpoint_every", 64))
        if K < 1:
            raise ValueError("adjoint_checkpoint_every must be at least 1")
        times, levels = dynamic_series(model, U0_)
        U, hmax, dts, rec, segments = _dyn_run(model, U0_, times, levels, duration, start, max_steps, cap, K,
                                               tape_budget(model, U0_))
        ctx.model, ctx.segments, ctx.rec = model, segments, rec
        ctx.dts = dts.cpu().tolist()
        ctx.wls_meta = (wls.shape, wls.dtype, wls.device)
        model.last_dt_series = dts
        model.last_adjoint_segments = [(kind, first, n) for kind, first, n, _ in segments]
        return U, hmax

    @staticmethod
    def backward(ctx, g_U, g_hmax):
        if g_hmax is not None and bool((g_hmax != 0).any()):
            raise NotImplementedError(
                "gradients through hmax of a device-driven run are not supported; set "
                "model.differentiable_hmax = True to run the per-step loop, whose hmax is differentiable")
        model, segments = ctx.model, ctx.segments
        step_level, step_index, step_weight = ctx.rec
        lib = _lib.load()
        some = segments[0][3]
        U0 = some[0][0] if segments[0][0] == "tape" else some
        dtype, dev = U0.dtype, U0.device
        want_bed = ctx.needs_input_grad[2]
        n = len(ctx.dts)
        level_grad = torch.zeros(n, dtype=dtype, device=dev)
        with _bed_gradient(model, U0, want_bed) as region:
            bed, rough = model._fields(U0)
            ws = model._workspace(U0, bed, rough)
            aws = _adj_ws(model, U0)
            stream = _lib.stream_ptr(dev)
            rows = int(model.rows_hint)
            lam = g_U.contiguous().clone() if g_U is not None else torch.zeros_like(U0)
            lam_tmp = torch.empty_like(lam)
            no_gauges = torch.zeros((0, 2), dtype=torch.int32, device=dev)
            g_obs = torch.zeros(1, dtype=dtype, device=dev)
            # every step is its own "chunk": its level is step_level[k] and its level
            # adjoint lands in level_grad[k]; no step ends an observed chunk
            host_dts = (ctypes.c_double * n)(*ctx.dts)
            host_chunk = (ctypes.c_int32 * n)(*range(n))
            host_ends = (ctypes.c_int32 * n)()
            scratch = None
            for kind, first, m, data in reversed(segments):
                dptr = ctypes.addressof(host_dts) + 8 * first
                cptr = ctypes.addressof(host_chunk) + 4 * first
                eptr = ctypes.addressof(host_ends) + 4 * first
                if kind == "tape":
                    tU, tk1, tk2 = data
                    _lib.check(_fn(lib, "swe_segment_reverse_chunked", dtype)(
                        ws["grid"], ws["cfg"], _lib.ptr(tU), _lib.ptr(tk1), _lib.ptr(tk2), _lib.ptr(bed),
                        _lib.ptr(rough), _lib.ptr(step_level), dptr, cptr, eptr, m, _lib.ptr(no_gauges), 0,
                        _lib.ptr(g_obs), _lib.ptr(lam), _lib.ptr(lam_tmp), _lib.ptr(level_grad), _lib.ptr(aws["buf"]),
                        rows, stream), "segment_reverse_chunked", _ADJ_MSG)
                else:
                    need = lib.swe_segment_tape_bytes(ws["grid"], m, U0.element_size()) // U0.element_size()
                    if scratch is None or scratch.numel() < need:
                        scratch = torch.empty(need, dtype=dtype, device=dev)
                    _lib.check(_fn(lib, "swe_segment_adjoint_chunked", dtype)(
                        ws["grid"], ws["cfg"], _lib.ptr(data), _lib.ptr(bed), _lib.ptr(rough), _lib.ptr(step_level),
                        dptr, cptr, eptr, m, ws["parity"], _lib.ptr(no_gauges), 0, _lib.ptr(g_obs), _lib.ptr(lam),
                        _lib.ptr(lam_tmp), _lib.ptr(level_grad), _lib.ptr(scratch), _lib.ptr(ws["buf"]),
                        _lib.ptr(aws["buf"]), rows, stream), "segment_adjoint_chunked", _ADJ_MSG)
                    ws["parity"] ^= m & 1
        ctx.segments = None
        g_wls = None
        if ctx.needs_input_grad[1]:
            # level_k = w0 + weight_k * (w1 - w0) with w0 = wls[i_k - 1], w1 = wls[i_k]
            # (solver.py:409-415): d/dw0 = g - g * weight, d/dw1 = g * weight
            shape, wdtype, wdev = ctx.wls_meta
            nw = shape[0]
            i1 = step_index.long()
            i0 = torch.where(i1 > 0, i1 - 1, torch.full_like(i1, nw - 1))
            gw = level_grad * step_weight
            g = torch.zeros(nw, dtype=dtype, device=dev)
            g.index_add_(0, i0, level_grad - gw)
            g.index_add_(0, i1, gw)
            g_wls = g.to(dtype=wdtype, device=wdev).reshape(shape)
        g_bed = _bed_grad(region, ctx.bed_meta) if want_bed else None
        return lam, g_wls, g_bed, None, None, None, None, None


def dynamic_forward(model, U0, duration, start, max_steps, cap):
    """Differentiable device-driven forward with the model's dynamic boundary series."""
    return _DynamicFn.apply(U0, model.bc_wls, _bed_input(model), model, float(duration), float(start),
                            int(max_steps), cap)
