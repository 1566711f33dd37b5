"""Tile-shape autotuner for the stage kernels.

The one free tile parameter of the forward kernels is S, the number of rows a
wavefront streams through (its strip is always 64 columns wide).  Small S
fills the GPU on small grids but pays the 4-row halo more often; large S
amortises the halo.  ``tune_rows`` times both stage launches with HIP events
for each candidate on the caller's own state and keeps the fastest; results
are cached per (device, grid, batch, dtype) and saved to ``tuned_rows.json``
next to this file, which ``SWE2D`` and the ensemble consult when their
``rows_hint`` is 0.
"""
from __future__ import annotations

import json
import os
from pathlib import Path

import torch

from . import _lib

TABLE = Path(__file__).resolve().parent / "tuned_rows.json"
CANDIDATES = (4, 8, 16, 32, 64, 128)
_cache = None


def _key(device, ny, nx, batch, dtype):
    name = torch.cuda.get_device_name(device) if torch.cuda.is_available() else "cpu"
    return f"{name}|{ny}x{nx}|B{batch}|{str(dtype).replace('torch.', '')}"


def _load():
    global _cache
    if _cache is None:
        _cache = json.loads(TABLE.read_text()) if TABLE.exists() else {}
    return _cache


def lookup(device, ny, nx, batch, dtype) -> int:
    """Tuned rows for this grid, or 0 (library heuristic) when not tuned."""
    return int(_load().get(_key(device, ny, nx, batch, dtype), {}).get("rows", 0))


def tune_rows(model, U, candidates=CANDIDATES, reps=8, save=True):
    """Time stage 1 + stage 2 for each S; set model.rows_hint to the fastest and return a report."""
    if U.dtype != torch.float32:
        raise TypeError("the tuner times the FP32 stage kernels")
    lib = _lib.load()
    U = U.detach().contiguous()
    bed, rough = model._fields(U)
    ws = model._workspace(U, bed, rough)
    lvl, stride = model._level_arg(model.cfg.constant_level, U)
    out = torch.empty_like(U)
    stream = _lib.stream_ptr(U.device)
    ny = U.shape[-2]
    report = {}
    for S in candidates:
        if S > ny:
            continue
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        for it in range(reps + 2):
            if it == 2:
                ev[0].record()
            _lib.check(lib.swe_step_f32(ws["grid"], ws["cfg"], _lib.ptr(U), _lib.ptr(bed), _lib.ptr(rough),
                                        _lib.ptr(lvl), stride, 1e30, ws["parity"], _lib.ptr(out),
                                        _lib.ptr(ws["buf"]), int(S), stream), "tune step")
            ws["parity"] ^= 1
        ev[1].record()
        torch.cuda.synchronize(U.device)
        report[S] = ev[0].elapsed_time(ev[1]) / reps
    best = min(report, key=report.get)
    model.rows_hint = int(best)
    if save:
        table = _load()
        table[_key(U.device, *U.shape[-2:], U.shape[0], U.dtype)] = {
            "rows": int(best), "ms_per_step": {str(k): round(v, 4) for k, v in report.items()}}
        try:
            TABLE.write_text(json.dumps(table, indent=1, sort_keys=True))
        except OSError:
            pass
    return best, report
