"""Tile-shape autotuner for the stage kernels.

The one free tile parameter of the forward kernels is S, the number of rows a
wavefront streams through (its strip is always 64 columns wide).  Small S
fills the GPU on small grids but pays the 4-row halo more often; large S
amortises the halo.  ``tune_rows`` times both stage launches with HIP events
for each candidate on the caller's own state and keeps the fastest; results
are cached per (device, grid, batch, dtype) and saved to ``tuned_rows.json``
next to this file, which ``SWE2D`` consults when its ``rows_hint`` is 0.

``tune_ensemble_rows`` does the same for an ensemble (csrc/swe_ensemble.hip,
``Ensemble``): one S per grid size, chosen by timing the whole ensemble over a
few global steps; ``Ensemble`` reads the result for every scenario that names
no ``rows`` of its own, and leaves the choice to the library's heuristic when
that mix of sizes was never tuned.
"""
from __future__ import annotations

import json
from pathlib import Path

import torch

from . import _lib

TABLE = Path(__file__).resolve().parent / "tuned_rows.json"
CANDIDATES = (4, 8, 16, 32, 48, 64, 128)
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


ENSEMBLE_CANDIDATES = (8, 16, 32, 64)


def ensemble_mix(scenarios):
    """{(ny, nx): number of scenarios of that size}."""
    mix = {}
    for sc in scenarios:
        shape = tuple(int(v) for v in sc["bed"].shape)
        mix[shape] = mix.get(shape, 0) + 1
    return mix


def _ens_key(device, mix, arithmetic="reference"):
    name = torch.cuda.get_device_name(device) if torch.cuda.is_available() else "cpu"
    sizes = ",".join(f"{ny}x{nx}:{c}" for (ny, nx), c in sorted(mix.items()))
    return f"{name}|ensemble|{sizes}|float32|{arithmetic}"


def lookup_ensemble(device, mix, arithmetic="reference") -> dict:
    """Tuned rows {(ny, nx): S} for an ensemble of this mix of grid sizes and arithmetic, or {}."""
    entry = _load().get(_ens_key(device, mix, arithmetic), {})
    return {tuple(int(v) for v in k.split("x")): int(S) for k, S in entry.get("rows", {}).items()}


def _time_ensemble(cfg, scenarios, rows, device, steps, arithmetic):
    from .ensemble import Ensemble

    ens = Ensemble(cfg, [dict(sc, rows=rows[tuple(int(v) for v in sc["bed"].shape)]) for sc in scenarios],
                   device, graph=False, arithmetic=arithmetic)
    stream = _lib.stream_ptr(device)
    ens._advance(0, 2, stream)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record()
    ens._advance(2, steps, stream)
    ev[1].record()
    torch.cuda.synchronize(device)
    return ev[0].elapsed_time(ev[1]) / steps


def tune_ensemble_rows(cfg, scenarios, device, candidates=ENSEMBLE_CANDIDATES, steps=32, save=True,
                       arithmetic="reference"):
    """Rows per strip for each grid size of an ensemble; returns ({(ny, nx): S}, report).

    The whole ensemble is timed as it will run (every scenario, one launch per
    stage), `steps` global steps per trial after two untimed ones, in a scratch
    ensemble built from copies of the inputs.  One pass of coordinate descent,
    largest share of the cells first: each size's S is varied over the
    candidates with the others held, and the fastest kept.  With ``save`` the
    result goes to the table under (device, mix of sizes, arithmetic)."""
    if steps % 2:
        raise ValueError("steps must be even (the ensemble's buffers alternate every step)")
    device = torch.device(device)
    mix = ensemble_mix(scenarios)
    shapes = sorted(mix, key=lambda sh: -sh[0] * sh[1] * mix[sh])
    current = {sh: max([c for c in candidates if c <= max(sh[0] // 8, min(candidates))] or [min(candidates)])
               for sh in shapes}
    timed = {}

    def trial(rows):
        key = tuple(rows[sh] for sh in shapes)
        if key not in timed:
            timed[key] = _time_ensemble(cfg, scenarios, rows, device, steps, arithmetic)
        return timed[key]

    best_ms = trial(current)
    for sh in shapes:
        for S in candidates:
            if S > sh[0] or S == current[sh]:
                continue
            rows = dict(current)
            rows[sh] = S
            ms = trial(rows)
            if ms < best_ms:
                best_ms, current = ms, rows
    report = {"rows": {f"{ny}x{nx}": S for (ny, nx), S in current.items()}, "ms_per_global_step": round(best_ms, 4),
              "trials": {",".join(f"{ny}x{nx}:{S}" for (ny, nx), S in zip(shapes, key)): round(ms, 4)
                         for key, ms in timed.items()}}
    if save:
        _load()[_ens_key(device, mix, arithmetic)] = report
        try:
            TABLE.write_text(json.dumps(_load(), indent=1, sort_keys=True))
        except OSError:
            pass
    return dict(current), report
