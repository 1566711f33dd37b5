"""Ensemble of forward scenarios with uneven grids and per-scenario dt.

Each scenario integrates ``n_chunks`` chunks of ``chunk`` seconds with a
piecewise-constant boundary level per chunk and samples h + z at its gauges at
every chunk end, exactly like ``SWE2D.forward(U, t_end=chunk, boundary_level=
levels[k])`` called chunk after chunk (src/solver.py:680-779 driven as in
Examples/inverse/.../bc_inversion_train.py:332-410).  All scenarios share one
launch sequence (three kernels per global step, see csrc/swe_ensemble.hip),
captured once into a HIP graph and replayed until every scenario is done.
"""
from __future__ import annotations

import ctypes
import struct
import time

import numpy as np
import torch

from . import _lib, autotune


class Ensemble:
    def __init__(self, cfg, scenarios, device, graph=True, steps_per_graph=64, arithmetic="reference"):
        """scenarios: dicts with bed [ny,nx], rough [ny,nx], dx, dy, U0 [3,ny,nx], levels [n_chunks],
        chunk (s), gauges [(iy, ix)], optional rows (rows per wavefront strip; without it the
        rows tuned for that mix of grid sizes and this arithmetic (``autotune.tune_ensemble_rows``),
        else the library's heuristic).

        arithmetic: "reference" (the reference's rounding; every scenario equals its batch-1
        ``SWE2D.forward`` run bit for bit) or "fast" (``SWE2D.arithmetic = "fast"``'s
        tolerance-bounded FP32 arithmetic; equal to batch-1 runs made with it)."""
        if arithmetic not in ("reference", "fast"):
            raise ValueError(f"arithmetic must be 'reference' or 'fast', not {arithmetic!r}")
        self.arithmetic = arithmetic
        if steps_per_graph % 2:
            raise ValueError("steps_per_graph must be even (buffers alternate every step)")
        self.lib = _lib.load()
        self.device = torch.device(device)
        self.cfg = _lib.make_config(cfg)
        self.n = len(scenarios)
        self.graph_enabled = graph
        self.spg = steps_per_graph
        self._keep = []
        self.rows = []  # rows per strip asked for each scenario (0: the library's heuristic)
        table = (_lib.SweScenario * self.n)()
        self.views = []
        tuned = autotune.lookup_ensemble(self.device, autotune.ensemble_mix(scenarios), arithmetic)
        f32 = dict(dtype=torch.float32, device=self.device)
        for i, sc in enumerate(scenarios):
            ny, nx = np.asarray(sc["bed"]).shape
            Ua = torch.as_tensor(np.asarray(sc["U0"]), **f32).reshape(3, ny, nx).contiguous()
            Ub = torch.empty_like(Ua)
            k1 = torch.empty_like(Ua)
            bed = torch.as_tensor(np.asarray(sc["bed"]), **f32).contiguous()
            rough = torch.as_tensor(np.asarray(sc["rough"]), **f32).contiguous()
            levels = torch.as_tensor(np.asarray(sc["levels"]), **f32).contiguous()
            gauges = torch.as_tensor(np.asarray(sc["gauges"], dtype=np.int32).reshape(-1, 2), dtype=torch.int32,
                                     device=self.device).contiguous()
            for iy, ix in gauges.cpu().tolist():
                if not (0 <= iy < ny and 0 <= ix < nx):
                    raise IndexError(f"scenario {i}: gauge {(iy, ix)} outside the {ny}x{nx} grid")
            gout = torch.full((len(levels), gauges.shape[0]), float("nan"), **f32)
            t = table[i]
            t.ny, t.nx, t.dx, t.dy = ny, nx, float(sc["dx"]), float(sc.get("dy", sc["dx"]))
            t.U_a, t.U_b, t.k1 = Ua.data_ptr(), Ub.data_ptr(), k1.data_ptr()
            t.bed, t.roughness, t.levels = bed.data_ptr(), rough.data_ptr(), levels.data_ptr()
            t.n_chunks, t.chunk_seconds = len(levels), float(sc["chunk"])
            t.gauges, t.n_gauges, t.gauge_out = gauges.data_ptr(), gauges.shape[0], gout.data_ptr()
            t.rows_hint = int(sc.get("rows", 0)) or tuned.get((ny, nx), 0)
            self.rows.append(int(t.rows_hint))
            self._keep += [Ua, Ub, k1, bed, rough, levels, gauges, gout]
            self.views.append(dict(Ua=Ua, Ub=Ub, gauge_out=gout))
        nbytes = self.lib.swe_ensemble_workspace_bytes(self.n)
        self.ws = torch.zeros(nbytes, dtype=torch.uint8, device=self.device)
        total = ctypes.c_longlong(0)
        stream = _lib.stream_ptr(self.device)
        rc = self.lib.swe_ensemble_init(ctypes.byref(self.cfg), self.n, table, _lib.ptr(self.ws), nbytes,
                                        ctypes.byref(total), stream)
        if rc != 0:
            raise RuntimeError(f"ensemble init failed: {self.lib.swe_ensemble_last_error_message().decode()}")
        self.total_waves = total.value
        self.steps = 0
        self._graph = None

    def _advance(self, first, count, stream):
        rc = self.lib.swe_ensemble_advance(_lib.ptr(self.ws), self.n, self.total_waves, first, count,
                                           int(self.arithmetic == "fast"), stream)
        if rc != 0:
            raise RuntimeError(f"ensemble advance failed: {self.lib.swe_ensemble_last_error_message().decode()}")

    def n_done(self) -> int:
        return int(self.ws[4:8].view(torch.int32).item())

    def run(self, max_steps=10_000_000, log=None, log_every=2048, timeline=None):
        """Advance until every scenario has finished its chunks; returns global steps taken.

        ``log`` (optional) receives a progress line every ``log_every`` global steps.
        ``timeline`` (a list, optional) receives (global steps, time.perf_counter(),
        scenarios done) after every replay; the done count is read back from the
        device, so each entry follows the replay's completion."""
        stream = _lib.stream_ptr(self.device)
        if self.graph_enabled and self._graph is None and self.steps % 2 == 0:
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self._advance(0, self.spg, _lib.stream_ptr(self.device))
            self._graph = g
        while self.steps < max_steps:
            if self._graph is not None:
                self._graph.replay()
                self.steps += self.spg
            else:
                self._advance(self.steps, self.spg, stream)
                self.steps += self.spg
            done = self.n_done()
            if timeline is not None:
                timeline.append((self.steps, time.perf_counter(), done))
            if done >= self.n:
                break
            if log and self.steps % log_every < self.spg:
                log(f"ensemble: {self.steps} global steps, {done}/{self.n} scenarios done")
        torch.cuda.synchronize(self.device)
        return self.steps

    def clocks(self):
        off = self.lib.swe_ensemble_clock_offset(self.n)
        size = self.lib.swe_ensemble_clock_bytes()
        raw = self.ws[off:off + size * self.n].cpu().numpy().tobytes()
        out = []
        for i in range(self.n):
            elapsed, chunk, done, steps, status, final_parity = struct.unpack_from("<diiiii", raw, i * size)
            out.append(dict(elapsed=elapsed, chunk=chunk, done=done, steps=steps, status=status,
                            final_parity=final_parity))
        return out

    def results(self):
        """Per scenario: gauge series [n_chunks, n_gauges], final state [3, ny, nx], own step count."""
        res = []
        for v, c in zip(self.views, self.clocks()):
            if c["status"] != 0:
                raise FloatingPointError(f"scenario failed with status {c['status']}")
            U = v["Ua"] if c["final_parity"] == 0 else v["Ub"]
            res.append(dict(gauges=v["gauge_out"], U=U, steps=c["steps"], done=bool(c["done"])))
        return res


def phase_rates(timeline, cells, steps):
    """Cell-update rates over a run's timeline, split at the global steps where a grid size finishes.

    cells[i], steps[i]: cells and own step count of scenario i.  Returns one dict per
    phase (global step range, scenarios still running, seconds, Gcell-updates/s)."""
    cells = np.asarray(cells, dtype=np.float64)
    steps = np.asarray(steps, dtype=np.int64)
    ends = sorted(set(int(s) for s in steps))
    t_prev, s_prev = timeline[0][1], timeline[0][0]
    out, bound = [], 0
    acc = dict(cells=0.0, seconds=0.0, start=s_prev)
    for s, t, _ in timeline[1:]:
        done_cells = float((cells * (np.minimum(steps, s) - np.minimum(steps, s_prev))).sum())
        acc["cells"] += done_cells
        acc["seconds"] += t - t_prev
        t_prev, s_prev = t, s
        while bound < len(ends) and s >= ends[bound]:
            bound += 1
            if acc["seconds"] > 0:
                out.append(dict(global_steps=(acc["start"], s), running=int((steps > acc["start"]).sum()),
                                seconds=round(acc["seconds"], 4),
                                gcell_per_s=round(acc["cells"] / acc["seconds"] / 1e9, 2)))
            acc = dict(cells=0.0, seconds=0.0, start=s)
    return out


if __name__ == "__main__":
    # Timer of the run-up ensemble (seeded scenarios over the three run-up grids):
    # construction, the graph-replayed run, and the throughput of each phase of
    # the run (all grid sizes running, then the finer ones alone).
    import argparse
    import json

    from . import scenarios
    from .warmstart import runup_config

    ap = argparse.ArgumentParser()
    ap.add_argument("--scenarios", type=int, default=256)
    ap.add_argument("--chunks", type=int, default=450)
    ap.add_argument("--tune", action="store_true", help="tune rows per grid size first (cached in the table)")
    ap.add_argument("--arithmetic", default="reference", choices=("reference", "fast"))
    a = ap.parse_args()
    dev = torch.device("cuda")
    sc, _ = scenarios.ensemble_scenarios(n=a.scenarios, n_chunks=a.chunks)
    tune_s, tuned = 0.0, None
    if a.tune:
        t0 = time.perf_counter()
        _, tuned = autotune.tune_ensemble_rows(runup_config(), sc, dev, arithmetic=a.arithmetic)
        tune_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    ens = Ensemble(runup_config(), sc, dev, arithmetic=a.arithmetic)
    torch.cuda.synchronize(dev)
    t1 = time.perf_counter()
    tl = [(0, t1, 0)]
    steps = ens.run(timeline=tl)
    t2 = time.perf_counter()
    res = ens.results()
    cells = [s["bed"].size for s in sc]
    own = [r["steps"] for r in res]
    total = float(np.dot(cells, own))
    print(json.dumps(dict(
        scenarios=a.scenarios, chunks=a.chunks, arithmetic=a.arithmetic, global_steps=steps, cell_updates=total,
        rows={f"{s['bed'].shape[0]}x{s['bed'].shape[1]}": r for s, r in zip(sc, ens.rows)},
        tune_seconds=round(tune_s, 2), tuning=tuned,
        construct_seconds=round(t1 - t0, 3), run_seconds=round(t2 - t1, 3),
        run_gcell_per_s=round(total / (t2 - t1) / 1e9, 2),
        with_construction_gcell_per_s=round(total / (t2 - t0) / 1e9, 2),
        phases=phase_rates(tl, cells, own))))
