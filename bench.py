"""Benchmark of the HIP shallow-water time step on one MI355X.

Workload: the 8192 x 8192 tsunami basin (dx = 100 m, fractal bathymetry,
~15 % land, 2 m Gaussian hump, Manning n = 0.025, transmissive sides), FP32.
A step is one adaptive midpoint step (two fused launches) over the whole grid.

    python bench.py [--gpus 1] [--steps 2000] [--warmup 20]

Prints one JSON line.  ``value`` is cell-updates per second over the K timed
steps (inputs resident in HBM, device-driven loop, no host sync inside).
``roofline`` prices the dominant kernel with the bytes it must move; a second
figure prices the whole step at 32 B per cell-update (read h, hu, hv, z, n and
write h, hu, hv once).  ``inversion`` runs the reference's Monai Valley
boundary inversion to its early stop and sets its time to the best misfit
beside the published run's.  ``warmstart`` runs the coastal run-up inversion
plainly and surrogate-started and compares their times to the plain run's best
misfit.  ``cpu_baseline`` times the CPU oracle (an eager PyTorch restatement of
the reference step) on the full basin on the host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
MIX_PROBE_BLOCKS, MIX_PROBE_ITERS = 8192, 512  # the VALU mix probe launch timed (and profiled) by bench.py
METRIC = "Gcell-updates/s fwd+adj FP32 & %HBM peak; inversion time-to-misfit vs plain"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=8192, help="grid side (default: the 8192^2 basin)")
    ap.add_argument("--rows", type=int, default=0, help="rows per wavefront strip (0 = autotuned / heuristic)")
    ap.add_argument("--kernel-steps", type=int, default=50, help="steps timed launch by launch with events")
    ap.add_argument("--cpu-steps", type=int, default=2, help="timed oracle steps of the full basin (after 1 warm-up)")
    ap.add_argument("--warmstart-scenarios", type=int, default=256,
                    help="ensemble size of the warm-start leg (the run-up inversion, plain vs surrogate-started); "
                         "0 to skip")
    ap.add_argument("--warmstart-epochs", type=int, default=40, help="adjoint epochs of the plain inversion")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-small", action="store_true", help="skip the dam-break and run-up workloads")
    ap.add_argument("--small-cpu-steps", type=int, default=40,
                    help="timed oracle steps of each smaller workload (after 1 warm-up)")
    ap.add_argument("--no-fast", action="store_true",
                    help="skip the leg with the library's tolerance-bounded FP32 arithmetic")
    ap.add_argument("--adjoint-steps", type=int, default=10,
                    help="steps of the adjoint leg (taped forward, reverse sweep, checkpointed backward); 0 to skip")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run a 1 GiB device copy (calibrates rocprof FETCH_SIZE / WRITE_SIZE)")
    ap.add_argument("--inversion-epochs", type=int, default=500,
                    help="epoch limit of the inversion leg (the reference's Monai Valley boundary inversion, which "
                         "stops early like the example does); 0 to skip")
    ap.add_argument("--small-grids", action="store_true",
                    help="instead of the basin: time forward steps at the reference examples' grid sizes")
    return ap.parse_args()


def inversion_leg(dev, max_epochs):
    """The reference's Monai Valley boundary inversion (68 x 42, FP64, 149 chunks, 20 spline controls,
    laboratory gauges), run to its early stop with the device-driven chunked run and the hand-written
    adjoint; time to the best training misfit against the time the published run took to the same
    epoch (its loss_history.txt, kept with the test fixtures)."""
    import numpy as np

    from diffswe2d_amd import inversion

    golden = ROOT / "tests" / "golden"
    d = inversion.monai_valley_setup(golden / "monai_bathy.nc", golden / "monai_wave_gauges_02s.txt", dev)
    _, best_epoch, hist = inversion.spline_inversion(d["problem"], d["basis"], d["target"], d["raw0"], train=(0, 1),
                                                     test=(2,), bounds=d["bounds"], max_epochs=max_epochs)
    pub = np.loadtxt(golden / "monai_inversion_loss_history.txt", skiprows=1)  # epoch, ..., seconds last
    best, pub_best = float("inf"), 0
    for e, loss in enumerate(pub[:, 2]):  # the example's rule: an improvement beyond 1e-8
        if loss < best - 1.0e-8:
            best, pub_best = loss, e
    ours = float(sum(r["elapsed_seconds"] for r in hist[:best_epoch]))
    theirs = float(pub[:pub_best + 1, -1].sum())
    return {"workload": "Monai Valley boundary inversion (the reference's example): 68x42 FP64, 149 chunks of "
                        "0.2 s, 20 spline controls, laboratory gauges",
            "epochs": len(hist), "best_epoch": best_epoch,
            "best_train_misfit": hist[best_epoch - 1]["train_loss"],
            "seconds_to_best": round(ours, 3),
            "seconds_per_epoch": round(float(np.median([r["elapsed_seconds"] for r in hist])), 4),
            "published": {"hardware": "A100 (its log)", "epochs": int(pub.shape[0]), "best_epoch": pub_best + 1,
                          "best_train_misfit": float(pub[pub_best, 2]), "seconds_to_best": round(theirs, 1)},
            "speedup_vs_published": round(theirs / ours, 1)}


SMALL_GRIDS = ((42, 68, "f64", "Monai Valley 0.08 m size"), (254, 200, "f64", "CEA 2008 size"),
               (344, 544, "f32", "coastal run-up"), (1024, 1024, "f32", "1024 x 1024"))


def small_grids(dev):
    """Forward steps of the run-up scenario at small grid sizes, device-driven loop, event-timed.

    At these sizes a launch holds a few dozen to a few thousand wavefronts, so a
    step costs about the latency of one wavefront's strip, not bandwidth."""
    import torch

    import diffswe2d_amd as swe
    from diffswe2d_amd import scenarios

    rows = []
    for ny, nx, name, label in SMALL_GRIDS:
        dtype = torch.float64 if name == "f64" else torch.float32
        dx = 5.44 / nx
        sc = scenarios.coastal_runup(nx=nx, ny=ny, dx=dx)
        cfg = swe.SWEConfig(boundary_left="water_level", boundary_right="wall", boundary_top="wall",
                            boundary_bottom="wall")
        m = swe.SWE2D(torch.tensor(sc["bed"], dtype=dtype), torch.tensor(sc["rough"], dtype=dtype), dx, dx,
                      config=cfg, train_dem=False).to(dev)
        U = torch.tensor(sc["U"], dtype=dtype, device=dev)
        with torch.no_grad():
            m(U, t_end=3.0, boundary_level=0.01)  # same length: first-use kernel loads happen here
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            torch.cuda.synchronize()
            e0.record()
            m(U, t_end=3.0, boundary_level=0.01)
            e1.record()
            torch.cuda.synchronize()
        steps = len(m.last_dt_series)
        ms = e0.elapsed_time(e1)
        rows.append({"grid": f"{nx}x{ny}", "dtype": name, "case": label, "steps": steps,
                     "us_per_step": round(1e3 * ms / steps, 2),
                     "gcell_updates_per_s": round(nx * ny * steps / (ms * 1e-3) / 1e9, 4)})
        print(f"# {rows[-1]}", file=sys.stderr, flush=True)
    return rows


def basin_model(n, device):
    import torch

    import diffswe2d_amd as swe
    from diffswe2d_amd import scenarios

    t0 = time.time()
    sc = scenarios.tsunami_basin(n=n)
    cfg = swe.SWEConfig(boundary_left="transmissive", boundary_right="transmissive",
                        boundary_top="transmissive", boundary_bottom="transmissive")
    bed = torch.from_numpy(sc["bed"])
    rough = torch.from_numpy(sc["rough"])
    model = swe.SWE2D(bed, rough, sc["dx"], sc["dy"], config=cfg, train_dem=False)
    model = model.to(device)
    U = torch.from_numpy(sc["U"]).to(device)
    print(f"# basin {n}x{n} built in {time.time() - t0:.1f} s (land {sc['land']:.3f})", file=sys.stderr,
          flush=True)
    return model, U, sc, cfg


def timed_run(model, U, warmup, steps, fast, events=None):
    """Device-driven run of warmup + steps steps from U; times the last `steps`.

    Returns (seconds, dt series, final state).  fast: the library's
    tolerance-bounded FP32 arithmetic instead of the reference's rounding.
    events: a list that receives the HIP-event time (ms) of the timed steps,
    recorded on the stream the step kernels are launched on."""
    import torch

    from diffswe2d_amd import _lib

    lib = _lib.load()
    bed, rough = model._fields(U)
    ws = model._workspace(U, bed, rough, fast=fast)
    lvl, stride = model._level_arg(0.0, U)
    stream = _lib.stream_ptr(U.device)
    Ua, Ub = U.clone(), torch.empty_like(U)
    dts = torch.empty(steps + warmup + 8, dtype=torch.float32, device=U.device)

    def run(first, count):
        _lib.check(lib.swe_advance_f32(ws["grid"], ws["cfg"], _lib.ptr(bed), _lib.ptr(rough), _lib.ptr(lvl),
                                       stride, _lib.ptr(Ua), _lib.ptr(Ub), first, count, None, _lib.ptr(dts),
                                       dts.numel(), None, 0, None, None, 0, _lib.ptr(ws["buf"]),
                                       int(model.rows_hint), stream), "advance")

    _lib.check(lib.swe_run_begin(_lib.ptr(ws["buf"]), 1e30, -1.0, 4, stream), "run_begin")
    run(0, warmup)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record()
    run(warmup, steps)
    e1.record()
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    if events is not None:
        events.append(e0.elapsed_time(e1))
    ctl = _lib.read_control(ws["buf"])
    from diffswe2d_amd.solver import sync_parity

    sync_parity(ws, ctl)
    ws["cfg"].arithmetic = 0
    if ctl.status != 0 or ctl.steps != warmup + steps:
        raise RuntimeError(f"run failed: status {ctl.status}, steps {ctl.steps}")
    return sec, dts[:warmup + steps], (Ua if (warmup + steps) % 2 == 0 else Ub)


def valu_mix_probe_seconds(dev, blocks=8192, iters=512, reps=3):
    """Best launch time of the library's VALU mix probe (the stage kernels' x-sweep arithmetic on
    register values); its VALU instruction count per launch comes from a committed SQ_INSTS_VALU run."""
    import torch

    from diffswe2d_amd import _lib

    lib = _lib.load()
    out = torch.empty(blocks * 256, dtype=torch.float32, device=dev)
    stream = _lib.stream_ptr(dev)
    _lib.check(lib.swe_valu_mix_probe(_lib.ptr(out), blocks, 8, stream), "valu_mix_probe")
    best = None
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        _lib.check(lib.swe_valu_mix_probe(_lib.ptr(out), blocks, iters, stream), "valu_mix_probe")
        e1.record()
        torch.cuda.synchronize(dev)
        ms = e0.elapsed_time(e1)
        best = ms if best is None else min(best, ms)
    if not bool(torch.isfinite(out).all()):
        raise RuntimeError("VALU mix probe produced non-finite values")
    return best * 1e-3


def valu_peak(dev, blocks=8192, iters=4096, reps=3):
    """Measured FP32 VALU issue ceiling (lane-instructions / s) from the library's probe kernel."""
    import torch

    from diffswe2d_amd import _lib

    lib = _lib.load()
    out = torch.empty(blocks * 256, dtype=torch.float32, device=dev)
    stream = _lib.stream_ptr(dev)
    _lib.check(lib.swe_valu_probe(_lib.ptr(out), blocks, 16, stream), "valu_probe")
    best = None
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        _lib.check(lib.swe_valu_probe(_lib.ptr(out), blocks, iters, stream), "valu_probe")
        e1.record()
        torch.cuda.synchronize(dev)
        ms = e0.elapsed_time(e1)
        best = ms if best is None else min(best, ms)
    wave_instr = blocks * 4 * iters * 16  # 4 wavefronts per block, 16 VALU instructions per iteration
    return wave_instr * 64 / (best * 1e-3)


def time_kernels(model, U, steps, rows, fast=False):
    """Average duration of each launch, bracketed by events on torch's current stream."""
    import torch

    from diffswe2d_amd import _lib

    lib = _lib.load()
    bed, rough = model._fields(U)
    ws = model._workspace(U, bed, rough, fast=fast)
    lvl, stride = model._level_arg(0.0, U)
    stream = _lib.stream_ptr(U.device)
    A, Bf = U.clone(), torch.empty_like(U)
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(steps)]
    torch.cuda.synchronize()
    for k in range(steps):
        e = evs[k]
        e[0].record()
        _lib.check(lib.swe_step_stage_f32(ws["grid"], ws["cfg"], 1, _lib.ptr(A), _lib.ptr(bed), _lib.ptr(rough),
                                          _lib.ptr(lvl), stride, 1e30, ws["parity"], _lib.ptr(Bf),
                                          _lib.ptr(ws["buf"]), rows, stream), "stage 1")
        e[1].record()
        _lib.check(lib.swe_step_stage_f32(ws["grid"], ws["cfg"], 2, _lib.ptr(A), _lib.ptr(bed), _lib.ptr(rough),
                                          _lib.ptr(lvl), stride, 1e30, ws["parity"], _lib.ptr(Bf),
                                          _lib.ptr(ws["buf"]), rows, stream), "stage 2")
        e[2].record()
        ws["parity"] ^= 1
        A, Bf = Bf, A
    torch.cuda.synchronize()
    ws["cfg"].arithmetic = 0
    s1 = sum(e[0].elapsed_time(e[1]) for e in evs) / steps
    s2 = sum(e[1].elapsed_time(e[2]) for e in evs) / steps
    return s1, s2


def time_adjoint(model, U, steps):
    """Per-step costs of the differentiable run's pieces on `steps` steps.

    forward_tape: device-driven forward that records the segment tape (states,
    k1, k2); reverse: reverse sweep of that tape; replay_reverse: a checkpointed
    segment's backward (replay from its first state onto a scratch tape, then
    the reverse sweep), all through the library's segment entry points."""
    import ctypes

    import torch

    from diffswe2d_amd import _lib, adjoint

    lib = _lib.load()
    bed, rough = model._fields(U)
    ws = model._workspace(U, bed, rough)
    aws = adjoint._adj_ws(model, U)
    lvl, stride = model._level_arg(0.0, U)
    stream = _lib.stream_ptr(U.device)
    rows = int(model.rows_hint)
    n = steps
    tape = torch.empty((3 * n + 1, *U.shape), dtype=U.dtype, device=U.device)
    dts = torch.empty(n, dtype=U.dtype, device=U.device)
    hmax = U[:, 0].clone()

    def forward_tape():
        tape[0].copy_(U)
        _lib.check(lib.swe_run_begin(_lib.ptr(ws["buf"]), 1e30, -1.0, U.element_size(), stream), "run_begin")
        _lib.check(lib.swe_advance_tape_f32(ws["grid"], ws["cfg"], _lib.ptr(bed), _lib.ptr(rough), _lib.ptr(lvl),
                                            stride, _lib.ptr(tape[0]), _lib.ptr(tape[n + 1]),
                                            _lib.ptr(tape[2 * n + 1]), 0, 0, n, _lib.ptr(hmax), _lib.ptr(dts), n,
                                            _lib.ptr(ws["buf"]), rows, stream), "advance_tape")

    forward_tape()  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    forward_tape()
    torch.cuda.synchronize()
    t_fwd = (time.perf_counter() - t0) / n
    ctl = _lib.read_control(ws["buf"])
    if ctl.steps != n or ctl.status != 0:
        raise RuntimeError(f"taped forward failed: steps {ctl.steps}, status {ctl.status}")
    from diffswe2d_amd.solver import sync_parity

    sync_parity(ws, ctl)
    host_dts = (ctypes.c_double * n)(*dts.cpu().tolist())
    lam0 = torch.ones_like(U)
    lam, lam_tmp = torch.empty_like(U), torch.empty_like(U)
    lam_level = torch.zeros(U.shape[0], dtype=U.dtype, device=U.device)
    scratch = torch.empty(lib.swe_segment_tape_bytes(ws["grid"], n, 4) // 4, dtype=U.dtype, device=U.device)

    def reverse():
        lam.copy_(lam0)
        _lib.check(lib.swe_segment_reverse_f32(
            ws["grid"], ws["cfg"], _lib.ptr(tape[0]), _lib.ptr(tape[n + 1]), _lib.ptr(tape[2 * n + 1]),
            _lib.ptr(bed), _lib.ptr(rough), _lib.ptr(lvl), stride, ctypes.addressof(host_dts), n, _lib.ptr(lam),
            _lib.ptr(lam_tmp), _lib.ptr(lam_level), _lib.ptr(aws["buf"]), rows, stream),
            "segment_reverse", "swe_adjoint_last_error_message")

    def replay_reverse():
        lam.copy_(lam0)
        _lib.check(lib.swe_segment_adjoint_f32(
            ws["grid"], ws["cfg"], _lib.ptr(tape[0]), _lib.ptr(bed), _lib.ptr(rough), _lib.ptr(lvl), stride,
            ctypes.addressof(host_dts), n, ws["parity"], _lib.ptr(lam), _lib.ptr(lam_tmp), _lib.ptr(lam_level),
            _lib.ptr(scratch), _lib.ptr(ws["buf"]), _lib.ptr(aws["buf"]), rows, stream),
            "segment_adjoint", "swe_adjoint_last_error_message")
        ws["parity"] ^= n & 1

    out = {}
    for name, fn in (("reverse", reverse), ("replay_reverse", replay_reverse)):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        out[name] = (time.perf_counter() - t0) / n * 1e3
        if not bool(torch.isfinite(lam).all()):
            raise RuntimeError("adjoint produced non-finite values")
    return t_fwd * 1e3, out["reverse"], out["replay_reverse"]


def cpu_model_name():
    """The host CPU's model name (/proc/cpuinfo), or None."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(sc, cfg, steps, run_steps):
    """Oracle (eager PyTorch CPU, FP32, the reference's op sequence) on the full basin.

    One warm-up step, then `steps` timed steps of the whole grid; the per-step
    time is also extrapolated to the run's `run_steps` steps (labelled as such)."""
    import torch

    from oracle import swe_oracle as O

    U = torch.from_numpy(sc["U"])
    bed = torch.from_numpy(sc["bed"])
    rough = torch.from_numpy(sc["rough"])
    threads = torch.get_num_threads()
    n = U.shape[-1]
    with torch.no_grad():
        t0 = time.perf_counter()
        U, _ = O.step(U, bed, rough, cfg, sc["dx"], sc["dy"], 1e30)  # warm-up
        warm = time.perf_counter() - t0
        print(f"# cpu baseline: warm-up step {warm:.1f} s", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        for _ in range(steps):
            U, _ = O.step(U, bed, rough, cfg, sc["dx"], sc["dy"], 1e30)
        sec = time.perf_counter() - t0
    per_step = sec / steps
    return dict(value=n * n / per_step / 1e9, unit="Gcell-updates/s", cores=threads, kind="port",
                sample=f"oracle FP32 (eager PyTorch, the reference's op sequence) on the full {n}x{n} basin: "
                       f"1 warm-up + {steps} timed steps, {per_step:.2f} s/step on {threads} threads",
                seconds_per_step=round(per_step, 3),
                extrapolated_run_seconds={"steps": run_steps, "seconds": round(per_step * run_steps, 1),
                                          "note": "extrapolated from the timed steps, not measured"},
                os_cpu_count=os.cpu_count(), cpu_model=cpu_model_name(),
                threads_note="torch threads = min(os.cpu_count(), OMP_NUM_THREADS): the host's CPU share "
                             "for one GPU")


def oracle_rate(U, bed, rough, cfg, dx, dy, steps, level=None):
    """Oracle (eager PyTorch CPU, FP32) steps per second on one grid: 1 warm-up + `steps` timed."""
    import torch

    from oracle import swe_oracle as O

    with torch.no_grad():
        U, _ = O.step(U, bed, rough, cfg, dx, dy, 1e30, level=level)
        t0 = time.perf_counter()
        for _ in range(steps):
            U, _ = O.step(U, bed, rough, cfg, dx, dy, 1e30, level=level)
        return steps / (time.perf_counter() - t0)


def small_workloads(dev, cpu_steps):
    """The two smaller forward workloads, each on the GPU next to the oracle on the host:
    the 256 x 256 dam break (500 steps, walls, flat bed, n = 0) and the full 544 x 344 coastal
    run-up (450 chunks of 0.05 s with the west level of its pulse train, 6 gauges sampled on the
    device).  The 32-B roofline fraction is reported for each; both are latency-bound, one
    wavefront's strip per SIMD or less."""
    import torch

    import diffswe2d_amd as swe
    from diffswe2d_amd import chunked, scenarios

    out = {}
    # dam break
    sc = scenarios.dam_break(n=256)
    cfg = swe.SWEConfig(boundary_left="wall", boundary_right="wall", boundary_top="wall", boundary_bottom="wall")
    bed, rough = torch.tensor(sc["bed"], dtype=torch.float32), torch.tensor(sc["rough"], dtype=torch.float32)
    m = swe.SWE2D(bed, rough, sc["dx"], sc["dy"], config=cfg, train_dem=False).to(dev)
    U = torch.tensor(sc["U"], dtype=torch.float32, device=dev)
    cells = U.shape[-1] * U.shape[-2]
    timed_run(m, U, 1, sc["steps"], fast=False)  # warm-up run
    ev = []
    sec, _, _ = timed_run(m, U, 1, sc["steps"], fast=False, events=ev)
    gpu = cells * sc["steps"] / (ev[0] * 1e-3) / 1e9
    rate = (oracle_rate(torch.tensor(sc["U"], dtype=torch.float32), bed, rough, cfg, sc["dx"], sc["dy"], cpu_steps)
            if cpu_steps > 0 else None)
    out["dam_break_256x256"] = {
        "steps": sc["steps"], "ms_per_step": round(ev[0] / sc["steps"], 4), "gcell_updates_per_s": round(gpu, 3),
        "hbm_frac_32B": round(gpu * 32 / HBM_PEAK_GBS, 4),
        "cpu_gcell_updates_per_s": None if rate is None else round(cells * rate / 1e9, 5),
        "cpu_sample": f"oracle FP32, 1 warm-up + {cpu_steps} steps from the initial state"}
    # coastal run-up, the whole 22.5 s record
    sc = scenarios.coastal_runup()
    cfg = swe.SWEConfig(boundary_left="water_level", boundary_right="wall", boundary_top="wall",
                        boundary_bottom="wall")
    bed, rough = torch.tensor(sc["bed"], dtype=torch.float32), torch.tensor(sc["rough"], dtype=torch.float32)
    m = swe.SWE2D(bed, rough, sc["dx"], sc["dy"], config=cfg, train_dem=False).to(dev)
    U = torch.tensor(sc["U"], dtype=torch.float32, device=dev)
    levels = torch.tensor(sc["chunk_levels"], dtype=torch.float32, device=dev)
    cells = U.shape[-1] * U.shape[-2]
    with torch.no_grad():
        chunked.gauge_run(m, U, levels, sc["sample_dt"], sc["gauges"])  # warm-up run
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        chunked.gauge_run(m, U, levels, sc["sample_dt"], sc["gauges"])
        e1.record()
        torch.cuda.synchronize()
    steps = len(m.last_dt_series)
    ms = e0.elapsed_time(e1)
    gpu = cells * steps / (ms * 1e-3) / 1e9
    rate = (oracle_rate(torch.tensor(sc["U"], dtype=torch.float32), bed, rough, cfg, sc["dx"], sc["dy"], cpu_steps,
                        level=float(sc["chunk_levels"][0])) if cpu_steps > 0 else None)
    out["coastal_runup_544x344"] = {
        "chunks": len(sc["chunk_levels"]), "steps": steps, "seconds": round(ms * 1e-3, 4),
        "ms_per_step": round(ms / steps, 4), "gcell_updates_per_s": round(gpu, 3),
        "hbm_frac_32B": round(gpu * 32 / HBM_PEAK_GBS, 4),
        "cpu_gcell_updates_per_s": None if rate is None else round(cells * rate / 1e9, 5),
        "cpu_sample": f"oracle FP32, 1 warm-up + {cpu_steps} steps from the initial state (first chunk's level)"}
    return out


def per_workload(small, gcells, achieved, cpu, inv, warm, adj):
    """Gcell-updates/s and the 32-B roofline fraction of every workload the run measured,
    with the oracle's rate on the host where it was timed."""
    out = dict(small or {})
    out["tsunami_basin_8192x8192"] = {
        "gcell_updates_per_s": round(gcells, 3), "hbm_frac_32B": round(achieved / HBM_PEAK_GBS, 4),
        "cpu_gcell_updates_per_s": None if cpu is None else round(cpu["value"], 5)}
    if adj is not None:
        out["tsunami_basin_8192x8192_adjoint"] = {
            "reverse_gcell_updates_per_s": adj["value"],
            "taped_backward_over_forward": adj["taped_backward_over_forward"],
            "checkpointed_backward_over_forward": adj["checkpointed_backward_over_forward"]}
    if inv is not None:
        out["monai_boundary_inversion_68x42_fp64"] = {"seconds_per_epoch": inv["seconds_per_epoch"],
                                                     "speedup_vs_published": inv["speedup_vs_published"]}
    if warm is not None:
        out["runup_boundary_inversion_544x344"] = {
            "seconds_per_epoch": round(warm["plain"]["seconds"] / max(1, warm["plain"]["epochs"]), 3),
            "warm_over_plain": warm["ratio_warm_over_plain"], "with_one_off": warm["ratio_with_one_off"]}
        out["ensemble_256_scenarios"] = {"gcell_updates_per_s": warm["one_off"]["ensemble_gcell_per_s"],
                                         "arithmetic": warm["one_off"]["ensemble_arithmetic"],
                                         "hbm_frac_32B": round(warm["one_off"]["ensemble_gcell_per_s"] * 32
                                                               / HBM_PEAK_GBS, 4)}
    return out


def warmstart_leg(dev, n_scenarios, epochs):
    """The learned warm start on the coastal run-up inversion (544 x 344, 64 spline controls,
    450 chunks): plain adjoint inversion from the prior vs surrogate-started inversion, time to
    the plain run's best gauge misfit, with and without the one-off ensemble and training."""
    from diffswe2d_amd import warmstart

    rep = warmstart.run(dev, n_scenarios=n_scenarios, epochs_plain=epochs, log=None)
    plain, warm = rep["plain"], rep["warm"]
    one_off = rep["ensemble"]["seconds"] + rep["ensemble"].get("tune_seconds", 0.0) + rep["surrogate"]["seconds"]
    t_warm = warm["seconds_to_bar"]
    return {"workload": f"coastal run-up boundary inversion 544x344 FP32, 64 spline controls, 450 chunks; "
                        f"{n_scenarios}-scenario ensemble (272x172 / 544x344 / 1088x688), 1D-conv surrogate",
            "plain": {"epochs": plain["epochs"], "seconds": round(plain["seconds"], 2),
                      "best_misfit": plain["best_misfit"]},
            "warm": {"reached_plain_misfit": warm["reached"], "seconds_to_plain_misfit": t_warm,
                     "epochs": warm["epochs"], "start_misfit": warm["start_misfit"],
                     "prior_misfit": warm["prior_misfit"], "fallback_to_prior": warm["fallback"]},
            "one_off": {"ensemble_seconds": round(rep["ensemble"]["seconds"], 2),
                        "ensemble_rows_tuning_seconds": round(rep["ensemble"].get("tune_seconds", 0.0), 2),
                        "ensemble_gcell_per_s": round(rep["ensemble"]["gcell_per_s"], 2),
                        "ensemble_arithmetic": rep["ensemble"].get("arithmetic"),
                        "training_seconds": round(rep["surrogate"]["seconds"], 2)},
            "bar": "the plain run's best gauge misfit after its epochs (fewer epochs, a looser bar); the warm run "
                   "starts from the surrogate and stops at that bar",
            "ratio_warm_over_plain": None if t_warm is None else round(t_warm / plain["seconds"], 4),
            "ratio_with_one_off": None if t_warm is None else round((t_warm + one_off) / plain["seconds"], 4)}


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "the benchmark runs on one GPU; --gpus must be 1 (multi-GPU sharding is "
                                   "not implemented)"}))
        sys.exit(2)
    import torch

    if not torch.cuda.is_available():
        print(json.dumps({"error": "no GPU visible"}))
        sys.exit(1)
    dev = torch.device("cuda:0")
    if a.small_grids:
        print(json.dumps({"small_grids": small_grids(dev)}))
        return
    model, U, sc, cfg = basin_model(a.n, dev)
    tuned = None
    if a.rows:
        model.rows_hint = a.rows
    else:
        from diffswe2d_amd import autotune

        # tune on the state the timed steps start from: the initial state, whose momenta are
        # exactly zero, steps ~6 % faster than the states a few steps on
        _, _, U_tune = timed_run(model, U, 0, max(a.warmup, 1), fast=False)
        best, rep = autotune.tune_rows(model, U_tune, candidates=(16, 32, 40, 48, 56, 64, 128), reps=10,
                                       save=False)
        del U_tune
        tuned = {str(k): round(v, 4) for k, v in rep.items()}
        print(f"# autotuned rows per strip: {best} ({tuned})", file=sys.stderr, flush=True)
    n = a.n
    cells = n * n

    from diffswe2d_amd import _lib

    ev_ms = []
    sec, dts, U_end = timed_run(model, U, a.warmup, a.steps, fast=False, events=ev_ms)
    ms = sec / a.steps * 1e3
    gcells = cells * a.steps / sec / 1e9
    dt_mean = float(dts[a.warmup:a.warmup + a.steps].double().mean())
    print(f"# {a.steps} steps in {sec:.3f} s: {ms:.3f} ms/step, {gcells:.2f} Gcell/s, mean dt {dt_mean:.4f} s",
          file=sys.stderr, flush=True)

    fast = None
    if not a.no_fast:
        # the same run with the tolerance-bounded arithmetic, and its distance from the
        # reference-rounding run above after the same steps
        sec_f, dts_f, U_f = timed_run(model, U, a.warmup, a.steps, fast=True)
        f1, f2 = time_kernels(model, U, a.kernel_steps, int(model.rows_hint), fast=True)
        d = dts.double()
        diff = float(torch.linalg.norm((U_f - U_end).double()) / torch.linalg.norm(U_end.double()))
        fast = {"value": round(cells * a.steps / sec_f / 1e9, 4), "unit": "Gcell-updates/s",
                "ms_per_step": round(sec_f / a.steps * 1e3, 4),
                "kernel_ms": {"stage1": round(f1, 4), "stage2": round(f2, 4)},
                "rel_l2_vs_reference_rounding": diff,
                "dt_max_rel_vs_reference_rounding": float(((dts_f.double() - d).abs() / d).max()),
                "wet_dry_mismatch_vs_reference_rounding": int(((U_f[:, 0] > 1e-6) != (U_end[:, 0] > 1e-6)).sum()),
                "note": "SWE2D.arithmetic = 'fast' (swe_config.arithmetic = SWE_ARITH_FAST): division with one "
                        "correction, hardware sqrt, h^(1/3) from log2/exp2, no velocity re-division; "
                        "opt-in, forward runs without gradients only; parity bars in tests/test_fast_arith_gpu.py"}
        print(f"# fast arithmetic: {fast['ms_per_step']:.3f} ms/step, {fast['value']:.2f} Gcell/s, events stage1 "
              f"{f1:.3f} stage2 {f2:.3f} ms; rel-L2 vs reference rounding after {a.warmup + a.steps} steps "
              f"{diff:.3e}", file=sys.stderr, flush=True)

    s1, s2 = time_kernels(model, U, a.kernel_steps, int(model.rows_hint))
    # One cell-update moves 32 algorithmic bytes (read h, hu, hv, z, n; write h, hu, hv).
    # The unit of work is one step = the two launches back to back, so the roofline
    # prices the step's duration from the HIP events around the timed region (the
    # device-driven loop above, on the launch stream).  s1 and s2 bracket single
    # launches with events for the per-stage split; the event records between
    # launches add their own cost, so their sum runs a few per cent above the step.
    step_kernel_ms = ev_ms[0] / a.steps
    achieved = 32.0 * cells / (step_kernel_ms * 1e-3) / 1e9
    print(f"# events: step {step_kernel_ms:.3f} ms over the timed region; single launches stage1 {s1:.3f} ms, "
          f"stage2 {s2:.3f} ms", file=sys.stderr, flush=True)
    if a.copy_probe:
        src = torch.empty(1 << 28, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print("# copy probe: 3 x 1 GiB device copies", file=sys.stderr, flush=True)

    adj = None
    if a.adjoint_steps > 0:
        fwd_tape_ms, rev_ms, ckpt_ms = time_adjoint(model, U, a.adjoint_steps)
        # checkpointed backward (K-step segments): replay onto a scratch tape + reverse sweep.
        # taped backward: the forward kept the tape (forward_tape_ms per step), the backward only reverses.
        adj = {"value": round(cells / (rev_ms * 1e-3) / 1e9, 4), "unit": "Gcell-updates/s",
               "reverse_ms_per_step": round(rev_ms, 4),
               "checkpointed_backward_ms_per_step": round(ckpt_ms, 4),
               "checkpointed_backward_over_forward": round(ckpt_ms / ms, 3),
               "taped_forward_ms_per_step": round(fwd_tape_ms, 4),
               "taped_backward_over_forward": round(rev_ms / ms, 3),
               "taped_gradient_over_forward": round((fwd_tape_ms + rev_ms) / ms, 3),
               "steps": a.adjoint_steps}
        print(f"# adjoint: taped forward {fwd_tape_ms:.3f} ms, reverse {rev_ms:.3f} ms, checkpointed backward "
              f"{ckpt_ms:.3f} ms per step ({rev_ms / ms:.2f}x / {ckpt_ms / ms:.2f}x forward)",
              file=sys.stderr, flush=True)

    # HBM bytes per step from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE runs of this workload
    traffic, traffic_src = None, None
    tfile = ROOT / "profiles" / "basin8192_fp32_traffic.json"
    valu = None
    if n == 8192 and tfile.exists():
        prof = json.loads(tfile.read_text())
        # the committed counters of this run's strip height when profiled, else the default's
        rows_run = int(model.rows_hint)
        at = prof.get("by_rows", {}).get(str(rows_run))
        rows_prof = rows_run if at is not None else prof.get("rows")
        at = at or prof
        traffic = round(at["step_hbm_bytes"])
        # second roof: VALU issue.  Instructions per cell from the committed SQ_INSTS_VALU run of
        # this workload, the ceiling from the probe kernel measured here.
        per_cell = at["stage1"]["valu_lane_instr_per_cell"] + at["stage2"]["valu_lane_instr_per_cell"]
        got = per_cell * cells / (step_kernel_ms * 1e-3)
        mix = prof.get("valu_mix_probe")
        sec_probe = valu_mix_probe_seconds(dev, MIX_PROBE_BLOCKS, MIX_PROBE_ITERS)
        if mix is not None and (mix["blocks"], mix["iters"]) == (MIX_PROBE_BLOCKS, MIX_PROBE_ITERS):
            # ceiling: the issue rate of the kernels' own arithmetic without memory (the mix probe)
            peak = mix["SQ_INSTS_VALU"] * 64 / sec_probe
            src = ("profiles/basin8192_fp32_traffic.json (SQ_INSTS_VALU of the stage kernels and of "
                   "swe_valu_mix_probe); the probe timed here")
            note = ("peak is the issue rate of swe_valu_mix_probe: the stage kernels' x-sweep arithmetic "
                    "(divisions, square roots, limiter, HLL, DPP shifts) on register values, no memory")
        else:
            peak = valu_peak(dev)
            src = "profiles/basin8192_fp32_traffic.json (SQ_INSTS_VALU); peak from swe_valu_probe"
            note = "peak is the measured issue rate of a plain FP32 mul/add stream"
        valu = {"bound": "valu", "achieved": round(got / 1e12, 2), "peak": round(peak / 1e12, 2),
                "unit": "T lane-instr/s", "frac": round(got / peak, 4),
                "lane_instr_per_cell_update": round(per_cell, 1), "rows_per_strip_profiled": rows_prof,
                "source": src, "note": note}
        print(f"# VALU: {got / 1e12:.2f} of {peak / 1e12:.2f} T lane-instr/s ({got / peak:.2f})", file=sys.stderr,
              flush=True)
        traffic_src = (f"profiles/basin8192_fp32_traffic.json at {rows_prof} rows per strip (FETCH_SIZE "
                       f"x{prof['fetch_factor']} + WRITE_SIZE x{prof['write_factor']}, copy-calibrated)")

    # the host's CPU share for one GPU, for every oracle leg below
    torch.set_num_threads(max(1, min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "16")))))
    small = None
    if not a.no_small:
        small = small_workloads(dev, 0 if a.no_cpu else a.small_cpu_steps)
        for name, r in small.items():
            print(f"# {name}: {r['steps']} steps, {r['ms_per_step']:.4f} ms/step, {r['gcell_updates_per_s']:.3f} "
                  f"Gcell/s ({100 * r['hbm_frac_32B']:.2f} % of 8 TB/s); oracle "
                  f"{r['cpu_gcell_updates_per_s'] or float('nan'):.5f} Gcell/s", file=sys.stderr, flush=True)

    inv = None
    if a.inversion_epochs > 0:
        inv = inversion_leg(dev, a.inversion_epochs)
        print(f"# inversion: {inv['epochs']} epochs, best at {inv['best_epoch']} after {inv['seconds_to_best']} s "
              f"(published run: {inv['published']['seconds_to_best']} s; {inv['speedup_vs_published']}x)",
              file=sys.stderr, flush=True)

    warm = None
    if a.warmstart_scenarios > 0:
        warm = warmstart_leg(dev, a.warmstart_scenarios, a.warmstart_epochs)
        print(f"# warm start: plain {warm['plain']['seconds']} s to misfit {warm['plain']['best_misfit']:.3e}; warm "
              f"{warm['warm']['seconds_to_plain_misfit']} s (ratio {warm['ratio_warm_over_plain']}, with the one-off "
              f"ensemble and training {warm['ratio_with_one_off']})", file=sys.stderr, flush=True)

    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(sc, cfg, a.cpu_steps, sc["steps"])
        print(f"# cpu baseline {cpu['value']:.4f} Gcell/s ({cpu['sample']})", file=sys.stderr, flush=True)

    out = {
        "metric": METRIC, "value": round(gcells, 4), "unit": "Gcell-updates/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"tsunami_basin_{n}x{n}_dx100m_manning0.025_transmissive_cfl0.5",
                   "cells": cells, "steps_of_run": a.steps, "launches_per_step": 2,
                   "rows_per_strip": int(model.rows_hint), "autotune_ms_per_step": tuned},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_source": traffic_src,
                     "kernel": "step (stage1 + stage2)", "algorithmic_bytes_per_cell_update": 32,
                     "kernel_ms": round(step_kernel_ms, 4),
                     "kernel_ms_source": "HIP events around the timed steps (device-driven loop, launch stream)"},
        "roofline_valu": valu,
        "arith_fast": fast,
        "kernel_ms": {"stage1": round(s1, 4), "stage2": round(s2, 4),
                      "note": "each launch bracketed by its own events (per-stage split)"},
        "adjoint": adj,
        "inversion": inv,
        "warmstart": warm,
        "cpu_baseline": cpu,
        "per_workload": per_workload(small, gcells, achieved, cpu, inv, warm, adj),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
