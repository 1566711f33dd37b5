"""HIP forward step against the CPU oracle and the reference's golden vectors.

Bars (FP32): tendency and states bitwise or within a few ulps of the oracle on
small grids; after the full 500-step dam break rel-L2 <= 1e-5, identical
wet/dry masks and dt series within 1e-6 relative.
"""
import ctypes
import math

import numpy as np
import pytest
import torch

import diffswe2d_amd as swe
from diffswe2d_amd import scenarios
from oracle import swe_oracle as O

pytestmark = pytest.mark.gpu

DT = {"f32": torch.float32, "f64": torch.float64}

STEP_CASES = [
    dict(bc=("wall", "wall", "wall", "wall"), friction=True, level=0.0),
    dict(bc=("water_level", "wall", "transmissive", "water_level"), friction=True, level=0.45),
    dict(bc=("periodic", "periodic", "absorbing", "constant"), friction=False, level=0.3),
    dict(bc=("transmissive", "absorbing", "periodic", "periodic"), friction=True, level=0.2),
]


def relerr(a, b):
    a = torch.as_tensor(a, dtype=torch.float64).cpu()
    b = torch.as_tensor(b, dtype=torch.float64).cpu()
    return float(torch.linalg.norm(a - b) / max(float(torch.linalg.norm(b)), 1e-300))


def make_model(d, case, dtype, dev):
    c = STEP_CASES[case]
    cfg = swe.SWEConfig(boundary_left=c["bc"][0], boundary_right=c["bc"][1], boundary_top=c["bc"][2],
                        boundary_bottom=c["bc"][3], apply_friction=c["friction"], constant_value=0.05)
    m = swe.SWE2D(torch.tensor(d["bed"], dtype=dtype), torch.tensor(d["rough"], dtype=dtype), 0.1, 0.1,
                  config=cfg).to(dev)
    return m, cfg


@pytest.mark.parametrize("case", range(len(STEP_CASES)))
@pytest.mark.parametrize("name", ["f32", "f64"])
def test_spatial_operator_matches_golden(golden, cuda_device, case, name):
    d = golden("steps")
    dtype = DT[name]
    m, _ = make_model(d, case, dtype, cuda_device)
    U = torch.tensor(d["U0"], dtype=dtype, device=cuda_device)
    lvl = torch.tensor(STEP_CASES[case]["level"], dtype=dtype)
    rhs, ax, ay = m.spatial_operator(U, return_wave_speed=True, boundary_level=lvl)
    ref = torch.tensor(d[f"s{case}_{name}_rhs"])
    tol = 2e-6 if name == "f32" else 1e-13
    diff = (rhs.cpu().double() - ref.double()).abs().max().item()
    scale = ref.abs().max().item()
    print(f"case {case} {name}: max |rhs diff| {diff:.3e} (scale {scale:.3e})")
    assert diff <= tol * scale
    axay = d[f"s{case}_{name}_axay"]
    assert abs(float(ax) - axay[0]) <= tol * axay[0]
    assert abs(float(ay) - axay[1]) <= tol * axay[1]


@pytest.mark.parametrize("case", range(len(STEP_CASES)))
@pytest.mark.parametrize("name", ["f32", "f64"])
def test_steps_match_golden(golden, cuda_device, case, name):
    d = golden("steps")
    dtype = DT[name]
    m, _ = make_model(d, case, dtype, cuda_device)
    U = torch.tensor(d["U0"], dtype=dtype, device=cuda_device)
    lvl = torch.tensor(STEP_CASES[case]["level"], dtype=dtype)
    tol = 1e-6 if name == "f32" else 1e-12
    for k in range(12):
        U, dt = m.step(U, 0.0, 1e30, boundary_level=lvl)
        ref = d[f"s{case}_{name}_U"][k]
        e = relerr(U, ref)
        assert e <= tol, f"step {k}: rel-L2 {e:.3e}"
        assert abs(dt - d[f"s{case}_{name}_dt"][k]) <= tol * d[f"s{case}_{name}_dt"][k]
        wet = (U[:, 0] > 1e-6).cpu().numpy()
        assert np.array_equal(wet, ref[:, 0] > 1e-6), f"wet/dry mask differs at step {k}"


def test_dam_break_full_run(golden, cuda_device):
    """Config: 256^2 circular dam break, walls, 500 steps, FP32 (rel-L2 <= 1e-5 after the run)."""
    d = golden("dam_break")
    sc = scenarios.dam_break()
    cfg = swe.SWEConfig(boundary_left="wall", boundary_right="wall", boundary_top="wall", boundary_bottom="wall")
    m = swe.SWE2D(torch.tensor(sc["bed"], dtype=torch.float32), torch.tensor(sc["rough"], dtype=torch.float32),
                  sc["dx"], sc["dy"], config=cfg).to(cuda_device)
    U = torch.tensor(sc["U"], dtype=torch.float32, device=cuda_device)
    dts = []
    for k in range(500):
        U, dt = m.step(U, 0.0, 1e30)
        dts.append(dt)
        if k + 1 in (1, 10, 100):
            e = relerr(U, d[f"U{k + 1}_f32"])
            print(f"step {k + 1}: rel-L2 vs reference FP32 {e:.3e}")
    dts = np.array(dts)
    e32 = relerr(U, d["U500_f32"])
    e64 = relerr(U, d["U500_f64"])
    e_ref = relerr(d["U500_f32"], d["U500_f64"])
    dt_err = float(np.max(np.abs(dts - d["dt_f32"]) / d["dt_f32"]))
    print(f"dam break 500 steps: rel-L2 vs ref FP32 {e32:.3e}; vs ref FP64 {e64:.3e} "
          f"(ref FP32 vs FP64 {e_ref:.3e}); dt max rel {dt_err:.3e}")
    assert e32 <= 1e-5
    assert dt_err <= 1e-6
    wet = (U[:, 0] > 1e-6).cpu().numpy()
    assert np.array_equal(wet, d["U500_f32"][:, 0] > 1e-6)
    # symmetry of the circular problem: h(i, j) = h(j, i), hu(i, j) = hv(j, i)
    h = U[0, 0].cpu()
    assert torch.equal(h, h.T)
    assert torch.equal(U[0, 1].cpu(), U[0, 2].cpu().T)
    # no dry cell appears (the positivity limiter never fires) ...
    assert float(U[0, 0].min()) > 0.2
    # ... and the mass change over the run (the reference's wall faces are not
    # exactly conservative: +2.55e-4 relative after 500 steps) is the reference's
    m0 = float(np.asarray(sc["U"][0, 0], dtype=np.float64).sum())
    dm_gpu = float(U[0, 0].double().sum()) - m0
    dm_ref = float(d["U500_f32"][0, 0].astype(np.float64).sum()) - m0
    assert abs(dm_gpu - dm_ref) / m0 < 1e-7


def test_forward_device_loop_matches_host_steps(cuda_device):
    sc = scenarios.coastal_runup(nx=68, ny=43, dx=0.08)
    cfg = swe.SWEConfig(boundary_left="water_level", boundary_right="wall", boundary_top="wall",
                        boundary_bottom="wall")
    m = swe.SWE2D(torch.tensor(sc["bed"], dtype=torch.float32), torch.tensor(sc["rough"], dtype=torch.float32),
                  0.08, 0.08, config=cfg).to(cuda_device)
    U0 = torch.tensor(sc["U"], dtype=torch.float32, device=cuda_device)
    lvl = torch.tensor(0.01)
    m.track_gauges = True
    m.gauge_indices = {f"g{i}": g for i, g in enumerate(sc["gauges"])}
    m.reset_gauge_tracking()
    Ud, hd = m.forward(U0, t_end=0.5, start_time=1.0, boundary_level=lvl)
    dts_dev = list(m.ts_dt)
    g_dev = list(m.ts_h["g0"])
    m.reset_gauge_tracking()
    Uh, hh = m.forward(U0, t_end=0.5, start_time=1.0, boundary_level=lvl, callback=lambda *a: None)
    assert torch.equal(Ud, Uh)
    assert torch.equal(hd, hh)
    assert dts_dev == m.ts_dt
    assert g_dev == m.ts_h["g0"]
    assert abs(m.ts_time[-1] - 1.5) < 1e-12


@pytest.mark.parametrize("n_steps_parity", [0, 1])
def test_step_after_device_run_sees_only_its_own_wave_speeds(cuda_device, n_steps_parity):
    """A single step right after a device-driven run must not inherit that run's wave speeds.

    The run integrates a deep, fast state (an odd or even number of steps); the
    step then starts from a shallow calm state whose CFL dt is much larger.  Its
    dt must equal the oracle's dt for that state alone."""
    n = 24
    cfg = swe.SWEConfig(boundary_left="wall", boundary_right="wall", boundary_top="wall", boundary_bottom="wall",
                        apply_friction=False)
    bed = torch.zeros(1, 1, n, n)
    rough = torch.zeros(1, 1, n, n)
    m = swe.SWE2D(bed, rough, 0.1, 0.1, config=cfg, train_dem=False).to(cuda_device)
    fast = torch.zeros(1, 3, n, n)
    fast[0, 0] = 5.0
    fast[0, 0, : n // 2] = 8.0
    Uf = fast.to(cuda_device)
    d0 = float(O.step(fast, bed, rough, cfg, 0.1, 0.1, 1e30, 0.0)[1])  # the fast state's first dt
    steps_wanted = 3 + n_steps_parity  # the run ends after an odd / even number of steps
    m(Uf, t_end=(steps_wanted - 0.5) * d0)
    assert len(m.last_dt_series) >= 3
    calm = torch.zeros(1, 3, n, n)
    calm[0, 0] = 0.01
    out, dt = m.step(calm.to(cuda_device), 0.0, 1e30)
    out_ref, dt_ref = O.step(calm, bed, rough, cfg, 0.1, 0.1, 1e30, 0.0)
    print(f"run of {len(m.last_dt_series)} steps, then calm step dt {dt:.6e} vs oracle {float(dt_ref):.6e}")
    assert abs(dt - float(dt_ref)) <= 1e-6 * float(dt_ref)


def oracle_runup_ieee(d):
    """The oracle with correctly rounded FP32 sqrt on the same run-up (dt series, final state)."""
    import oracle.swe_oracle as OO

    cfg = swe.SWEConfig(boundary_left="water_level", boundary_right="wall", boundary_top="wall",
                        boundary_bottom="wall")
    old = OO.SQRT_MODE
    OO.SQRT_MODE = "ieee"
    try:
        U = torch.tensor(d["U0"], dtype=torch.float32)
        bed, rough = torch.tensor(d["bed"], dtype=torch.float32), torch.tensor(d["rough"], dtype=torch.float32)
        dts = []
        with torch.no_grad():
            for k in range(len(d["levels"])):
                U, _ = OO.forward(U, bed, rough, cfg, 0.04, 0.04, 0.05, start_time=0.05 * k,
                                  boundary_level=float(d["levels"][k]), dt_log=dts)
    finally:
        OO.SQRT_MODE = old
    return np.array(dts), U


def test_forward_matches_oracle_runup(golden, cuda_device):
    """Reduced coastal run-up (136 x 86, dx 0.04), 40 chunks of 0.05 s: moving wet/dry front.

    Bars against the reference (CPU, FP32): rel-L2 <= 1e-5, identical wet/dry
    masks, same step count.  The dt series is held to 1e-6 against the oracle
    with IEEE square roots; against the CPU reference it is held to 2e-6,
    because torch's CPU float32 sqrt is itself off by up to one ulp and moves
    the reference's dt series by 1.66e-6 on this run (measured: IEEE-sqrt
    oracle vs reference).
    """
    d = golden("runup")
    cfg = swe.SWEConfig(boundary_left="water_level", boundary_right="wall", boundary_top="wall",
                        boundary_bottom="wall")
    m = swe.SWE2D(torch.tensor(d["bed"], dtype=torch.float32), torch.tensor(d["rough"], dtype=torch.float32),
                  0.04, 0.04, config=cfg).to(cuda_device)
    U = torch.tensor(d["U0"], dtype=torch.float32, device=cuda_device)
    m.track_gauges = True
    m.gauge_indices = {f"g{i}": tuple(int(v) for v in g) for i, g in enumerate(d["gauges"])}
    m.reset_gauge_tracking()
    masks = []
    gauge = []
    for k in range(len(d["levels"])):
        U, _ = m.forward(U, t_end=0.05, start_time=0.05 * k, boundary_level=float(d["levels"][k]))
        gauge.append([float(U[0, 0, iy, ix]) for iy, ix in d["gauges"]])
        if (k + 1) % 10 == 0:
            masks.append((U[0, 0] > 1e-6).cpu().numpy())
    dts = np.array(m.ts_dt)
    ref_dt = d["dt_f32"]
    assert len(dts) == len(ref_dt)
    dt_err = float(np.max(np.abs(dts - ref_dt) / ref_dt))
    e32 = relerr(U, d["U_f32"])
    e64 = relerr(U, d["U_f64"])
    ref_masks = np.unpackbits(d["mask_f32"])[: np.stack(masks).size].reshape(np.stack(masks).shape).astype(bool)
    mismatch = int((np.stack(masks) != ref_masks).sum())
    ieee_dt, ieee_U = oracle_runup_ieee(d)
    dt_err_ieee = float(np.max(np.abs(dts - ieee_dt) / ieee_dt)) if len(ieee_dt) == len(dts) else float("inf")
    e_ieee = relerr(U, ieee_U)
    print(f"run-up: {len(dts)} steps, rel-L2 vs ref FP32 {e32:.3e}, vs FP64 {e64:.3e}, vs IEEE-sqrt oracle "
          f"{e_ieee:.3e}; dt max rel vs ref {dt_err:.3e}, vs IEEE-sqrt oracle {dt_err_ieee:.3e}; "
          f"mask mismatches {mismatch}")
    assert e32 <= 1e-5
    assert mismatch == 0
    assert dt_err_ieee <= 1e-6
    assert dt_err <= 2e-6


def test_error_paths(cuda_device):
    cfg = swe.SWEConfig(boundary_left="wall", boundary_right="wall", boundary_top="wall", boundary_bottom="wall")
    z = torch.zeros(1, 1, 8, 8)
    m = swe.SWE2D(z, z, 0.1, 0.1, config=cfg).to(cuda_device)
    U = torch.zeros(1, 3, 8, 8, device=cuda_device)
    U[:, 0] = 1.0
    with pytest.raises(FloatingPointError, match="Non-positive timestep"):
        m.step(U, 0.0, 0.0)
    Un = U.clone()
    Un[0, 1, 3, 3] = float("nan")
    with pytest.raises(FloatingPointError, match="non-finite"):
        m.step(Un, 0.0, 1.0)
    with pytest.raises(FloatingPointError, match="non-finite"):
        m.forward(Un, t_end=1.0)
    with pytest.raises(RuntimeError, match="max_steps=2 reached"):
        m.forward(U, t_end=10.0, max_steps=2)
    with pytest.raises(ValueError, match="cannot be negative"):
        m.forward(U, t_end=-1.0)
    with pytest.raises(RuntimeError, match="ROCm device only"):
        m.step(U.cpu(), 0.0, 1.0)


def test_torch_compile_wrapper_runs_the_library_path(cuda_device):
    """torch.compile(model) must not break the drop-in (the library calls run eagerly)."""
    n = 24
    m = swe.SWE2D(torch.zeros(1, 1, n, n), torch.zeros(1, 1, n, n), 0.1, 0.1, train_dem=False).to(cuda_device)
    U = torch.zeros(1, 3, n, n, device=cuda_device)
    U[:, 0] = 1.0
    U[:, 0, : n // 2] = 1.5
    compiled = torch.compile(m)
    a, ha = compiled(U, t_end=0.05)
    b, hb = m(U, t_end=0.05)
    assert torch.equal(a, b) and torch.equal(ha, hb)


def test_results_do_not_depend_on_strip_height(cuda_device):
    """The strip height S only tiles the work: every S gives the same bits."""
    sc = scenarios.coastal_runup(nx=68, ny=43, dx=0.08)
    cfg = swe.SWEConfig(boundary_left="water_level", boundary_right="wall", boundary_top="wall",
                        boundary_bottom="wall")
    m = swe.SWE2D(torch.tensor(sc["bed"], dtype=torch.float32), torch.tensor(sc["rough"], dtype=torch.float32),
                  0.08, 0.08, config=cfg, train_dem=False).to(cuda_device)
    U0 = torch.tensor(sc["U"], dtype=torch.float32, device=cuda_device)
    outs = []
    for rows in (1, 2, 5, 16, 43):
        m.rows_hint = rows
        with torch.no_grad():
            outs.append(m(U0, t_end=0.3, boundary_level=0.01))
    for U, h in outs[1:]:
        assert torch.equal(U, outs[0][0]) and torch.equal(h, outs[0][1])


def test_autotune_picks_a_candidate(cuda_device):
    from diffswe2d_amd import autotune

    sc = scenarios.coastal_runup(nx=68, ny=43, dx=0.08)
    m = swe.SWE2D(torch.tensor(sc["bed"], dtype=torch.float32), torch.tensor(sc["rough"], dtype=torch.float32),
                  0.08, 0.08, train_dem=False).to(cuda_device)
    U0 = torch.tensor(sc["U"], dtype=torch.float32, device=cuda_device)
    best, rep = autotune.tune_rows(m, U0, candidates=(2, 4, 8), reps=2, save=False)
    print(f"tuned rows {best}: {rep}")
    assert best in (2, 4, 8) and m.rows_hint == best
    assert set(rep) == {2, 4, 8} and all(v > 0 for v in rep.values())


def test_dynamic_boundary_series_matches_oracle_fp64(cuda_device):
    """set_dynamic_boundary: the level is interpolated at every step's start time (solver.py:391-415)."""
    import oracle.swe_oracle as OO

    sc = scenarios.coastal_runup(nx=68, ny=43, dx=0.08)
    cfg = swe.SWEConfig(boundary_left="water_level", boundary_right="wall", boundary_top="wall",
                        boundary_bottom="wall")
    bed = torch.tensor(sc["bed"])
    rough = torch.tensor(sc["rough"])
    m = swe.SWE2D(bed, rough, 0.08, 0.08, config=cfg, train_dem=False).to(cuda_device)
    times = torch.tensor([0.0, 0.2, 0.5, 1.0], dtype=torch.float64)
    levels = torch.tensor([0.0, 0.02, -0.01, 0.03], dtype=torch.float64)
    m.set_dynamic_boundary(times, levels)
    U0 = torch.tensor(sc["U"])
    old = OO.SQRT_MODE
    OO.SQRT_MODE = "ieee"
    try:
        ref, href = OO.forward(U0, bed, rough, cfg, 0.08, 0.08, 0.6, start_time=0.1, dynamic=(times, levels))
    finally:
        OO.SQRT_MODE = old
    with torch.no_grad():
        U, h = m(U0.to(cuda_device), t_end=0.6, start_time=0.1)
    e = relerr(U, ref)
    eh = relerr(h, href)
    print(f"dynamic boundary: rel-L2 U {e:.2e}, hmax {eh:.2e}")
    assert e <= 1e-10 and eh <= 1e-10
    m.clear_dynamic_boundary()
    assert not m.use_dynamic_bc


@pytest.mark.parametrize("name", ["f32", "f64"])
def test_stable_dt_matches_golden(golden, cuda_device, name):
    """SWE2D.stable_dt against the reference's wave speeds (solver.py:537-545) on the golden step case."""
    d = golden("steps")
    m, cfg = make_model(d, 1, DT[name], cuda_device)
    ax, ay = (float(v) for v in d[f"s1_{name}_axay"])
    ref = cfg.cfl / max(ax / 0.1 + ay / 0.1, cfg.epsilon)
    U = torch.tensor(d["U0"], dtype=DT[name], device=cuda_device)
    got = float(m.stable_dt(U, boundary_level=torch.tensor(STEP_CASES[1]["level"], dtype=DT[name])))
    print(f"{name}: stable dt {got:.9e} vs {ref:.9e}")
    assert abs(got - ref) <= (1e-6 if name == "f32" else 1e-13) * ref


@pytest.mark.parametrize("name", ["f32", "f64"])
def test_water_volume_fp64_sum(cuda_device, name):
    """SWE2D.water_volume: FP64 sum of h * dx * dy per member, the same bits on every call."""
    dtype = DT[name]
    g = torch.Generator().manual_seed(11)
    for B, ny, nx in ((2, 37, 53), (1, 700, 1000), (3, 2, 2)):
        bed = torch.zeros(1, 1, ny, nx, dtype=dtype)
        m = swe.SWE2D(bed, torch.zeros_like(bed), 0.3, 0.7, train_dem=False).to(cuda_device)
        U = torch.rand(B, 3, ny, nx, generator=g, dtype=dtype).to(cuda_device)
        v1, v2 = m.water_volume(U), m.water_volume(U)
        ref = U[:, 0].double().cpu().numpy().sum(axis=(1, 2)) * 0.3 * 0.7
        err = float(np.max(np.abs(v1.cpu().numpy() - ref) / ref))
        print(f"{name} {B}x{ny}x{nx}: volume rel err {err:.2e}")
        assert v1.dtype == torch.float64 and v1.shape == (B,)
        assert torch.equal(v1, v2) and err <= 1e-13


@pytest.mark.parametrize("name,bar", [("f64", 1e-12), ("f32", 1e-6)])
def test_volume_is_conserved_on_a_periodic_basin(cuda_device, name, bar):
    """All-periodic boundaries, no dry cells, no friction: the finite-volume update conserves
    the water volume up to rounding (the FP64 sum measures it)."""
    dtype = DT[name]
    n = 96
    y, x = torch.meshgrid(torch.arange(n, dtype=torch.float64), torch.arange(n, dtype=torch.float64), indexing="ij")
    bed = (0.2 * torch.sin(2 * math.pi * x / n) * torch.cos(2 * math.pi * y / n))[None, None]
    h = 1.0 + 0.3 * torch.exp(-((x - 40) ** 2 + (y - 50) ** 2) / 60.0)[None, None] - bed
    U = torch.cat((h, 0.05 * torch.ones_like(h), torch.zeros_like(h)), dim=1).to(dtype)
    cfg = swe.SWEConfig(boundary_left="periodic", boundary_right="periodic", boundary_top="periodic",
                        boundary_bottom="periodic", apply_friction=False)
    m = swe.SWE2D(bed.to(dtype), torch.zeros_like(bed, dtype=dtype), 0.5, 0.5, config=cfg,
                  train_dem=False).to(cuda_device)
    U = U.to(cuda_device)
    v0 = float(m.water_volume(U)[0])
    with torch.no_grad():
        U, _ = m(U, t_end=2.0)
    v1 = float(m.water_volume(U)[0])
    steps = len(m.last_dt_series)
    print(f"{name}: {steps} steps, volume {v0:.15e} -> {v1:.15e}, rel change {abs(v1 - v0) / v0:.2e}")
    assert steps > 50 and float(U[:, 0].min()) > 0.1
    assert abs(v1 - v0) <= bar * v0


def test_device_division_equals_ieee_division(cuda_device):
    """The stage kernels' FP32 division (v_rcp_f32 estimate, one Newton step, two remainder
    corrections; and its host-reciprocal form for dx, dy) against the compiler's IEEE division
    on the device: bit for bit over 16 M random normal-range operand pairs (numerators
    log-uniform over [2^-100, 2^40], divisors over [2^-30, 2^30], both signs, some zeros)."""
    from diffswe2d_amd import _lib

    n = 1 << 24
    g = torch.Generator().manual_seed(3)
    ea = torch.randint(-100, 40, (n,), generator=g)
    eb = torch.randint(-30, 30, (n,), generator=g)
    a = (1 + torch.rand(n, generator=g, dtype=torch.float64)) * torch.pow(2.0, ea.double())
    b = (1 + torch.rand(n, generator=g, dtype=torch.float64)) * torch.pow(2.0, eb.double())
    a = torch.where(torch.rand(n, generator=g) < 0.5, -a, a)
    b = torch.where(torch.rand(n, generator=g) < 0.5, -b, b)
    a[torch.rand(n, generator=g) < 1 / 64] = 0.0
    a = a.float().to(cuda_device)
    b = b.float().to(cuda_device)
    q_rn, q_by, q_ieee = (torch.empty_like(a) for _ in range(3))
    _lib.check(_lib.load().swe_division_probe(_lib.ptr(a), _lib.ptr(b), n, _lib.ptr(q_rn), _lib.ptr(q_by),
                                              _lib.ptr(q_ieee), _lib.stream_ptr(cuda_device)), "division probe")
    torch.cuda.synchronize()
    bits = lambda t: t.view(torch.int32)  # noqa: E731
    zero = q_ieee == 0
    bad_rn = int(((bits(q_rn) != bits(q_ieee)) & ~zero).sum())
    bad_by = int(((bits(q_by) != bits(q_ieee)) & ~zero).sum())
    print(f"division: {n} pairs, div_rn mismatches {bad_rn}, div_by mismatches {bad_by}")
    assert bad_rn == 0 and bad_by == 0
    assert bool((q_rn[zero] == 0).all()) and bool((q_by[zero] == 0).all())


def test_device_square_root_equals_ieee_sqrt(cuda_device):
    """The kernels' FP32 square root against the compiler's IEEE sqrt on the device: bit for bit over
    16 M arguments log-uniform in [2^-96, 2^100] plus +-0, +inf and NaN."""
    from diffswe2d_amd import _lib

    n = 1 << 24
    g = torch.Generator().manual_seed(9)
    e = torch.randint(-96, 100, (n,), generator=g)
    x = ((1 + torch.rand(n, generator=g, dtype=torch.float64)) * torch.pow(2.0, e.double())).float()
    x[:4] = torch.tensor([0.0, -0.0, float("inf"), float("nan")])
    x = x.to(cuda_device)
    s_rn, s_ieee = torch.empty_like(x), torch.empty_like(x)
    _lib.check(_lib.load().swe_sqrt_probe(_lib.ptr(x), n, _lib.ptr(s_rn), _lib.ptr(s_ieee),
                                          _lib.stream_ptr(cuda_device)), "sqrt probe")
    torch.cuda.synchronize()
    bad = int((s_rn.view(torch.int32) != s_ieee.view(torch.int32))[4:].sum())
    print(f"sqrt: {n} arguments, mismatches {bad}; specials {s_rn[:4].tolist()}")
    assert bad == 0
    assert s_rn[0].item() == 0.0 and str(s_rn[1].item()) == "-0.0" and s_rn[2].item() == float("inf")
    assert math.isnan(s_rn[3].item())


def test_device_friction_power_rounds_to_nearest(cuda_device):
    """The kernels' FP32 h^third (Manning friction) on the device against h^third rounded from an FP64
    evaluation on the host, over 4 M depths log-uniform in [1e-6, 1e4]: at most one ulp away, and
    for at most 1 in 10^5 arguments; powf's own count is reported beside it."""
    from diffswe2d_amd import _lib

    n = 1 << 22
    g = torch.Generator().manual_seed(13)
    x64 = torch.exp(torch.empty(n, dtype=torch.float64).uniform_(math.log(1e-6), math.log(1e4), generator=g))
    x = x64.float()
    third = float(np.float32(1.0 / 3.0))
    want = torch.pow(x.double(), third).float()
    xd = x.to(cuda_device)
    y_fast, y_powf = torch.empty_like(xd), torch.empty_like(xd)
    _lib.check(_lib.load().swe_pow_probe(_lib.ptr(xd), n, ctypes.c_float(third), _lib.ptr(y_fast), _lib.ptr(y_powf),
                                         _lib.stream_ptr(cuda_device)), "pow probe")
    torch.cuda.synchronize()
    d_fast = (y_fast.cpu().view(torch.int32).long() - want.view(torch.int32).long()).abs()
    d_powf = (y_powf.cpu().view(torch.int32).long() - want.view(torch.int32).long()).abs()
    print(f"pow_third: {int((d_fast == 1).sum())} one ulp off, {int((d_fast > 1).sum())} further, of {n}; "
          f"powf: {int((d_powf == 1).sum())} one ulp off, {int((d_powf > 1).sum())} further")
    assert int((d_fast > 1).sum()) == 0 and int((d_fast == 1).sum()) <= n // 100_000


@pytest.mark.gpu
def test_device_lane_shifts(cuda_device):
    """The kernels' one-lane wavefront shifts (DPP moves) against the shift they stand for, per 64-lane
    wavefront of consecutive elements: up[i] = x[i - 1], dn[i] = x[i + 1], 0 in the lane with no source;
    FP64 values bit for bit (both 32-bit halves move) and FP32 values of (float)x."""
    from diffswe2d_amd import _lib

    n = 1024
    g = torch.Generator().manual_seed(5)
    x = torch.randn(n, dtype=torch.float64, generator=g) * 1e3
    xd = x.to(cuda_device)
    up64, dn64 = torch.empty_like(xd), torch.empty_like(xd)
    up32 = torch.empty(n, dtype=torch.float32, device=cuda_device)
    dn32 = torch.empty_like(up32)
    _lib.check(_lib.load().swe_lane_shift_probe(_lib.ptr(xd), n, _lib.ptr(up64), _lib.ptr(dn64), _lib.ptr(up32),
                                                _lib.ptr(dn32), _lib.stream_ptr(cuda_device)), "lane shift probe")
    torch.cuda.synchronize()
    lane = torch.arange(n) % 64
    idx = torch.arange(n)
    want_up = torch.where(lane > 0, x[(idx - 1).clamp(min=0)], 0.0)
    want_dn = torch.where(lane < 63, x[(idx + 1).clamp(max=n - 1)], 0.0)
    assert torch.equal(up64.cpu().view(torch.int64), want_up.view(torch.int64))
    assert torch.equal(dn64.cpu().view(torch.int64), want_dn.view(torch.int64))
    assert torch.equal(up32.cpu(), want_up.float())
    assert torch.equal(dn32.cpu(), want_dn.float())
