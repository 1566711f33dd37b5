"""Write golden vectors from the reference's own hot-path modules.

Run in the development container only (it reads the reference tree):

    python tests/golden/make_golden.py [--reference /root/reference/src]

The reference package's ``__init__`` and ``solver.py`` import ``io_netcdf``,
whose module top imports xarray, which this image does not have.  The script
therefore imports the six hot-path modules that need only torch
(config, numerics, riemann, reconstruction, boundary, friction) as
submodules of a namespace package pointing at the reference ``src``
directory, without executing the package ``__init__``.  Nothing is stubbed.

The solver glue of ``src/solver.py:448-663`` (ghost cells -> x/y reconstruction
-> HLL -> rhs assembly, CFL dt, positivity clean-up, RK2 midpoint, friction)
is restated below, line for line with the file it cites, and calls the
reference functions for every numerical kernel.  The fixtures are small
``.npz`` files next to this script; each stores inputs and expected outputs.
"""
from __future__ import annotations

import argparse
import importlib
import sys
import time
import types
from pathlib import Path

import numpy as np
import torch

HERE = Path(__file__).resolve().parent
REPO = HERE.parent.parent
sys.path.insert(0, str(REPO))

from diffswe2d_amd import scenarios  # noqa: E402


def load_reference(src: Path):
    pkg = types.ModuleType("_swe_reference")
    pkg.__path__ = [str(src)]
    sys.modules["_swe_reference"] = pkg
    names = ("config", "numerics", "riemann", "reconstruction", "boundary", "friction")
    return types.SimpleNamespace(**{n: importlib.import_module(f"_swe_reference.{n}") for n in names})


# ---------------------------------------------------------------------------
# solver glue, restated from src/solver.py (see module docstring)
# ---------------------------------------------------------------------------
class RefStep:
    def __init__(self, ref, cfg, bed, rough, dx, dy):
        self.ref, self.cfg, self.bed, self.rough = ref, cfg, bed, rough
        self.dx, self.dy = float(dx), float(dy)

    def level(self, value, U):  # solver.py:378-389
        return value.to(dtype=U.dtype) if torch.is_tensor(value) else U.new_tensor(value)

    def spatial_operator(self, U, level, wave=False):  # solver.py:448-531
        cfg, r = self.cfg, self.ref
        B, _, ny, nx = U.shape
        bed = self.bed if self.bed.shape[0] == B else self.bed.expand(B, -1, -1, -1)
        Up, zp = r.boundary.add_ghost_cells(
            U, bed, ng=2, boundary_left=cfg.boundary_left, boundary_right=cfg.boundary_right,
            boundary_top=cfg.boundary_top, boundary_bottom=cfg.boundary_bottom,
            constant_value=cfg.constant_value, constant_level=level)
        args = (cfg.gravity, cfg.epsilon, cfg.dry_depth, cfg.limiter_theta, cfg.water_slope_reset)
        ULx, URx, sxL, sxR = r.reconstruction.bflood_cn_reconstruction(Up, zp, -1, *args)
        ULy, URy, syL, syR = r.reconstruction.bflood_cn_reconstruction(Up, zp, -2, *args)
        Fx, ax = r.riemann.bflood_hllc_flux(ULx, URx, "x", cfg.gravity, cfg.epsilon, cfg.dry_depth)
        Gy, ay = r.riemann.bflood_hllc_flux(ULy, URy, "y", cfg.gravity, cfg.epsilon, cfg.dry_depth)
        Fx_left, Fx_right = Fx.clone(), Fx.clone()
        Gy_bottom, Gy_top = Gy.clone(), Gy.clone()
        Fx_left[:, 1:2] -= sxL
        Fx_right[:, 1:2] -= sxR
        Gy_bottom[:, 2:3] -= syL
        Gy_top[:, 2:3] -= syR
        rhs = (Fx_right[:, :, 2:2 + ny, 1:1 + nx] - Fx_left[:, :, 2:2 + ny, 2:2 + nx]) / self.dx
        rhs += (Gy_top[:, :, 1:1 + ny, 2:2 + nx] - Gy_bottom[:, :, 2:2 + ny, 2:2 + nx]) / self.dy
        h = U[:, 0:1]
        rain = torch.zeros_like(h)
        loss = torch.clamp(torch.zeros_like(h), min=0.0)
        rhs = torch.cat((rhs[:, 0:1] + rain - loss, rhs[:, 1:2], rhs[:, 2:3]), dim=1)
        if wave:
            return rhs, ax[:, :, 2:2 + ny, 1:2 + nx].amax(), ay[:, :, 1:2 + ny, 2:2 + nx].amax()
        return rhs

    def clean(self, old, cand):  # solver.py:547-570
        cfg = self.cfg
        if cfg.positivity_limiter:
            delta = cand - old
            dh = delta[:, 0:1]
            safe_change = torch.clamp(-dh, min=cfg.epsilon)
            theta = torch.where(dh < 0.0, torch.clamp(old[:, 0:1] / safe_change, max=1.0),
                                torch.ones_like(dh))
            cand = old + theta * delta
        h = torch.clamp(cand[:, 0:1], min=0.0)
        wet = h > cfg.dry_depth
        zero = torch.zeros_like(h)
        return torch.cat((h, torch.where(wet, cand[:, 1:2], zero), torch.where(wet, cand[:, 2:3], zero)), 1)

    def step(self, U, max_dt, level):  # solver.py:572-663
        cfg = self.cfg
        max_dt = U.new_tensor(max_dt)
        k1, ax, ay = self.spatial_operator(U, level, wave=True)
        den = torch.clamp(ax / self.dx + ay / self.dy, min=cfg.epsilon)
        dt = torch.minimum(cfg.cfl / den, max_dt).detach()
        assert bool(torch.isfinite(dt)) and float(dt) > 0.0
        pred = self.clean(U, U + 0.5 * dt * k1)
        k2 = self.spatial_operator(pred, level)
        new = self.clean(U, U + dt * k2)
        if cfg.apply_friction:
            rough = self.rough if self.rough.shape[0] == U.shape[0] else self.rough.expand(U.shape[0], -1, -1, -1)
            kw = dict(dry_depth=cfg.dry_depth, epsilon=cfg.epsilon)
            if cfg.frictionmodel == "manning":
                new = self.ref.friction.apply_manning_friction(new, rough, dt, **kw)
            else:
                new = self.ref.friction.apply_roughness_length_friction(new, rough, dt, **kw)
        return new, float(dt)

    def advance(self, U, duration, level, dt_log):  # solver.py:719-779, without hmax/gauges
        elapsed, tol = 0.0, max(1e-12, 1e-12 * abs(duration))
        while True:
            remaining = duration - elapsed
            if remaining <= tol:
                return U
            U, dtv = self.step(U, remaining, self.level(level, U))
            elapsed += dtv
            if elapsed > duration and elapsed - duration <= tol:
                elapsed = duration
            dt_log.append(dtv)


# ---------------------------------------------------------------------------
# fixture writers
# ---------------------------------------------------------------------------
def rand_state(rng, B, ny, nx, dry_frac=0.25, dtype=np.float64):
    h = rng.uniform(0.0, 1.5, (B, 1, ny, nx))
    dry = rng.uniform(size=h.shape) < dry_frac
    h = np.where(dry, rng.choice([0.0, 1e-7, -1e-3, 5e-7], size=h.shape), h)
    hu = rng.normal(0.0, 0.6, h.shape)
    hv = rng.normal(0.0, 0.6, h.shape)
    return np.concatenate([h, hu, hv], axis=1).astype(dtype)


def write_limiter(ref, out):
    rng = np.random.default_rng(10)
    data = {}
    q = rng.normal(size=(2, 1, 9, 11))
    q[0, 0, 2:5, 3:7] = 0.25          # flat patch: every difference is zero
    q[1, 0, :, 4] = q[1, 0, :, 5]     # equal neighbours: sign ties
    q[1, 0, 6, :] = np.linspace(-1, 1, 11)  # linear row: centred slope wins
    for name, dt in (("f32", torch.float32), ("f64", torch.float64)):
        t = torch.tensor(q, dtype=dt)
        data[f"q_{name}"] = t.numpy()
        for dim in (-1, -2):
            data[f"d{-dim}_{name}"] = ref.numerics.limited_increment(t, dim, 1.3).numpy()
    np.savez_compressed(out / "limiter.npz", **data)


def write_hll(ref, out):
    rng = np.random.default_rng(11)
    data = {}
    UL = rand_state(rng, 2, 6, 7)
    UR = rand_state(rng, 2, 6, 7)
    UR[0, :, 0, :] = UL[0, :, 0, :]              # identical states: consistency
    UR[1, 1, 1, :] = -UL[1, 1, 1, :]             # mirrored normal momentum: wall faces
    UR[1, 0, 1, :] = UL[1, 0, 1, :]
    for name, dt in (("f32", torch.float32), ("f64", torch.float64)):
        a, b = torch.tensor(UL, dtype=dt), torch.tensor(UR, dtype=dt)
        data[f"UL_{name}"], data[f"UR_{name}"] = a.numpy(), b.numpy()
        h, u, v, c = ref.riemann.primitive(a, 9.81, 1e-12, 1e-6)
        data[f"prim_{name}"] = torch.cat((h, u, v, c), 1).numpy()
        for normal in ("x", "y"):
            f, s = ref.riemann.bflood_hllc_flux(a, b, normal, 9.81, 1e-12, 1e-6)
            data[f"flux_{normal}_{name}"], data[f"speed_{normal}_{name}"] = f.numpy(), s.numpy()
    np.savez_compressed(out / "hll.npz", **data)


def write_recon(ref, out):
    rng = np.random.default_rng(12)
    data = {}
    Up = rand_state(rng, 2, 10, 12)
    Up[:, 0] = np.abs(Up[:, 0]) * (rng.uniform(size=Up[:, 0].shape) > 0.2)
    zp = np.cumsum(rng.normal(0, 0.2, (2, 1, 10, 12)), axis=-1) + np.cumsum(rng.normal(0, 0.2, (2, 1, 10, 12)), axis=-2)
    Up[1, 0, 3:6, 3:8] = np.maximum(0.4 - zp[1, 0, 3:6, 3:8], 0.0)  # lake at rest patch
    Up[1, 1:, 3:6, 3:8] = 0.0
    for name, dt in (("f32", torch.float32), ("f64", torch.float64)):
        U, z = torch.tensor(Up, dtype=dt), torch.tensor(zp, dtype=dt)
        data[f"Up_{name}"], data[f"zp_{name}"] = U.numpy(), z.numpy()
        for dim in (-1, -2):
            for wsr in (True, False):
                res = ref.reconstruction.bflood_cn_reconstruction(U, z, dim, 9.81, 1e-12, 1e-6, 1.3, wsr)
                data[f"r{-dim}_{int(wsr)}_{name}"] = torch.cat(res, 1).numpy()
    np.savez_compressed(out / "recon.npz", **data)


GHOST_CASES = [
    ("wall", "wall", "wall", "wall"),
    ("water_level", "wall", "transmissive", "periodic"),
    ("periodic", "periodic", "periodic", "periodic"),
    ("constant", "absorbing", "water_level", "absorbing"),
    ("absorbing", "constant", "absorbing", "water_level"),
    ("transmissive", "transmissive", "transmissive", "transmissive"),
    ("periodic", "water_level", "wall", "constant"),
]


def write_ghost(ref, out):
    rng = np.random.default_rng(13)
    data = {}
    U = rand_state(rng, 2, 5, 6)
    bed1 = rng.normal(0, 0.3, (1, 1, 5, 6))
    bed2 = rng.normal(0, 0.3, (2, 1, 5, 6))
    level = np.array([0.35, -0.1])
    data["U"], data["bed1"], data["bed2"], data["level"] = U, bed1, bed2, level
    for i, case in enumerate(GHOST_CASES):
        for j, (bed, lvl) in enumerate(((bed1, 0.2), (bed2, level))):
            for name, dt in (("f32", torch.float32), ("f64", torch.float64)):
                Up, zp = ref.boundary.add_ghost_cells(
                    torch.tensor(U, dtype=dt), torch.tensor(bed, dtype=dt), 2, *case,
                    constant_value=0.07, constant_level=torch.tensor(lvl, dtype=dt))
                data[f"c{i}_b{j}_{name}_Up"], data[f"c{i}_b{j}_{name}_zp"] = Up.numpy(), zp.numpy()
    np.savez_compressed(out / "ghost.npz", **data)


def write_friction(ref, out):
    rng = np.random.default_rng(14)
    data = {}
    U = rand_state(rng, 2, 8, 9)
    U[0, 1:, 0, :] = 0.0                 # resting wet cells: zero speed branch
    n = rng.uniform(-0.01, 0.05, (1, 1, 8, 9))
    z0 = rng.uniform(1e-4, 0.2, (1, 1, 8, 9))
    data["U"], data["n"], data["z0"] = U, n, z0
    for name, dt in (("f32", torch.float32), ("f64", torch.float64)):
        t = torch.tensor(U, dtype=dt)
        d = torch.tensor(0.013, dtype=dt)
        data[f"manning_{name}"] = ref.friction.apply_manning_friction(
            t, torch.tensor(n, dtype=dt).expand(2, -1, -1, -1), d).numpy()
        data[f"rough_{name}"] = ref.friction.apply_roughness_length_friction(
            t, torch.tensor(z0, dtype=dt).expand(2, -1, -1, -1), d).numpy()
    np.savez_compressed(out / "friction.npz", **data)


STEP_CASES = [
    dict(bc=("wall", "wall", "wall", "wall"), friction=True, level=0.0),
    dict(bc=("water_level", "wall", "transmissive", "water_level"), friction=True, level=0.45),
    dict(bc=("periodic", "periodic", "absorbing", "constant"), friction=False, level=0.3),
    dict(bc=("transmissive", "absorbing", "periodic", "periodic"), friction=True, level=0.2),
]


def write_steps(ref, out):
    """Small wet/dry states through several steps, every BC type, B = 2."""
    rng = np.random.default_rng(15)
    data = {}
    ny, nx = 12, 10
    x = (np.arange(nx) + 0.5) * 0.1
    y = (np.arange(ny) + 0.5) * 0.1
    X, Y = np.meshgrid(x, y)
    bed = (0.25 * np.sin(3 * X) * np.cos(2 * Y) + 0.3 * X - 0.2)[None, None]
    eta = 0.3 + 0.15 * np.exp(-((X - 0.5) ** 2 + (Y - 0.6) ** 2) / 0.02)
    h = np.maximum(eta - bed[0, 0], 0.0)
    U = np.stack([np.stack([h, 0.05 * h, -0.03 * h]),
                  np.stack([h[::-1, :] * 0.8, np.zeros_like(h), 0.1 * h[::-1, :]])])
    rough = rng.uniform(0.0, 0.04, (1, 1, ny, nx))
    data["U0"], data["bed"], data["rough"] = U, bed, rough
    for i, case in enumerate(STEP_CASES):
        for name, dt in (("f32", torch.float32), ("f64", torch.float64)):
            cfg = ref.config.SWEConfig(boundary_left=case["bc"][0], boundary_right=case["bc"][1],
                                       boundary_top=case["bc"][2], boundary_bottom=case["bc"][3],
                                       apply_friction=case["friction"], constant_value=0.05)
            st = RefStep(ref, cfg, torch.tensor(bed, dtype=dt), torch.tensor(rough, dtype=dt), 0.1, 0.1)
            s = torch.tensor(U, dtype=dt)
            lvl = torch.tensor(case["level"], dtype=dt)
            k1, ax, ay = st.spatial_operator(s, lvl, wave=True)
            data[f"s{i}_{name}_rhs"] = k1.numpy()
            data[f"s{i}_{name}_axay"] = np.array([float(ax), float(ay)])
            dts, states = [], []
            for _ in range(12):
                s, dtv = st.step(s, 1e30, lvl)
                dts.append(dtv)
                states.append(s.numpy())
            data[f"s{i}_{name}_dt"] = np.array(dts)
            data[f"s{i}_{name}_U"] = np.stack(states)
    np.savez_compressed(out / "steps.npz", **data)


def write_dam_break(ref, out):
    sc = scenarios.dam_break()
    data = {}
    for name, dt in (("f32", torch.float32), ("f64", torch.float64)):
        cfg = ref.config.SWEConfig(boundary_left="wall", boundary_right="wall", boundary_top="wall",
                                   boundary_bottom="wall", apply_friction=True)
        st = RefStep(ref, cfg, torch.tensor(sc["bed"], dtype=dt), torch.tensor(sc["rough"], dtype=dt),
                     sc["dx"], sc["dy"])
        s = torch.tensor(sc["U"], dtype=dt)
        lvl = s.new_tensor(0.0)
        dts, hmin = [], []
        t0 = time.time()
        for k in range(sc["steps"]):
            s, dtv = st.step(s, 1e30, lvl)
            dts.append(dtv)
            hmin.append(float(s[:, 0].min()))
            if k in (0, 9, 99):
                data[f"U{k + 1}_{name}"] = s.numpy().astype(np.float32 if name == "f32" else np.float64)
        print(f"dam break {name}: {time.time() - t0:.1f} s", flush=True)
        data[f"dt_{name}"] = np.array(dts)
        data[f"hmin_{name}"] = np.array(hmin)
        data[f"U500_{name}"] = s.numpy()
    np.savez_compressed(out / "dam_break.npz", **data)


def write_runup(ref, out, nx=136, ny=86, dx=0.04, n_chunks=40):
    """Reduced coastal run-up (same 5.44 x 3.44 m basin, coarser grid), 2 s of flow."""
    sc = scenarios.coastal_runup(nx=nx, ny=ny, dx=dx)
    data = dict(U0=sc["U"], bed=sc["bed"], rough=sc["rough"], levels=sc["chunk_levels"][:n_chunks],
                gauges=np.array(sc["gauges"]))
    for name, dt in (("f32", torch.float32), ("f64", torch.float64)):
        cfg = ref.config.SWEConfig(boundary_left="water_level", boundary_right="wall",
                                   boundary_top="wall", boundary_bottom="wall")
        st = RefStep(ref, cfg, torch.tensor(sc["bed"], dtype=dt), torch.tensor(sc["rough"], dtype=dt), dx, dx)
        s = torch.tensor(sc["U"], dtype=dt)
        dts, gauge, masks = [], [], []
        t0 = time.time()
        for k in range(n_chunks):
            s = st.advance(s, sc["sample_dt"], torch.tensor(sc["chunk_levels"][k], dtype=dt), dts)
            gauge.append([float(s[0, 0, iy, ix]) for iy, ix in sc["gauges"]])
            if (k + 1) % 10 == 0:
                masks.append((s[0, 0] > cfg.dry_depth).numpy())
        print(f"run-up {name}: {len(dts)} steps, {time.time() - t0:.1f} s", flush=True)
        data[f"dt_{name}"] = np.array(dts)
        data[f"gauge_{name}"] = np.array(gauge)
        data[f"mask_{name}"] = np.packbits(np.stack(masks))
        data[f"U_{name}"] = s.numpy()
    np.savez_compressed(out / "runup.npz", **data)


def write_runup_full(ref, out, name, n_chunks=450, mask_every=50):
    """The full coastal run-up: 544 x 344 at dx = 0.01 m, 450 chunks of 0.05 s (22.5 s).

    Driven like the reduced run (one duration-driven forward per chunk at the
    chunk's boundary level, solver.py:719-779).  Stores every step's dt, the six
    gauge depths at every chunk end, the wet/dry mask every ``mask_every``
    chunks and the final state.  One precision per file
    (``runup_full_f32.npz`` / ``runup_full_f64.npz``) so both can be written at
    once by two processes.
    """
    dt = dict(f32=torch.float32, f64=torch.float64)[name]
    sc = scenarios.coastal_runup()
    cfg = ref.config.SWEConfig(boundary_left="water_level", boundary_right="wall",
                               boundary_top="wall", boundary_bottom="wall")
    st = RefStep(ref, cfg, torch.tensor(sc["bed"], dtype=dt), torch.tensor(sc["rough"], dtype=dt),
                 sc["dx"], sc["dy"])
    s = torch.tensor(sc["U"], dtype=dt)
    dts, chunk_steps, gauge, masks = [], [], [], []
    t0 = time.time()
    for k in range(n_chunks):
        s = st.advance(s, sc["sample_dt"], torch.tensor(sc["chunk_levels"][k], dtype=dt), dts)
        chunk_steps.append(len(dts))
        gauge.append([float(s[0, 0, iy, ix]) for iy, ix in sc["gauges"]])
        if (k + 1) % mask_every == 0:
            masks.append((s[0, 0] > cfg.dry_depth).numpy())
            print(f"run-up full {name}: chunk {k + 1}/{n_chunks}, {len(dts)} steps, "
                  f"{time.time() - t0:.0f} s", flush=True)
    data = dict(levels=sc["chunk_levels"][:n_chunks], gauges=np.array(sc["gauges"]),
                mask_every=np.array(mask_every), dt=np.array(dts), chunk_steps=np.array(chunk_steps),
                gauge=np.array(gauge), mask=np.packbits(np.stack(masks)), U=s.numpy())
    np.savez_compressed(out / f"runup_full_{name}.npz", **data)


def write_runup_full_ieee(ref, out, n_chunks=450, mask_every=50):
    """The full run-up in FP32 through the oracle (oracle/swe_oracle.py) with
    correctly rounded square roots, stored like ``runup_full_f32.npz``.

    torch's CPU float32 sqrt is one ulp off for ~0.7 % of arguments and its
    float32 pow (Manning's h^(1/3)) for ~1.4 %; the GPU rounds both correctly.
    Over the run-up's ~11,000 steps that difference moves the reference's FP32
    trajectory by a few 1e-7 and, at the moving shoreline, can flip a cell
    between wet and dry.  This run separates the two effects: it is the
    reference's algorithm (the oracle, pinned to the reference's modules by the
    fixtures above) with correctly rounded square roots and friction powers."""
    from oracle import swe_oracle as O

    sc = scenarios.coastal_runup()
    cfg = ref.config.SWEConfig(boundary_left="water_level", boundary_right="wall",
                               boundary_top="wall", boundary_bottom="wall")
    bed = torch.tensor(sc["bed"], dtype=torch.float32)
    rough = torch.tensor(sc["rough"], dtype=torch.float32)
    s = torch.tensor(sc["U"], dtype=torch.float32)
    dts, chunk_steps, gauge, masks = [], [], [], []
    O.SQRT_MODE = "ieee"
    t0 = time.time()
    with torch.no_grad():
        for k in range(n_chunks):
            s, _ = O.forward(s, bed, rough, cfg, sc["dx"], sc["dy"], sc["sample_dt"], start_time=k * sc["sample_dt"],
                             boundary_level=float(sc["chunk_levels"][k]), dt_log=dts)
            chunk_steps.append(len(dts))
            gauge.append([float(s[0, 0, iy, ix]) for iy, ix in sc["gauges"]])
            if (k + 1) % mask_every == 0:
                masks.append((s[0, 0] > cfg.dry_depth).numpy())
                print(f"run-up full (oracle, IEEE sqrt): chunk {k + 1}/{n_chunks}, {len(dts)} steps, "
                      f"{time.time() - t0:.0f} s", flush=True)
    data = dict(levels=sc["chunk_levels"][:n_chunks], gauges=np.array(sc["gauges"]),
                mask_every=np.array(mask_every), dt=np.array(dts), chunk_steps=np.array(chunk_steps),
                gauge=np.array(gauge), mask=np.packbits(np.stack(masks)), U=s.numpy())
    np.savez_compressed(out / "runup_full_ieee_f32.npz", **data)


def bspline_basis(sample_times, n_controls, degree=3):
    """Clamped B-spline basis by the Cox-de Boor recursion.

    Restated from Examples/inverse/.../bc_inversion_train.py:189-231.
    """
    t0, t1 = sample_times[0], sample_times[-1]
    n_int = n_controls - degree - 1
    interior = (torch.linspace(t0, t1, n_int + 2, dtype=sample_times.dtype)[1:-1]
                if n_int > 0 else sample_times.new_empty(0))
    knots = torch.cat((t0.repeat(degree + 1), interior, t1.repeat(degree + 1)))
    x = sample_times[:, None]
    basis = ((x >= knots[:-1]) & (x < knots[1:])).to(sample_times.dtype)
    eps = torch.finfo(x.dtype).eps
    for order in range(1, degree + 1):
        cols = knots.numel() - order - 1
        ld = knots[order:order + cols] - knots[:cols]
        rd = knots[order + 1:order + 1 + cols] - knots[1:1 + cols]
        left = torch.where(ld > 0, (x - knots[:cols]) / ld.clamp_min(eps), 0.0)
        right = torch.where(rd > 0, (knots[order + 1:order + 1 + cols] - x) / rd.clamp_min(eps), 0.0)
        basis = left * basis[:, :cols] + right * basis[:, 1:cols + 1]
    basis[-1].zero_()
    basis[-1, -1] = 1.0
    return basis


def write_gradient(ref, out, nx=68, ny=43, dx=0.08, n_chunks=10, n_controls=64):
    """Reference-autograd gradient of a gauge-misfit loss w.r.t. spline controls.

    Reduced horizon (10 chunks of 0.2 s) on the coarse run-up domain; the level
    of chunk k is series[k] (bc_inversion_train.py:332-341), the observation is
    h + z at each gauge cell after each chunk (:401-410), and the loss is the
    mean over gauges of the per-gauge MSE (:427-431).  No state-gradient clip.
    """
    sc = scenarios.coastal_runup(nx=nx, ny=ny, dx=dx)
    chunk = 0.2
    data = dict(U0=sc["U"], bed=sc["bed"], rough=sc["rough"], gauges=np.array(sc["gauges"]))
    rng = np.random.default_rng(4)
    for name, dt in (("f64", torch.float64), ("f32", torch.float32)):
        times = torch.arange(n_chunks, dtype=dt) * chunk
        basis = bspline_basis(times, n_controls)
        truth_levels = torch.tensor(scenarios.pulse_level(sc["pulses"], times.numpy().astype(np.float64) + 2.0), dtype=dt)
        cfg = ref.config.SWEConfig(boundary_left="water_level", boundary_right="wall",
                                   boundary_top="wall", boundary_bottom="wall")
        st = RefStep(ref, cfg, torch.tensor(sc["bed"], dtype=dt), torch.tensor(sc["rough"], dtype=dt), dx, dx)
        bed = torch.tensor(sc["bed"], dtype=dt)

        def run(series, dts=None):
            s = torch.tensor(sc["U"], dtype=dt)
            obs = []
            for k in range(n_chunks):
                s = st.advance(s, chunk, series[k], dts if dts is not None else [])
                obs.append(torch.stack([s[0, 0, iy, ix] + bed[0, 0, iy, ix] for iy, ix in sc["gauges"]]))
            return torch.stack(obs)

        with torch.no_grad():
            target = run(truth_levels)
        controls = torch.tensor(rng.uniform(0.0, 0.02, n_controls) if name == "f64" else data["controls"],
                                dtype=dt, requires_grad=True)
        dts = []
        t0 = time.time()
        obs = run(basis @ controls, dts)
        loss = ((obs - target) ** 2).mean(dim=0).mean()
        loss.backward()
        print(f"gradient {name}: {len(dts)} steps, {time.time() - t0:.1f} s", flush=True)
        data["controls"] = controls.detach().numpy().astype(np.float64)
        data[f"basis_{name}"] = basis.numpy()
        data[f"target_{name}"] = target.numpy()
        data[f"loss_{name}"] = np.array(float(loss))
        data[f"grad_{name}"] = controls.grad.numpy()
        data[f"dt_{name}"] = np.array(dts)
    np.savez_compressed(out / "gradient.npz", **data)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--reference", default="/root/reference/src")
    ap.add_argument("--only", default="")
    ap.add_argument("--threads", type=int, default=8)
    a = ap.parse_args()
    torch.set_num_threads(a.threads)
    ref = load_reference(Path(a.reference))
    writers = dict(limiter=write_limiter, hll=write_hll, recon=write_recon, ghost=write_ghost,
                   friction=write_friction, steps=write_steps, dam_break=write_dam_break,
                   runup=write_runup, gradient=write_gradient,
                   runup_full_f32=lambda r, o: write_runup_full(r, o, "f32"),
                   runup_full_f64=lambda r, o: write_runup_full(r, o, "f64"),
                   runup_full_ieee=write_runup_full_ieee)
    # the full run-up takes 10-15 minutes per run: only on request
    on_request = ("runup_full_f32", "runup_full_f64", "runup_full_ieee")
    for name, fn in writers.items():
        if (a.only and name not in a.only.split(",")) or (not a.only and name in on_request):
            continue
        t0 = time.time()
        fn(ref, HERE)
        print(f"{name}: {time.time() - t0:.1f} s", flush=True)


if __name__ == "__main__":
    main()
