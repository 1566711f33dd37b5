"""The CPU oracle against the reference's golden vectors, closed forms and invariants.

Golden vectors come from the reference's own modules (tests/golden/make_golden.py).
FP32 comparisons are bitwise: the oracle performs the same ATen CPU operations
in the same order as the reference.
"""
import math

import numpy as np
import pytest
import torch

from oracle import swe_oracle as O

G, EPS, DRY = 9.81, 1e-12, 1e-6
DT = {"f32": torch.float32, "f64": torch.float64}


class Cfg:
    gravity, dry_depth, epsilon, cfl, limiter_theta = 9.81, 1e-6, 1e-12, 0.5, 1.3
    boundary_left = boundary_right = boundary_top = boundary_bottom = "wall"
    constant_value, constant_level = 0.0, 0.0
    water_slope_reset = positivity_limiter = apply_friction = True
    frictionmodel = "manning"

    def __init__(self, **kw):
        for k, v in kw.items():
            setattr(self, k, v)


def exact(a, b):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape
    np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("name", ["f32", "f64"])
def test_limiter_golden(golden, name):
    d = golden("limiter")
    q = torch.tensor(d[f"q_{name}"])
    exact(O.limited_increment(q, -1, 1.3), d[f"d1_{name}"])
    exact(O.limited_increment(q, -2, 1.3), d[f"d2_{name}"])


def test_limiter_closed_forms():
    # constant -> 0, linear -> the centred slope (theta > 1), extremum -> 0, odd symmetry
    x = torch.arange(9, dtype=torch.float64)[None, None, None, :]
    assert torch.all(O.limited_increment(torch.full_like(x, 3.0), -1) == 0)
    lin = O.limited_increment(0.7 * x - 2.0, -1)
    assert torch.allclose(lin[..., 1:-1], torch.full_like(lin[..., 1:-1], 0.7))
    bump = torch.tensor([0.0, 1.0, 3.0, 1.0, 0.0], dtype=torch.float64)[None, None, None]
    assert O.limited_increment(bump, -1)[..., 2] == 0
    r = torch.randn(1, 1, 5, 7, dtype=torch.float64)
    assert torch.equal(O.limited_increment(-r, -1), -O.limited_increment(r, -1))


@pytest.mark.parametrize("name", ["f32", "f64"])
def test_hll_golden(golden, name):
    d = golden("hll")
    UL, UR = torch.tensor(d[f"UL_{name}"]), torch.tensor(d[f"UR_{name}"])
    exact(torch.cat(O.primitive(UL, G, EPS, DRY), 1), d[f"prim_{name}"])
    for normal in ("x", "y"):
        f, s = O.hll_flux(UL, UR, normal, G, EPS, DRY)
        exact(f, d[f"flux_{normal}_{name}"])
        exact(s, d[f"speed_{normal}_{name}"])


def test_hll_closed_forms():
    # consistency F(U,U) = physical flux, and a symmetric wall face carries no mass
    U = torch.tensor([[[[1.3]], [[0.4]], [[-0.2]]]], dtype=torch.float64)
    f, _ = O.hll_flux(U, U, "x", G, EPS, DRY)
    h, u, v = 1.3, 0.4 / 1.3, -0.2 / 1.3
    phys = [h * u, h * u * u + 0.5 * G * h * h, h * u * v]
    assert np.allclose(f.flatten().numpy(), phys, rtol=1e-13)
    Um = U.clone()
    Um[:, 1] = -U[:, 1]
    f, s = O.hll_flux(U, Um, "x", G, EPS, DRY)
    assert abs(float(f[0, 0])) < 1e-15
    both_dry = torch.zeros(1, 3, 1, 1, dtype=torch.float64)
    f, s = O.hll_flux(both_dry, both_dry, "y", G, EPS, DRY)
    assert torch.all(f == 0) and torch.all(s == 0)


@pytest.mark.parametrize("name", ["f32", "f64"])
def test_reconstruction_golden(golden, name):
    d = golden("recon")
    U, z = torch.tensor(d[f"Up_{name}"]), torch.tensor(d[f"zp_{name}"])
    for dim in (-1, -2):
        for wsr in (True, False):
            res = O.reconstruct(U, z, dim, G, EPS, DRY, 1.3, wsr)
            exact(torch.cat(res, 1), d[f"r{-dim}_{int(wsr)}_{name}"])


def test_lake_at_rest_rhs_vanishes():
    # C-property: flat free surface over a rough bed, zero velocity -> zero tendency
    ny, nx = 14, 11
    y, x = torch.meshgrid(torch.arange(ny, dtype=torch.float64), torch.arange(nx, dtype=torch.float64),
                          indexing="ij")
    bed = (0.3 * torch.sin(0.7 * x) * torch.cos(0.5 * y) - 0.5)[None, None]
    h = torch.clamp(0.1 - bed, min=0.0)
    U = torch.cat((h, torch.zeros_like(h), torch.zeros_like(h)), 1)
    cfg = Cfg(boundary_left="wall", boundary_right="transmissive", boundary_top="wall",
              boundary_bottom="water_level", constant_level=0.1)
    rhs = O.spatial_operator(U, bed, cfg, 0.1, 0.1)
    assert float(rhs.abs().max()) < 1e-12


GHOST_CASES = [
    ("wall", "wall", "wall", "wall"),
    ("water_level", "wall", "transmissive", "periodic"),
    ("periodic", "periodic", "periodic", "periodic"),
    ("constant", "absorbing", "water_level", "absorbing"),
    ("absorbing", "constant", "absorbing", "water_level"),
    ("transmissive", "transmissive", "transmissive", "transmissive"),
    ("periodic", "water_level", "wall", "constant"),
]


@pytest.mark.parametrize("case", range(len(GHOST_CASES)))
@pytest.mark.parametrize("name", ["f32", "f64"])
def test_ghost_golden(golden, case, name):
    d = golden("ghost")
    U = torch.tensor(d["U"], dtype=DT[name])
    for j, (bed, lvl) in enumerate(((d["bed1"], 0.2), (d["bed2"], d["level"]))):
        Up, zp = O.ghost_pad(U, torch.tensor(bed, dtype=DT[name]), *GHOST_CASES[case],
                             constant_value=0.07, constant_level=torch.tensor(lvl, dtype=DT[name]))
        exact(Up, d[f"c{case}_b{j}_{name}_Up"])
        exact(zp, d[f"c{case}_b{j}_{name}_zp"])


def test_ghost_brute_force_wall_and_level():
    # hand-derived ghosts on a 3 x 4 grid
    U = torch.arange(36, dtype=torch.float64).reshape(1, 3, 3, 4)
    bed = torch.full((1, 1, 3, 4), -0.5, dtype=torch.float64)
    Up, zp = O.ghost_pad(U, bed, "wall", "water_level", "transmissive", "wall", constant_level=1.0)
    # left wall: hu ghosts mirror columns 1, 0 with sign flip; h replicated
    assert Up[0, 1, 2, 0] == -U[0, 1, 0, 1] and Up[0, 1, 2, 1] == -U[0, 1, 0, 0]
    assert Up[0, 0, 3, 0] == U[0, 0, 1, 0]
    # right water level: h = level - z = 1.5, momentum replicated
    assert Up[0, 0, 3, -1] == 1.5 and Up[0, 1, 3, -1] == U[0, 1, 1, -1]
    # bottom wall: hv mirrored
    assert Up[0, 2, -1, 3] == -U[0, 2, 1, 1] and Up[0, 2, -2, 3] == -U[0, 2, 2, 1]


@pytest.mark.parametrize("name", ["f32", "f64"])
def test_friction_golden(golden, name):
    d = golden("friction")
    U = torch.tensor(d["U"], dtype=DT[name])
    dt = torch.tensor(0.013, dtype=DT[name])
    n = torch.tensor(d["n"], dtype=DT[name]).expand(2, -1, -1, -1)
    z0 = torch.tensor(d["z0"], dtype=DT[name]).expand(2, -1, -1, -1)
    exact(O.apply_friction(U, n, dt, "manning"), d[f"manning_{name}"])
    exact(O.apply_friction(U, z0, dt, "roughness"), d[f"rough_{name}"])


def test_friction_closed_form():
    h, hu, hv, n, dt = 0.8, 0.3, -0.4, 0.03, 0.05
    U = torch.tensor([[[[h]], [[hu]], [[hv]]]], dtype=torch.float64)
    out = O.apply_friction(U, torch.full((1, 1, 1, 1), n, dtype=torch.float64), dt)
    speed = math.hypot(hu, hv) / h
    factor = 1.0 / (1.0 + 9.81 * n * n * speed * dt / h ** (4.0 / 3.0))
    assert abs(float(out[0, 1]) - hu * factor) < 1e-15
    assert abs(float(out[0, 2]) - hv * factor) < 1e-15


STEP_CASES = [
    dict(bc=("wall", "wall", "wall", "wall"), friction=True, level=0.0),
    dict(bc=("water_level", "wall", "transmissive", "water_level"), friction=True, level=0.45),
    dict(bc=("periodic", "periodic", "absorbing", "constant"), friction=False, level=0.3),
    dict(bc=("transmissive", "absorbing", "periodic", "periodic"), friction=True, level=0.2),
]


@pytest.mark.parametrize("case", range(len(STEP_CASES)))
@pytest.mark.parametrize("name", ["f32", "f64"])
def test_step_golden(golden, case, name):
    d = golden("steps")
    c = STEP_CASES[case]
    cfg = Cfg(boundary_left=c["bc"][0], boundary_right=c["bc"][1], boundary_top=c["bc"][2],
              boundary_bottom=c["bc"][3], apply_friction=c["friction"], constant_value=0.05)
    dt_ = DT[name]
    bed, rough = torch.tensor(d["bed"], dtype=dt_), torch.tensor(d["rough"], dtype=dt_)
    U = torch.tensor(d["U0"], dtype=dt_)
    lvl = torch.tensor(c["level"], dtype=dt_)
    k1, ax, ay = O.spatial_operator(U, bed, cfg, 0.1, 0.1, lvl, wave_speed=True)
    exact(k1, d[f"s{case}_{name}_rhs"])
    exact([float(ax), float(ay)], d[f"s{case}_{name}_axay"])
    dts = []
    for k in range(12):
        U, dtv = O.step(U, bed, rough, cfg, 0.1, 0.1, 1e30, lvl)
        dts.append(dtv)
        exact(U, d[f"s{case}_{name}_U"][k])
    exact(dts, d[f"s{case}_{name}_dt"])


def test_first_dam_break_dt_closed_form():
    # minmod flattens the jump, so the fastest face speed is
    # S = max(cL, (cL+cR)/2 + cL - cR) with cL = sqrt(2g), cR = sqrt(0.5g):
    # dt0 = cfl / (2 S / dx), 0.0045152362 s in FP32.
    from diffswe2d_amd import scenarios

    sc = scenarios.dam_break(n=64, radius=2.0)
    cfg = Cfg()
    U = torch.tensor(sc["U"], dtype=torch.float64)
    _, dt = O.step(U, torch.zeros(1, 1, 64, 64, dtype=torch.float64),
                   torch.zeros(1, 1, 64, 64, dtype=torch.float64), cfg, 0.1, 0.1, 1e30)
    cL, cR = math.sqrt(2.0 * 9.81), math.sqrt(0.5 * 9.81)
    S = max(cL, 0.5 * (cL + cR) + cL - cR)
    assert abs(dt - 0.5 / (2 * S / 0.1)) < 1e-14
    assert abs(dt - 0.004515236409857) < 1e-14


def test_error_paths():
    cfg = Cfg()
    bed = torch.zeros(1, 1, 6, 6)
    U = torch.zeros(1, 3, 6, 6)
    U[:, 0] = 1.0
    with pytest.raises(FloatingPointError, match="Non-positive timestep"):
        O.step(U, bed, bed, cfg, 0.1, 0.1, 0.0)
    Un = U.clone()
    Un[0, 1, 2, 2] = float("nan")  # NaN momentum reaches the wave speeds
    with pytest.raises(FloatingPointError, match="non-finite"):
        O.step(Un, bed, bed, cfg, 0.1, 0.1, 1.0)
    with pytest.raises(RuntimeError, match="max_steps=1 reached"):
        O.forward(U, bed, bed, cfg, 0.1, 0.1, t_end=10.0, max_steps=1)
    with pytest.raises(ValueError, match="cannot be negative"):
        O.forward(U, bed, bed, cfg, 0.1, 0.1, t_end=-1.0)
