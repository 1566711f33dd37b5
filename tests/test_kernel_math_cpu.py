"""The kernels' scalar math, run on the host, against the oracle and its autograd.

The same __host__ __device__ functions that the gfx950 kernels inline are
compiled for the host (tests/hostcheck) and driven with plain loops.  Forward
values must equal the oracle's (FP64 to rounding, FP32 bitwise or within an
ulp or two); the hand-written reverse must equal torch autograd through the
oracle, which reproduces the reference's autograd conventions.
"""
import sys
from pathlib import Path

import numpy as np
import pytest
import torch

sys.path.insert(0, str(Path(__file__).resolve().parent / "hostcheck"))
import hc  # noqa: E402

import diffswe2d_amd as swe  # noqa: E402
from oracle import swe_oracle as O  # noqa: E402

STEP_CASES = [
    dict(bc=("wall", "wall", "wall", "wall"), friction=True, level=0.0),
    dict(bc=("water_level", "wall", "transmissive", "water_level"), friction=True, level=0.45),
    dict(bc=("periodic", "periodic", "absorbing", "constant"), friction=False, level=0.3),
    dict(bc=("transmissive", "absorbing", "periodic", "periodic"), friction=True, level=0.2),
    dict(bc=("absorbing", "water_level", "wall", "absorbing"), friction=True, level=0.35),
]


def cfg_of(c, **kw):
    return swe.SWEConfig(boundary_left=c["bc"][0], boundary_right=c["bc"][1], boundary_top=c["bc"][2],
                         boundary_bottom=c["bc"][3], apply_friction=c["friction"], constant_value=0.05, **kw)


def case_inputs(golden, dtype):
    d = golden("steps")
    return (torch.tensor(d["U0"], dtype=dtype), torch.tensor(d["bed"], dtype=dtype),
            torch.tensor(d["rough"], dtype=dtype))


@pytest.mark.parametrize("case", range(len(STEP_CASES)))
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_forward_math_matches_oracle(golden, case, dtype):
    U, bed, rough = case_inputs(golden, dtype)
    c = STEP_CASES[case]
    cfg = cfg_of(c)
    lvl = torch.tensor(c["level"], dtype=dtype)
    rhs, axay = hc.spatial_operator(cfg, U, bed, 0.1, 0.1, lvl)
    ref, ax, ay = O.spatial_operator(U, bed, cfg, 0.1, 0.1, lvl, wave_speed=True)
    scale = float(ref.abs().max())
    tol = 0.0 if dtype == torch.float64 else 2e-6
    assert np.abs(rhs - ref.numpy()).max() <= tol * scale + (1e-13 * scale if dtype == torch.float64 else 0)
    assert abs(axay[0] - float(ax)) <= tol * float(ax) + 1e-15
    assert abs(axay[1] - float(ay)) <= tol * float(ay) + 1e-15
    Uo, dt = O.step(U, bed, rough, cfg, 0.1, 0.1, 1e30, lvl)
    out, _, _ = hc.step_fwd(cfg, U, bed, rough, 0.1, 0.1, lvl, dt)
    err = np.abs(out - Uo.numpy()).max() / float(Uo.abs().max())
    assert err <= (1e-14 if dtype == torch.float64 else 2e-6)


def autograd_spatial(cfg, U, bed, level, lam):
    Ur = U.clone().requires_grad_(True)
    lv = level.clone().requires_grad_(True)
    rhs = O.spatial_operator(Ur, bed, cfg, 0.1, 0.1, lv)
    gU, gl = torch.autograd.grad(rhs, (Ur, lv), grad_outputs=lam, allow_unused=True)
    return gU, (torch.zeros(()) if gl is None else gl)


def relerr(a, b):
    a, b = np.asarray(a, dtype=np.float64), np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("case", range(len(STEP_CASES)))
def test_spatial_operator_adjoint_matches_autograd(golden, case):
    U, bed, _ = case_inputs(golden, torch.float64)
    c = STEP_CASES[case]
    cfg = cfg_of(c)
    lvl = torch.tensor(c["level"], dtype=torch.float64)
    lam = torch.randn(U.shape, dtype=torch.float64, generator=torch.Generator().manual_seed(case))
    gU_ref, gl_ref = autograd_spatial(cfg, U, bed, lvl, lam)
    gU, gl = hc.spatial_operator_bwd(cfg, U, bed, 0.1, 0.1, lvl, lam)
    e = relerr(gU, gU_ref)
    print(f"case {case}: rel err dU {e:.2e}; dlevel {float(gl.sum()):.6e} vs {float(gl_ref):.6e}")
    assert e <= 1e-12
    assert abs(float(gl.sum()) - float(gl_ref)) <= 1e-10 * max(1.0, abs(float(gl_ref)))


def autograd_step(cfg, U, bed, rough, level, lam):
    Ur = U.clone().requires_grad_(True)
    lv = level.clone().requires_grad_(True)
    out, dt = O.step(Ur, bed, rough, cfg, 0.1, 0.1, 1e30, lv)
    gU, gl = torch.autograd.grad(out, (Ur, lv), grad_outputs=lam, allow_unused=True)
    return gU, (torch.zeros(()) if gl is None else gl), dt


@pytest.mark.parametrize("case", range(len(STEP_CASES)))
def test_step_adjoint_matches_autograd(golden, case):
    U, bed, rough = case_inputs(golden, torch.float64)
    c = STEP_CASES[case]
    cfg = cfg_of(c)
    lvl = torch.tensor(c["level"], dtype=torch.float64)
    lam = torch.randn(U.shape, dtype=torch.float64, generator=torch.Generator().manual_seed(10 + case))
    gU_ref, gl_ref, dt = autograd_step(cfg, U, bed, rough, lvl, lam)
    gU, gl = hc.step_bwd(cfg, U, bed, rough, 0.1, 0.1, lvl, dt, lam)
    e = relerr(gU, gU_ref)
    print(f"case {case}: step rel err dU {e:.2e}; dlevel {float(gl.sum()):.6e} vs {float(gl_ref):.6e}")
    assert e <= 1e-12
    assert abs(float(gl.sum()) - float(gl_ref)) <= 1e-10 * max(1.0, abs(float(gl_ref)))


def test_step_adjoint_roughness_length_friction(golden):
    U, bed, rough = case_inputs(golden, torch.float64)
    c = STEP_CASES[1]
    cfg = cfg_of(c, frictionmodel="roughness")
    z0 = 0.002 + rough
    lvl = torch.tensor(c["level"], dtype=torch.float64)
    lam = torch.randn(U.shape, dtype=torch.float64, generator=torch.Generator().manual_seed(99))
    gU_ref, gl_ref, dt = autograd_step(cfg, U, bed, z0, lvl, lam)
    gU, gl = hc.step_bwd(cfg, U, bed, z0, 0.1, 0.1, lvl, dt, lam)
    assert relerr(gU, gU_ref) <= 1e-12


def test_per_member_levels(golden):
    U, bed, rough = case_inputs(golden, torch.float64)
    cfg = cfg_of(STEP_CASES[1])
    lvl = torch.tensor([0.45, 0.3], dtype=torch.float64)
    lam = torch.randn(U.shape, dtype=torch.float64, generator=torch.Generator().manual_seed(5))
    gU_ref, gl_ref, dt = autograd_step(cfg, U, bed, rough, lvl, lam)
    gU, gl = hc.step_bwd(cfg, U, bed, rough, 0.1, 0.1, lvl, dt, lam)
    assert relerr(gU, gU_ref) <= 1e-12
    assert np.allclose(gl, gl_ref.numpy(), rtol=1e-10, atol=1e-14)


def test_division_sequence_rounds_like_ieee_division():
    """div_rn / div_by (the FP32 division of the kernels) equal IEEE division bit for bit over
    random normal-range operands, with the reciprocal estimate exact or one ulp off either way
    (the hardware estimate's bound); zero numerators compare as values."""
    import ctypes

    lib = hc.lib()
    out = (ctypes.c_longlong * 6)()
    lib.hc_div_check(ctypes.c_longlong(2_000_000), ctypes.c_uint(7), out)
    print("mismatches (div_rn exact/-1/+1 ulp estimate, div_by exact/-1/+1):", list(out))
    assert list(out) == [0] * 6


def test_square_root_pick_rounds_like_ieee_sqrt():
    """sqrt_rn_from (the FP32 square root of the kernels) equals IEEE sqrt over random arguments in
    [2^-96, 2^100], whether the estimate is the exact root or one ulp off either way."""
    import ctypes

    lib = hc.lib()
    out = (ctypes.c_longlong * 3)()
    lib.hc_sqrt_check(ctypes.c_longlong(2_000_000), ctypes.c_uint(5), out)
    print("mismatches (exact / -1 ulp / +1 ulp estimate):", list(out))
    assert list(out) == [0] * 3


def test_friction_cube_root_rounds_to_nearest():
    """pow_third_from (the FP32 h^third of the Manning friction) against h^third rounded from a double
    evaluation over depths 1e-6..1e4, with log2 / cube-root estimates 1e-6 off: never more than one ulp
    away, and one ulp away for at most 1 in 10^5 arguments."""
    import ctypes

    lib = hc.lib()
    out = (ctypes.c_longlong * 3)()
    n = 4_000_000
    lib.hc_pow_check(ctypes.c_longlong(n), ctypes.c_uint(11), out)
    print(f"pow_third: {out[0]} one ulp off, {out[1]} further, of {n} (host powf differs for {out[2]})")
    assert out[1] == 0 and out[0] <= n // 100_000
