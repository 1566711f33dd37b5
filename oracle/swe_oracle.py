"""CPU oracle for the DiffSWE2d per-time-step finite-volume update.

TEST INFRASTRUCTURE ONLY.  Results are checked against it: only ``tests/``, the
package's smoke check and the ``cpu_baseline`` leg of ``bench.py`` may import
it.  The product path (``diffswe2d_amd``) never calls it and fails
loudly when its HIP library is missing.

It is an eager PyTorch (CPU) restatement of the reference algorithm, written so
that every floating-point operation happens in the same order, on the same
operands and with the same dtype as the reference.  That is what makes it
rounding-faithful: the reference's own numbers come from these same ATen CPU
kernels.  Python-side constants fold in double before a single cast, exactly as
in the reference (``0.5 * gravity`` becomes ``4.905``).

Followed files (reference tree):
  * ``src/numerics.py:4-27``        -- minmod-theta limited increment
  * ``src/riemann.py:6-106``        -- primitive variables, two-wave HLL flux
  * ``src/reconstruction.py:12-108``-- hydrostatic "CN" MUSCL reconstruction
  * ``src/boundary.py:113-413``     -- ghost cells (ng = 2)
  * ``src/friction.py:4-131``       -- semi-implicit Manning / roughness-length friction
  * ``src/solver.py:367-779``       -- boundary level, spatial operator, CFL dt,
                                       positivity clean-up, RK2 midpoint step, forward loop

Autograd through these functions reproduces the reference's autograd tie and
branch conventions (``torch.minimum`` ties split 0.5/0.5, ``clamp`` passes on
the inclusive side, ``torch.where`` follows the selected branch, ``sign`` has a
zero derivative), because the same differentiable ops are used in the same
nesting.

Parity pins (see ``tests/test_oracle.py``): golden vectors written from the
reference's own hot-path modules (``tests/golden/``), closed forms
(first dam-break dt, friction factor, lake at rest), invariants (mass
conservation, x/y transpose symmetry) and brute force on tiny grids.
"""
from __future__ import annotations

import math
from numbers import Integral, Real
from typing import Optional

import torch
import torch.nn.functional as F

NG = 2
BC_TYPES = ("transmissive", "periodic", "constant", "water_level", "wall", "absorbing")

# torch 2.10's CPU sqrt is not correctly rounded (about 0.7 % of
# float32 and 0.9 % of float64 inputs land one ulp away from the IEEE result).
# SQRT_MODE "torch" keeps the reference's own rounding; "ieee" gives the
# correctly rounded square root that IEEE hardware (and the GPU) computes:
# float32 via one rounding of the float64 root, float64 via numpy's hardware
# sqrt, both with the sqrt derivative grad / (2 * result) for autograd.  "ieee"
# also rounds the float32 friction power h^(1/3) once from its FP64 value
# (_pow_third), as the GPU does.
SQRT_MODE = "torch"


class _IeeeSqrt(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x):
        if x.dtype == torch.float32:
            y = torch.sqrt(x.double()).float()
        else:
            import numpy as np

            y = torch.from_numpy(np.sqrt(x.detach().cpu().numpy())).to(x.device)
        ctx.save_for_backward(y)
        return y

    @staticmethod
    def backward(ctx, g):
        (y,) = ctx.saved_tensors
        return g / (2 * y)


def _sqrt(x):
    if SQRT_MODE == "ieee":
        return _IeeeSqrt.apply(x)
    return torch.sqrt(x)


# ---------------------------------------------------------------------------
# dtype-dependent floors (reference: riemann.py:6-17, reconstruction.py:12-15,
# friction.py:4-9)
# ---------------------------------------------------------------------------
def depth_floor(dtype, epsilon: float, dry_depth: float) -> float:
    """Positive depth floor used before divisions and square roots."""
    if dtype in (torch.float16, torch.bfloat16):
        prec = 1e-5
    elif dtype == torch.float32:
        prec = 1e-6
    else:
        prec = 1e-10
    return max(float(epsilon), float(dry_depth), prec)


def hll_floor(dtype, epsilon: float) -> float:
    """Floor of the HLL denominator SR - SL (riemann.py:13-17)."""
    return max(float(epsilon), 1e-6 if dtype == torch.float32 else 1e-10)


def recon_floor(dtype, epsilon: float, dry_depth: float) -> float:
    """Floor of the velocity-factor denominator (reconstruction.py:12-15)."""
    return max(float(epsilon), float(dry_depth), 1e-6 if dtype == torch.float32 else 1e-10)


# ---------------------------------------------------------------------------
# limiter (numerics.py:4-27)
# ---------------------------------------------------------------------------
def minmod3(a, b, c):
    """Three-argument minmod: the smallest magnitude when all signs agree, else 0.

    Written as sign(a) * min(|a|, min(|b|, |c|)) so that autograd sees the same
    op nesting (and therefore the same tie split) as numerics.py:4-13.
    """
    sa, sb, sc = torch.sign(a), torch.sign(b), torch.sign(c)
    agree = (sa == sb) & (sb == sc)
    smallest = torch.minimum(a.abs(), torch.minimum(b.abs(), c.abs()))
    return torch.where(agree, sa * smallest, torch.zeros_like(a))


def limited_increment(q, dim: int, theta: float = 1.3):
    """Limited cell increment Delta q (numerics.py:16-27).

    Neighbours are taken with a periodic shift over the padded array, like the
    reference; the wrapped values only reach faces that are cropped later.
    """
    left = torch.roll(q, 1, dim)
    right = torch.roll(q, -1, dim)
    return minmod3(theta * (q - left), 0.5 * (right - left), theta * (right - q))


# ---------------------------------------------------------------------------
# primitive variables and HLL flux (riemann.py:20-106)
# ---------------------------------------------------------------------------
def primitive(U, gravity: float, epsilon: float, dry_depth: float):
    """(h, u, v, c) from [h, hu, hv]; dry cells get zero velocity and celerity."""
    h = torch.clamp(U[:, 0:1], min=0.0)
    wet = h > dry_depth
    safe = torch.clamp(h, min=depth_floor(U.dtype, epsilon, dry_depth))
    zero = torch.zeros_like(h)
    u = torch.where(wet, U[:, 1:2] / safe, zero)
    v = torch.where(wet, U[:, 2:3] / safe, zero)
    c = torch.where(wet, _sqrt(float(gravity) * safe), zero)
    return h, u, v, c


def hll_flux(UL, UR, normal: str, gravity: float, epsilon: float, dry_depth: float):
    """Two-wave HLL flux with dry-front speeds (riemann.py:41-106).

    Returns (flux [B,3,...] in (h, x-momentum, y-momentum) order, face speed).
    """
    hL, uL, vL, cL = primitive(UL, gravity, epsilon, dry_depth)
    hR, uR, vR, cR = primitive(UR, gravity, epsilon, dry_depth)
    if normal == "x":
        nL, tL, nR, tR = uL, vL, uR, vR
    else:
        nL, tL, nR, tR = vL, uL, vR, uR

    um = 0.5 * (nL + nR) + cL - cR
    cm = 0.5 * (cL + cR) + 0.25 * (nL - nR)
    dryL = hL <= dry_depth
    dryR = hR <= dry_depth
    SL = torch.where(dryL, nR - 2.0 * cR, torch.minimum(nL - cL, um - cm))
    SR = torch.where(dryR, nL + 2.0 * cL, torch.maximum(nR + cR, um + cm))

    half_g = 0.5 * float(gravity)
    mL, mR = hL * nL, hR * nR
    pL = hL * (nL.square() + half_g * hL)
    pR = hR * (nR.square() + half_g * hR)

    den = torch.clamp(SR - SL, min=hll_floor(UL.dtype, epsilon))
    m_star = (SR * mL - SL * mR + SL * SR * (hR - hL)) / den
    p_star = (SR * pL - SL * pR + SL * SR * (hR * nR - hL * nL)) / den

    upwind_left = SL >= 0.0
    upwind_right = SR <= 0.0
    fh = torch.where(upwind_left, mL, torch.where(upwind_right, mR, m_star))
    fn = torch.where(upwind_left, pL, torch.where(upwind_right, pR, p_star))
    ft = torch.where(fh > 0.0, tL, tR) * fh

    flux = torch.cat((fh, fn, ft) if normal == "x" else (fh, ft, fn), dim=1)
    both = dryL & dryR
    flux = torch.where(both.expand_as(flux), torch.zeros_like(flux), flux)
    speed = torch.maximum(SL.abs(), SR.abs())
    speed = torch.where(both, torch.zeros_like(speed), speed)
    return flux, speed


# ---------------------------------------------------------------------------
# hydrostatic reconstruction (reconstruction.py:18-108)
# ---------------------------------------------------------------------------
def reconstruct(Up, zp, dim: int, gravity, epsilon, dry_depth, theta=1.3, water_slope_reset=True):
    """Face states on the face between each padded cell and its +dim neighbour.

    Returns (UL, UR, sL, sR): left/right conservative face states and the
    topographic source terms for the cell on each side of the face.
    """
    h, u, v, _ = primitive(Up, gravity, epsilon, dry_depth)
    eta = h + zp
    dh = limited_increment(h, dim, theta)
    deta = limited_increment(eta, dim, theta)
    dz = limited_increment(zp, dim, theta)
    du = limited_increment(u, dim, theta)
    dv = limited_increment(v, dim, theta)

    if water_slope_reset:
        # The bed implied at each side of the cell by the limited free surface.
        implied_minus = eta - h - 0.5 * (deta - dh)
        implied_plus = eta - h + 0.5 * (deta - dh)
        crosses = (implied_minus > torch.roll(eta, 1, dim)) | (implied_plus > torch.roll(eta, -1, dim))
        deta = torch.where((deta.square() > dz.square()) & crosses, dh + dz, deta)

    def nxt(q):
        return torch.roll(q, -1, dim)

    hR_cell, zR_cell, dhR = nxt(h), nxt(zp), nxt(dh)
    h_lin_L = h + 0.5 * dh
    h_lin_R = hR_cell - 0.5 * dhR
    eta_L = eta + 0.5 * deta
    eta_R = nxt(eta) - 0.5 * nxt(deta)
    z_L = zp + 0.5 * dz
    z_R = zR_cell - 0.5 * nxt(dz)

    z_face = torch.minimum(torch.maximum(z_L, z_R), torch.minimum(eta_L, eta_R))
    hcL = torch.clamp(torch.minimum(eta_L - z_face, h_lin_L), min=0.0)
    hcR = torch.clamp(torch.minimum(eta_R - z_face, h_lin_R), min=0.0)

    floor = recon_floor(Up.dtype, epsilon, dry_depth)
    kL = 1.0 - 0.5 * dh / torch.clamp(h, min=floor)
    kR = 1.0 + 0.5 * dhR / torch.clamp(hR_cell, min=floor)
    kL = torch.clamp(torch.where(h > dry_depth, kL, torch.ones_like(h)), -10.0, 10.0)
    kR = torch.clamp(torch.where(hR_cell > dry_depth, kR, torch.ones_like(hR_cell)), -10.0, 10.0)

    uL = u + 0.5 * kL * du
    vL = v + 0.5 * kL * dv
    uR = nxt(u) - 0.5 * kR * nxt(du)
    vR = nxt(v) - 0.5 * kR * nxt(dv)

    UL = torch.cat((hcL, hcL * uL, hcL * vL), dim=1)
    UR = torch.cat((hcR, hcR * uR, hcR * vR), dim=1)
    sL = 0.5 * gravity * (h + hcL) * (zp - z_face)
    sR = 0.5 * gravity * (hcR + hR_cell) * (zR_cell - z_face)
    return UL, UR, sL, sR


# ---------------------------------------------------------------------------
# ghost cells (boundary.py:113-413)
# ---------------------------------------------------------------------------
def as_boundary_value(value, ref):
    """Scalar / 0-d / one-element / per-member value, broadcastable to [B,1,*,*]."""
    t = value.to(dtype=ref.dtype, device=ref.device) if torch.is_tensor(value) else ref.new_tensor(value)
    if t.ndim == 0:
        return t
    if t.numel() == 1:
        return t.reshape(())
    if t.ndim == 1 and t.shape[0] == ref.shape[0]:
        return t.reshape(-1, 1, 1, 1)
    raise ValueError(
        "Boundary values must be a scalar, a one-element tensor, or a tensor with one "
        f"value per batch member. Received shape {tuple(t.shape)}."
    )


def ghost_pad(U, bed, left="transmissive", right="transmissive", top="transmissive",
              bottom="transmissive", constant_value=0.0, constant_level=0.0, ng=NG):
    """Pad state and bed with ng ghost layers; sides applied in the order L, R, T, B.

    "left/right" are columns (x, last dim); "top/bottom" are rows 0.. and -ng..
    The order matters only for corner ghosts, which never reach interior cells.
    """
    sides = {"left": left, "right": right, "top": top, "bottom": bottom}
    for side, kind in sides.items():
        if kind not in BC_TYPES:
            raise ValueError(
                f"Unsupported {side} boundary type {kind!r}. Valid types are {sorted(BC_TYPES)}."
            )
    Up = F.pad(U, (ng, ng, ng, ng), mode="replicate")
    zp = F.pad(bed, (ng, ng, ng, ng), mode="replicate")
    # periodic bed copies, interior span of each side only
    if left == "periodic":
        zp[:, :, ng:-ng, :ng] = bed[:, :, :, -ng:]
    if right == "periodic":
        zp[:, :, ng:-ng, -ng:] = bed[:, :, :, :ng]
    if top == "periodic":
        zp[:, :, :ng, ng:-ng] = bed[:, :, -ng:, :]
    if bottom == "periodic":
        zp[:, :, -ng:, ng:-ng] = bed[:, :, :ng, :]

    forced = None
    if any(k in ("water_level", "absorbing") for k in sides.values()):
        forced = torch.clamp(as_boundary_value(constant_level, U) - zp, min=0.0)
    cval = None
    if any(k == "constant" for k in sides.values()):
        cval = as_boundary_value(constant_value, U)

    # (ghost slice, mirrored-interior slice, momentum channel, outward sign, flip dim)
    geometry = {
        "left": ((slice(None), slice(0, ng)), (slice(None), slice(ng, 2 * ng)), 1, -1.0, -1),
        "right": ((slice(None), slice(-ng, None)), (slice(None), slice(-2 * ng, -ng)), 1, 1.0, -1),
        "top": ((slice(0, ng), slice(None)), (slice(ng, 2 * ng), slice(None)), 2, -1.0, -2),
        "bottom": ((slice(-ng, None), slice(None)), (slice(-2 * ng, -ng), slice(None)), 2, 1.0, -2),
    }
    periodic_src = {
        "left": ((slice(ng, -ng), slice(0, ng)), (slice(None), slice(-ng, None))),
        "right": ((slice(ng, -ng), slice(-ng, None)), (slice(None), slice(0, ng))),
        "top": ((slice(0, ng), slice(ng, -ng)), (slice(-ng, None), slice(None))),
        "bottom": ((slice(-ng, None), slice(ng, -ng)), (slice(0, ng), slice(None))),
    }
    for side in ("left", "right", "top", "bottom"):
        kind = sides[side]
        ghost, inner, mch, outward, fdim = geometry[side]
        if kind == "periodic":
            dst, src = periodic_src[side]
            Up[(slice(None), slice(None)) + dst] = U[(slice(None), slice(None)) + src]
        elif kind == "constant":
            Up[(slice(None), slice(None)) + ghost] = cval
        elif kind == "water_level":
            Up[(slice(None), slice(0, 1)) + ghost] = forced[(slice(None), slice(None)) + ghost]
        elif kind == "wall":
            Up[(slice(None), slice(mch, mch + 1)) + ghost] = -Up[
                (slice(None), slice(mch, mch + 1)) + inner].flip(dims=(fdim,))
        elif kind == "absorbing":
            h_in = Up[(slice(None), slice(0, 1)) + ghost]
            h_ref = forced[(slice(None), slice(None)) + ghost]
            h_c = torch.where(h_ref > 1e-4, h_ref, h_in)
            c = _sqrt(9.81 * torch.clamp(h_c, min=1e-6))
            flow = c * (h_in - h_ref) if outward > 0 else -c * (h_in - h_ref)
            Up[(slice(None), slice(mch, mch + 1)) + ghost] = flow
    return Up, zp


# ---------------------------------------------------------------------------
# friction (friction.py:12-131)
# ---------------------------------------------------------------------------
def _speed(h, hu, hv, dry_depth, epsilon):
    wet = h > dry_depth
    safe = torch.clamp(h, min=depth_floor(h.dtype, epsilon, dry_depth))
    zero = torch.zeros_like(h)
    u = torch.where(wet, hu, zero) / safe
    v = torch.where(wet, hv, zero) / safe
    s2 = u.square() + v.square()
    moving = wet & (s2 > 0.0)
    s2 = torch.clamp(s2, min=depth_floor(h.dtype, epsilon, 0.0) ** 2)
    return wet, safe, torch.where(moving, _sqrt(s2), zero)


def _pow_third(x):
    """x ** (1/3) as the reference computes it (torch pow), or, with SQRT_MODE "ieee",
    rounded once from the FP64 power for float32 x (the correctly rounded value the
    GPU's friction power gives; torch's CPU float32 pow is one ulp off for some
    arguments).  The exponent is the Python double 1/3, cast to x's dtype by pow."""
    if SQRT_MODE == "ieee" and x.dtype == torch.float32:
        third = float(torch.tensor(1.0 / 3.0, dtype=torch.float32))
        return x.double().pow(third).float()
    return x.pow(1.0 / 3.0)


def manning_coefficient(h, n, epsilon=1e-12):
    """Cf = 9.81 n^2 / h^(1/3) (friction.py:62-66; gravity hard-coded as 9.81)."""
    safe = torch.clamp(h, min=depth_floor(h.dtype, epsilon, 0.0))
    return 9.81 * torch.clamp(n, min=0.0).square() / _pow_third(safe)


def roughness_length_coefficient(h, z0, epsilon=1e-12):
    """BG_Flood smart-friction coefficient from h / z0 (friction.py:41-59)."""
    floor = depth_floor(h.dtype, epsilon, 0.0)
    ratio = torch.clamp(h, min=floor) / torch.clamp(z0, min=floor)
    shallow = 1.0 / torch.clamp(0.46 * ratio, min=floor)
    r = torch.clamp(ratio, min=floor)
    log_term = 2.5 * (torch.log(r) - 1.0 + 1.359 / r)
    deep = 1.0 / torch.clamp(log_term.square(), min=floor)
    return torch.where(ratio < math.e, shallow, deep)


def apply_friction(U, coeff_field, dt, model="manning", dry_depth=1e-6, epsilon=1e-12):
    """Semi-implicit momentum damping 1 / (1 + Cf |u| dt / h) (friction.py:69-131)."""
    h = torch.clamp(U[:, 0:1], min=0.0)
    hu, hv = U[:, 1:2], U[:, 2:3]
    wet, safe, speed = _speed(h, hu, hv, dry_depth, epsilon)
    if model == "manning":
        cf = manning_coefficient(h, coeff_field, epsilon)
    else:
        cf = roughness_length_coefficient(h, coeff_field, epsilon)
    factor = 1.0 / (1.0 + cf * speed * dt / safe)
    zero = torch.zeros_like(h)
    return torch.cat((h, torch.where(wet, hu * factor, zero), torch.where(wet, hv * factor, zero)), dim=1)


# ---------------------------------------------------------------------------
# solver (solver.py:367-779)
# ---------------------------------------------------------------------------
def _forcing(value, h):
    if value is None:
        return torch.zeros_like(h)
    t = value.to(dtype=h.dtype, device=h.device) if torch.is_tensor(value) else h.new_tensor(value)
    return torch.broadcast_to(t, h.shape)


def as_level(level, ref):
    """Boundary level on the model dtype/device, autograd-preserving (solver.py:378-389)."""
    if torch.is_tensor(level):
        return level.to(dtype=ref.dtype, device=ref.device)
    return ref.new_tensor(level)


def spatial_operator(U, bed, cfg, dx, dy, level=None, rain=None, infiltration=None, wave_speed=False):
    """Conservative tendency dU/dt (solver.py:448-531).

    bed is [1|B,1,ny,nx]; the wave speeds, when asked for, are maxima over all
    batch members of the interior faces.
    """
    if U.ndim != 4 or U.shape[1] != 3:
        raise ValueError(f"U must have shape [batch,3,ny,nx], got {tuple(U.shape)}.")
    B, _, ny, nx = U.shape
    if bed.shape[0] != B:
        bed = bed.expand(B, -1, -1, -1)
    lvl = as_level(cfg.constant_level if level is None else level, U)
    Up, zp = ghost_pad(U, bed, cfg.boundary_left, cfg.boundary_right, cfg.boundary_top,
                       cfg.boundary_bottom, cfg.constant_value, lvl)
    args = (cfg.gravity, cfg.epsilon, cfg.dry_depth, cfg.limiter_theta, cfg.water_slope_reset)
    ULx, URx, sxL, sxR = reconstruct(Up, zp, -1, *args)
    ULy, URy, syL, syR = reconstruct(Up, zp, -2, *args)
    Fx, ax = hll_flux(ULx, URx, "x", cfg.gravity, cfg.epsilon, cfg.dry_depth)
    Gy, ay = hll_flux(ULy, URy, "y", cfg.gravity, cfg.epsilon, cfg.dry_depth)

    # the cell left of a face sees F - sL, the cell right of it F - sR (normal channel only)
    FxL, FxR, GyL, GyR = Fx.clone(), Fx.clone(), Gy.clone(), Gy.clone()
    FxL[:, 1:2] -= sxL
    FxR[:, 1:2] -= sxR
    GyL[:, 2:3] -= syL
    GyR[:, 2:3] -= syR
    rhs = (FxR[:, :, 2:2 + ny, 1:1 + nx] - FxL[:, :, 2:2 + ny, 2:2 + nx]) / dx
    rhs += (GyR[:, :, 1:1 + ny, 2:2 + nx] - GyL[:, :, 2:2 + ny, 2:2 + nx]) / dy

    h = U[:, 0:1]
    rate = _forcing(rain, h)
    loss = torch.clamp(_forcing(infiltration, h), min=0.0)
    rhs = torch.cat((rhs[:, 0:1] + rate - loss, rhs[:, 1:2], rhs[:, 2:3]), dim=1)
    if wave_speed:
        return rhs, ax[:, :, 2:2 + ny, 1:2 + nx].amax(), ay[:, :, 1:2 + ny, 2:2 + nx].amax()
    return rhs


def cfl_dt(ax, ay, cfg, dx, dy):
    """CFL step cfl / max(ax/dx + ay/dy, eps) (solver.py:537-545, 620-623)."""
    return cfg.cfl / torch.clamp(ax / dx + ay / dy, min=cfg.epsilon)


def clean_and_limit(old, cand, cfg):
    """Positivity limiter and dry-cell momentum removal (solver.py:547-570)."""
    if cfg.positivity_limiter:
        delta = cand - old
        dh = delta[:, 0:1]
        shrink = torch.where(dh < 0.0,
                             torch.clamp(old[:, 0:1] / torch.clamp(-dh, min=cfg.epsilon), max=1.0),
                             torch.ones_like(dh))
        cand = old + shrink * delta
    h = torch.clamp(cand[:, 0:1], min=0.0)
    wet = h > cfg.dry_depth
    zero = torch.zeros_like(h)
    return torch.cat((h, torch.where(wet, cand[:, 1:2], zero), torch.where(wet, cand[:, 2:3], zero)), dim=1)


def step(U, bed, roughness, cfg, dx, dy, max_dt, level=None, rain_start=None, rain_half=None,
         infiltration=None, fixed_dt=None):
    """One adaptive midpoint step (solver.py:572-663); returns (U_new, float dt).

    ``fixed_dt`` (a test device, not in the reference) bypasses the CFL choice
    and uses that dt, so finite differences can be taken with the dt series
    frozen, matching the reference's detached dt.
    """
    max_dt = max_dt.to(dtype=U.dtype) if torch.is_tensor(max_dt) else U.new_tensor(max_dt)
    lvl = as_level(cfg.constant_level if level is None else level, U)
    k1, ax, ay = spatial_operator(U, bed, cfg, dx, dy, lvl, rain_start, infiltration, wave_speed=True)
    if fixed_dt is not None:
        dt = U.new_tensor(fixed_dt)
    else:
        dt = torch.minimum(cfl_dt(ax, ay, cfg, dx, dy), max_dt).detach()
    if not torch.isfinite(dt):
        raise FloatingPointError("The calculated timestep is non-finite.")
    dt_value = float(dt)
    if dt_value <= 0.0:
        raise FloatingPointError(f"Non-positive timestep: {dt_value}.")
    pred = clean_and_limit(U, U + 0.5 * dt * k1, cfg)
    k2 = spatial_operator(pred, bed, cfg, dx, dy, lvl, rain_half, infiltration)
    new = clean_and_limit(U, U + dt * k2, cfg)
    if cfg.apply_friction:
        rough = roughness if roughness.shape[0] == U.shape[0] else roughness.expand(U.shape[0], -1, -1, -1)
        new = apply_friction(new, rough, dt, cfg.frictionmodel, cfg.dry_depth, cfg.epsilon)
    return new, dt_value


def interpolate_level(times, levels, t, ref):
    """Dynamic boundary level at time t (solver.py:391-415); extrapolates at the ends."""
    times = times.to(dtype=ref.dtype, device=ref.device)
    levels = levels.to(dtype=ref.dtype, device=ref.device)
    tt = ref.new_tensor(t)
    i = torch.clamp(torch.searchsorted(times, tt, right=False), 1, times.numel() - 1)
    t0, t1 = times[i - 1], times[i]
    w = (tt - t0) / torch.clamp(t1 - t0, min=torch.finfo(times.dtype).eps)
    return levels[i - 1] + w * (levels[i] - levels[i - 1])


def scalar_float(value, name):
    if torch.is_tensor(value):
        if value.numel() != 1:
            raise ValueError(f"{name} must be scalar.")
        return float(value.detach().cpu())
    if isinstance(value, Real):
        return float(value)
    raise TypeError(f"{name} must be a real scalar or scalar tensor.")


def forward(U0, bed, roughness, cfg, dx, dy, t_end, start_time=0.0, max_steps=1_000_000,
            boundary_level=None, max_dt_cap=None, dynamic=None, infiltration=None,
            dt_log: Optional[list] = None, gauges=None, gauge_log: Optional[list] = None):
    """Duration-driven loop (solver.py:680-779); returns (U, hmax [B,ny,nx]).

    ``dynamic`` is an optional (times, levels) pair used when boundary_level is
    None.  ``dt_log`` receives every step's dt; ``gauge_log`` receives the
    depth at the ``gauges`` cells (batch member 0) after every step.
    """
    duration = scalar_float(t_end, "t_end")
    t0 = scalar_float(start_time, "start_time")
    if not isinstance(max_steps, Integral) or isinstance(max_steps, bool) or max_steps <= 0:
        raise ValueError(f"max_steps must be positive, got {max_steps!r}.")
    if duration < 0.0:
        raise ValueError(f"t_end duration cannot be negative: {duration}.")
    if duration == 0.0:
        return U0, U0[:, 0].clone()
    U, hmax, elapsed = U0, U0[:, 0].clone(), 0.0
    tol = max(1e-12, 1e-12 * abs(duration))
    for _ in range(max_steps):
        remaining = duration - elapsed
        if remaining <= tol:
            return U, hmax
        cap = remaining if max_dt_cap is None else min(remaining, scalar_float(max_dt_cap, "max_dt_cap"))
        now = t0 + elapsed
        if boundary_level is not None:
            lvl = as_level(boundary_level, U)
        elif dynamic is not None:
            lvl = interpolate_level(dynamic[0], dynamic[1], now, U)
        else:
            lvl = as_level(cfg.constant_level, U)
        U, dt_value = step(U, bed, roughness, cfg, dx, dy, U.new_tensor(cap), lvl,
                           infiltration=infiltration)
        hmax = torch.maximum(hmax, U[:, 0])
        elapsed += dt_value
        if elapsed > duration and elapsed - duration <= tol:
            elapsed = duration
        if dt_log is not None:
            dt_log.append(dt_value)
        if gauge_log is not None and gauges:
            gauge_log.append([float(U[0, 0, iy, ix]) for (iy, ix) in gauges])
    raise RuntimeError(
        f"max_steps={max_steps} reached after {elapsed:.8f} s; requested duration={duration:.8f} s."
    )
