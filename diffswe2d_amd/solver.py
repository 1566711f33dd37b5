"""Drop-in ``SWE2D`` whose time step runs on the gfx950 HIP library.

Same constructor, attributes and methods as the reference ``SWE2D``
(src/solver.py:34-801); the numerics of ``spatial_operator``, ``step`` and
``forward`` run in ``libswe2d_hip.so``.  There is no CPU path: tensors must be
on a ROCm device, and a missing library raises at the first call.

Differences a caller can see:
  * ``forward`` integrates on the device without a host synchronisation per
    step (the reference syncs 2-3 times per step), a dynamic boundary series
    included; a callback takes host-driven steps with identical numerics.
  * gradients flow through ``step``/``forward`` via the hand-written adjoint
    (``diffswe2d_amd.adjoint``) to U, the boundary level (or dynamic series) and
    the bed / trainable DEM correction.  Roughness and the source terms of
    ``step``/``forward`` are not differentiated (a tensor infiltration that
    requires grad raises); the hmax of a device-driven differentiable forward
    raises on backward unless ``differentiable_hmax`` is set.
"""
from __future__ import annotations

import ctypes
import math
from numbers import Integral, Real
from typing import Union

import torch
import torch.nn as nn

from . import _lib
from .config import SWEConfig

ScalarOrTensor = Union[float, int, torch.Tensor]

_DTYPES = {torch.float32: "f32", torch.float64: "f64"}


def sync_parity(ws, ctl):
    """Point the host-side step parity at the wave-speed bank a device-driven run left cleared.

    Step k accumulates its wave speeds into bank k & 1 and clears the other
    bank for step k + 1.  After a run that ended with ``ctl.steps`` steps the
    cleared bank is ``steps & 1``; a run stopped by a bad dt at step ``steps``
    had already cleared the bank of step ``steps + 1``."""
    ws["parity"] = (ctl.steps & 1) if ctl.status == 0 else 1 - (ctl.steps & 1)


class SWE2D(nn.Module):
    """B-Flood-style 2-D shallow-water solver; state channels are [h, hu, hv].

    ``forward(..., t_end=...)`` interprets ``t_end`` as the duration of the
    call, not an absolute end time (src/solver.py:37-46).
    """

    def __init__(self, bed_reference, roughness_length, dx, dy, config=None, train_dem=True,
                 maximum_dem_correction=None, valid_mask=None, x=None, y=None, rainfall=None):
        super().__init__()
        self.cfg = config or SWEConfig()
        self.dx, self.dy = float(dx), float(dy)
        if self.dx <= 0.0 or self.dy <= 0.0:
            raise ValueError(f"dx and dy must be positive, got {self.dx}, {self.dy}.")
        if bed_reference.ndim == 2:
            bed_reference = bed_reference[None, None]
        if roughness_length.ndim == 2:
            roughness_length = roughness_length[None, None]
        if bed_reference.ndim != 4 or bed_reference.shape[1] != 1:
            raise ValueError(
                f"bed_reference must have shape [ny,nx] or [B,1,ny,nx], got {tuple(bed_reference.shape)}.")
        if roughness_length.ndim != 4:
            raise ValueError(f"roughness_length must be 2-D or 4-D, got {tuple(roughness_length.shape)}.")
        if bed_reference.shape != roughness_length.shape:
            raise ValueError(
                f"DEM and roughness shapes differ: {bed_reference.shape} != {roughness_length.shape}.")
        if not torch.isfinite(bed_reference).all():
            raise ValueError("bed_reference contains non-finite values.")
        if not torch.isfinite(roughness_length).all():
            raise ValueError("roughness_length contains non-finite values.")
        self.register_buffer("bed_reference", bed_reference.clone())
        self.register_buffer("roughness_length", roughness_length.clone())
        mask = (torch.ones_like(bed_reference, dtype=torch.bool) if valid_mask is None
                else valid_mask.reshape_as(bed_reference).bool())
        self.register_buffer("valid_mask", mask)
        if x is not None:
            self.register_buffer("x", x.clone())
        if y is not None:
            self.register_buffer("y", y.clone())
        self.maximum_dem_correction = maximum_dem_correction
        self.dem_correction_parameter = nn.Parameter(torch.zeros_like(bed_reference),
                                                     requires_grad=bool(train_dem))
        self.train_dem = bool(train_dem)
        self.rainfall = rainfall
        self.use_dynamic_bc = False
        self.bc_times = None
        self.bc_wls = None
        self.track_gauges = False
        self.gauge_indices = {}
        self.bed_elevations = {}
        self.ts_time, self.ts_dt = [], []
        self.ts_h, self.ts_z = {}, {}
        self._ws = {}
        self._field_cache = {}
        self.rows_hint = 0  # wavefront rows per strip; 0 = library heuristic / autotuned value
        # "reference": the reference's rounding (default).  "fast": FP32 forward runs
        # without gradients use the library's tolerance-bounded arithmetic
        # (swe_config.arithmetic); differentiable paths always use the reference's.
        self.arithmetic = "reference"
        # With gradients, forward() runs device-driven and returns an hmax that carries
        # no gradient (a backward through it raises); True takes the per-step loop,
        # whose hmax = maximum(hmax, h) is differentiable as in src/solver.py:765.
        self.differentiable_hmax = False
        self.adjoint_checkpoint_every = 64  # K: forward checkpoints for the adjoint replay

    # ------------------------------------------------------------------ properties
    @property
    def bed(self):
        """Current bed: bed_reference + correction on valid cells (src/solver.py:134-147)."""
        correction = self.dem_correction_parameter
        if self.maximum_dem_correction is not None:
            correction = float(self.maximum_dem_correction) * torch.tanh(correction)
        return self.bed_reference + correction * self.valid_mask.to(self.bed_reference.dtype)

    @classmethod
    def from_netcdf(cls, dem_path, model_grid, rainfall_path=None, dem_variable=None, roughness_path=None,
                    roughness_variable=None, rainfall_variable=None, dem_x="x", dem_y="y", rain_x="x", rain_y="y",
                    rain_time="time", rainfall_units=None, dem_method="linear", default_roughness=None,
                    roughness_method="linear", dem_outside_domain="error", config=None, dtype=torch.float64,
                    device=None, train_dem=True, maximum_dem_correction=None):
        """Model on ``model_grid`` from a NetCDF-4 DEM, optional roughness and gridded rainfall
        (src/solver.py:149-238)."""
        from .grid import ModelGrid
        from .io_netcdf import load_dem_and_roughness_to_grid, load_rainfall_to_grid

        cfg = config or SWEConfig()
        if isinstance(model_grid, dict):
            model_grid = ModelGrid.from_bounds(dtype=dtype, device=device, **model_grid)
        if not isinstance(model_grid, ModelGrid):
            raise TypeError("model_grid must be a ModelGrid or bounds dictionary.")
        grid = load_dem_and_roughness_to_grid(dem_path, model_grid, dem_variable, roughness_path=roughness_path,
                                              roughness_variable=roughness_variable, x_name=dem_x, y_name=dem_y,
                                              dem_method=dem_method, roughness_method=roughness_method,
                                              outside_domain=dem_outside_domain, dtype=dtype, device=device)
        rainfall = None
        if rainfall_path is not None:
            rainfall = load_rainfall_to_grid(rainfall_path, model_grid, variable=rainfall_variable,
                                             time_name=rain_time, x_name=rain_x, y_name=rain_y,
                                             units=rainfall_units, outside_domain=cfg.rainfall_outside_domain,
                                             expected_crs=grid.crs, dtype=dtype, device=device)
        if roughness_variable is not None:
            roughness = grid.roughness_length
        else:
            roughness = torch.full_like(grid.bed, float(default_roughness) if default_roughness is not None
                                        else 0.025)
        model = cls(grid.bed, roughness, grid.dx, grid.dy, config=cfg, train_dem=train_dem,
                    maximum_dem_correction=maximum_dem_correction, valid_mask=grid.valid_mask, x=grid.x,
                    y=grid.y, rainfall=rainfall)
        model.model_grid = model_grid
        model.dem_path = str(dem_path)
        model.roughness_path = None if roughness_path is None else str(roughness_path)
        return model

    @classmethod
    def from_ascii(cls, dem_path, *, model_grid, roughness_path=None, config=None, dem_method="linear",
                   roughness_method="nearest", dem_outside_domain="error", roughness_outside_domain="error",
                   fill_dem_nodata=True, fill_roughness_nodata=True, default_roughness=None, train_dem=False,
                   maximum_dem_correction=None, device=None, dtype=torch.float32, **model_kwargs):
        """Model on ``model_grid`` from ESRI-ASCII DEM / roughness rasters (src/solver.py:240-323)."""
        from .io_ascii import ascii_to_tensor, parse_model_grid

        spec = parse_model_grid(model_grid)
        bed, grid = ascii_to_tensor(dem_path, spec, method=dem_method, outside_domain=dem_outside_domain,
                                    fill_internal_nodata=fill_dem_nodata, device=device, dtype=dtype)
        if not torch.isfinite(bed).all():
            raise ValueError("The regridded bed contains non-finite values.")
        if roughness_path is not None:
            rough, _ = ascii_to_tensor(roughness_path, spec, method=roughness_method,
                                       outside_domain=roughness_outside_domain,
                                       fill_internal_nodata=fill_roughness_nodata, device=device, dtype=dtype)
        else:
            rough = torch.full_like(bed, float(default_roughness) if default_roughness is not None else 0.03)
        if not torch.isfinite(rough).all():
            raise ValueError("The regridded roughness contains non-finite values.")
        model = cls(bed, rough, grid.resolution, grid.resolution, config=config, train_dem=train_dem,
                    maximum_dem_correction=maximum_dem_correction, **model_kwargs)
        model.model_grid, model.grid_spec = model_grid, grid
        model.dem_path = str(dem_path)
        model.roughness_path = None if roughness_path is None else str(roughness_path)
        model.dem_interpolation_method = dem_method
        model.roughness_interpolation_method = roughness_method
        model.dem_outside_domain = dem_outside_domain
        model.roughness_outside_domain = roughness_outside_domain
        return model

    # ------------------------------------------------------------------ configuration helpers
    def set_dynamic_boundary(self, times: torch.Tensor, water_levels: torch.Tensor, enabled: bool = True) -> None:
        """Time-varying boundary level (src/solver.py:329-354)."""
        if not torch.is_tensor(times) or not torch.is_tensor(water_levels):
            raise TypeError("times and water_levels must be tensors.")
        if times.ndim != 1 or water_levels.ndim != 1:
            raise ValueError("times and water_levels must be one-dimensional.")
        if times.numel() == 0 or times.numel() != water_levels.numel():
            raise ValueError("Boundary arrays must be non-empty and equal length.")
        if times.numel() > 1 and not torch.all(times[1:] > times[:-1]):
            raise ValueError("Dynamic boundary times must be strictly increasing.")
        if not torch.isfinite(times).all() or not torch.isfinite(water_levels).all():
            raise ValueError("Dynamic boundary arrays contain non-finite values.")
        self.bc_times, self.bc_wls = times, water_levels
        self.use_dynamic_bc = bool(enabled)

    def clear_dynamic_boundary(self) -> None:
        self.use_dynamic_bc = False
        self.bc_times = self.bc_wls = None

    def reset_gauge_tracking(self) -> None:
        self.ts_time, self.ts_dt = [], []
        self.ts_h = {name: [] for name in self.gauge_indices}
        self.ts_z = {name: [] for name in self.gauge_indices}

    @staticmethod
    def _expand(field: torch.Tensor, batch: int) -> torch.Tensor:
        if field.shape[0] == batch:
            return field
        if field.shape[0] != 1:
            raise ValueError(f"Field batch must be 1 or {batch}, got {field.shape[0]}.")
        return field.expand(batch, -1, -1, -1)

    @staticmethod
    def _prepare_boundary_level(level: ScalarOrTensor, reference: torch.Tensor) -> torch.Tensor:
        if torch.is_tensor(level):
            return level.to(dtype=reference.dtype, device=reference.device)
        return reference.new_tensor(level)

    def _interpolate_dynamic_boundary(self, current_time: float, reference: torch.Tensor) -> torch.Tensor:
        """Linear interpolation of the dynamic series at a step start (src/solver.py:391-415)."""
        if self.bc_times is None or self.bc_wls is None:
            raise RuntimeError("Dynamic boundary forcing is not configured.")
        times = self.bc_times.to(dtype=reference.dtype, device=reference.device)
        levels = self.bc_wls.to(dtype=reference.dtype, device=reference.device)
        time = reference.new_tensor(current_time)
        index = torch.clamp(torch.searchsorted(times, time, right=False), 1, times.numel() - 1)
        t0, t1 = times[index - 1], times[index]
        w0, w1 = levels[index - 1], levels[index]
        weight = (time - t0) / torch.clamp(t1 - t0, min=torch.finfo(times.dtype).eps)
        return w0 + weight * (w1 - w0)

    def _resolve_boundary_level(self, current_time, U, explicit):
        if explicit is not None:
            return self._prepare_boundary_level(explicit, U)
        if self.use_dynamic_bc:
            return self._interpolate_dynamic_boundary(current_time, U)
        return self._prepare_boundary_level(self.cfg.constant_level, U)

    # ------------------------------------------------------------------ device plumbing
    def _check_state(self, U, name="U"):
        if U.ndim != 4 or U.shape[1] != 3:
            raise ValueError(f"{name} must have shape [batch,3,ny,nx], got {tuple(U.shape)}.")
        if U.dtype not in _DTYPES:
            raise TypeError(f"{name} must be float32 or float64, got {U.dtype}.")
        if U.device.type != "cuda":
            raise RuntimeError(
                f"{name} is on {U.device}; the HIP time step runs on a ROCm device only (no CPU path).")
        if tuple(U.shape[-2:]) != tuple(self.bed_reference.shape[-2:]):
            raise ValueError(
                "U and bed must have identical spatial dimensions. "
                f"U has {tuple(U.shape[-2:])}; bed has {tuple(self.bed_reference.shape[-2:])}.")

    def _bed_trainable(self):
        """The DEM correction takes gradients (train_dem=True): the differentiable paths then carry the bed."""
        return bool(self.dem_correction_parameter.requires_grad)

    def _fields(self, U):
        """Contiguous bed and roughness on U's device/dtype (cached between calls)."""
        p = self.dem_correction_parameter
        key = (U.device, U.dtype, self.bed_reference.data_ptr(), self.bed_reference._version, p.data_ptr(),
               p._version, self.valid_mask._version, self.roughness_length.data_ptr(),
               self.roughness_length._version, self.maximum_dem_correction)
        hit = self._field_cache.get("fields")
        if hit is not None and hit[0] == key:
            return hit[1], hit[2]
        # cached tensors outlive the call: build them as normal tensors even under inference_mode
        with torch.inference_mode(False), torch.no_grad():
            bed = self.bed.detach().to(device=U.device, dtype=U.dtype).contiguous()
            rough = self.roughness_length.detach().to(device=U.device, dtype=U.dtype).contiguous()
        B = U.shape[0]
        self._expand(bed, B)
        self._expand(rough, B)
        self._field_cache["fields"] = (key, bed, rough)
        return bed, rough

    def _rows(self, U):
        if self.rows_hint:
            return int(self.rows_hint)
        from . import autotune

        return autotune.lookup(U.device, U.shape[-2], U.shape[-1], U.shape[0], U.dtype)

    @staticmethod
    def _check_source_grad(infiltration):
        """step / forward do not differentiate their source terms: refuse to drop a requested gradient."""
        if torch.is_grad_enabled() and torch.is_tensor(infiltration) and infiltration.requires_grad:
            raise NotImplementedError(
                "step() and forward() do not return gradients with respect to infiltration; "
                "spatial_operator() does (or pass a tensor that does not require grad)")

    def _fast(self):
        if self.arithmetic not in ("reference", "fast"):
            raise ValueError(f"arithmetic must be 'reference' or 'fast', got {self.arithmetic!r}")
        return self.arithmetic == "fast"

    def _workspace(self, U, bed, rough, fast=False):
        B, _, ny, nx = U.shape
        key = (U.device, U.dtype, B, ny, nx, bed.shape[0], rough.shape[0])
        ws = self._ws.get(key)
        if ws is None:
            lib = _lib.load()
            grid = _lib.make_grid(B, ny, nx, bed.shape[0], rough.shape[0], self.dx, self.dy)
            nbytes = lib.swe_workspace_bytes(grid, U.element_size())
            if nbytes == 0:
                raise ValueError(f"invalid grid {B}x{ny}x{nx}: {lib.swe_last_error_message().decode()}")
            with torch.inference_mode(False):
                buf = torch.empty(nbytes, dtype=torch.uint8, device=U.device)
            _lib.check(lib.swe_workspace_init(_lib.ptr(buf), nbytes, _lib.stream_ptr(U.device)), "workspace init")
            ws = dict(buf=buf, grid=grid, parity=0)
            self._ws[key] = ws
        ws["cfg"] = _lib.make_config(self.cfg)
        ws["cfg"].grad_bed = int(getattr(self, "_grad_bed", False))  # set by the adjoint during a backward
        ws["cfg"].arithmetic = 1 if fast else 0
        return ws

    def _level_arg(self, level, U):
        lvl = self._prepare_boundary_level(level, U).detach()
        B = U.shape[0]
        if lvl.numel() == 1:
            return lvl.reshape(1).contiguous(), 0
        if lvl.numel() == B and (lvl.ndim == 1 or lvl.shape == (B, 1, 1, 1)):
            return lvl.reshape(B).contiguous(), 1
        raise ValueError(
            "Boundary values must be a scalar, a one-element tensor, or a tensor with one value per batch "
            f"member. Received shape {tuple(lvl.shape)}.")

    def _forcing(self, U, infiltration, time):
        """Source terms of a step for the library (solver.py:520-525, 609-641), or None.

        Returns (swe_forcing struct, tensors it points at): the rainfall record
        (``self.rainfall.times`` [nt], ``.rates`` [nt,1,ny,nx]) is sampled on the
        device at each stage's model time; infiltration is broadcast to the
        depth shape (one plane when every member shares it)."""
        if self.rainfall is None and infiltration is None:
            return None, ()
        f = _lib.SweForcing()
        keep = []
        ny, nx = U.shape[-2:]
        if self.rainfall is not None:
            times = self.rainfall.times.to(device=U.device, dtype=U.dtype).contiguous()
            rates = self.rainfall.rates.to(device=U.device, dtype=U.dtype)
            nt = times.numel()
            if rates.shape != (nt, 1, ny, nx):
                raise ValueError(f"rainfall rates must have shape [{nt},1,{ny},{nx}], got {tuple(rates.shape)}")
            rates = rates.reshape(nt, ny, nx).contiguous()
            f.rain_times, f.rain_rates, f.rain_nt = times.data_ptr(), rates.data_ptr(), nt
            keep += [times, rates]
        if infiltration is not None:
            v = (infiltration.to(dtype=U.dtype, device=U.device) if torch.is_tensor(infiltration)
                 else U.new_tensor(infiltration))
            v = torch.broadcast_to(v, U[:, 0:1].shape)
            if U.shape[0] == 1 or bool((v == v[:1]).all()):
                plane, stride = v[:1].contiguous(), 0
            else:
                plane, stride = v.contiguous(), ny * nx
            f.infiltration, f.infil_bstride = plane.data_ptr(), stride
            keep.append(plane)
        f.time = float(time)
        return f, keep

    # ------------------------------------------------------------------ spatial operator
    @torch._dynamo.disable
    def spatial_operator(self, U, rainfall_rate=None, infiltration=None, return_wave_speed=False,
                         boundary_level=None):
        """Tendency dU/dt (src/solver.py:448-531); wave speeds are maxima over the batch."""
        self._check_state(U)
        level = self._prepare_boundary_level(
            self.cfg.constant_level if boundary_level is None else boundary_level, U)
        sources_grad = any(torch.is_tensor(v) and v.requires_grad for v in (rainfall_rate, infiltration))
        if torch.is_grad_enabled() and (U.requires_grad or level.requires_grad or self._bed_trainable()
                                        or sources_grad):
            from .adjoint import operator_adjoint
            rhs = operator_adjoint(self, U, level)
            if rainfall_rate is not None or infiltration is not None:
                # the source terms join the depth tendency after the fluxes, rain first
                # (src/solver.py:520-525), as the kernel adds them
                def source(v):
                    t = v.to(dtype=U.dtype, device=U.device) if torch.is_tensor(v) else U.new_tensor(v)
                    return torch.broadcast_to(t, U[:, 0:1].shape)

                h = rhs[:, 0:1]
                if rainfall_rate is not None:
                    h = h + source(rainfall_rate)
                if infiltration is not None:
                    h = h - torch.clamp(source(infiltration), min=0)
                rhs = torch.cat((h, rhs[:, 1:]), dim=1)
            if return_wave_speed:
                with torch.no_grad():
                    _, ax, ay = self.spatial_operator(U.detach(), return_wave_speed=True,
                                                      boundary_level=level.detach())
                return rhs, ax, ay
            return rhs
        U = U.detach().contiguous()
        bed, rough = self._fields(U)
        ws = self._workspace(U, bed, rough, fast=self._fast())
        lvl, stride = self._level_arg(level, U)
        lib = _lib.load()
        rhs = torch.empty_like(U)
        B = U.shape[0]

        def full(v):
            if v is None:
                return None
            t = v.to(dtype=U.dtype, device=U.device) if torch.is_tensor(v) else U.new_tensor(v)
            return torch.broadcast_to(t, U[:, 0:1].shape)

        rain_f, infil_f = full(rainfall_rate), full(infiltration)
        per_member = B > 1 and any(t is not None and not bool((t == t[:1]).all()) for t in (rain_f, infil_f))
        # shared fields go to the kernel; per-member ones are added after it, in the kernel's order
        # (flux tendency + rain - max(infiltration, 0), src/solver.py:520-525)
        rain = None if rain_f is None or per_member else rain_f[:1].contiguous()
        infil = None if infil_f is None or per_member else infil_f[:1].contiguous()
        fn = lib.swe_spatial_operator_f32 if U.dtype == torch.float32 else lib.swe_spatial_operator_f64
        _lib.check(fn(ws["grid"], ws["cfg"], _lib.ptr(U), _lib.ptr(bed), _lib.ptr(lvl), stride, _lib.ptr(rain),
                      _lib.ptr(infil), _lib.ptr(rhs), _lib.ptr(ws["buf"]), self._rows(U),
                      _lib.stream_ptr(U.device)), "spatial_operator")
        if per_member:
            h = rhs[:, 0:1]
            if rain_f is not None:
                h = h + rain_f
            if infil_f is not None:
                h = h - torch.clamp(infil_f, min=0)
            rhs = torch.cat((h, rhs[:, 1:]), dim=1)
        if return_wave_speed:
            ax, ay = self._bank_speeds(ws, U)
            return rhs, ax, ay
        return rhs

    def _bank_speeds(self, ws, U):
        nslots = 64
        word = 4 if U.dtype == torch.float32 else 8
        off = 256 + 2 * 2 * nslots * word
        raw = ws["buf"][off:off + 2 * nslots * word]
        vals = raw.view(U.dtype).reshape(2, nslots)
        bits = raw.view(torch.int32 if word == 4 else torch.int64).reshape(2, nslots)
        if word == 4:
            bits = bits.to(torch.int64) & 0xFFFFFFFF  # unsigned order, as the device atomics compare
        ix = bits.argmax(dim=1)  # largest bit pattern == largest non-negative value (NaN above all)
        return vals[0, ix[0]].clone(), vals[1, ix[1]].clone()

    def stable_dt(self, U, boundary_level=None):
        """CFL-stable timestep (src/solver.py:537-545)."""
        _, ax, ay = self.spatial_operator(U, return_wave_speed=True, boundary_level=boundary_level)
        denominator = torch.clamp(ax / self.dx + ay / self.dy, min=self.cfg.epsilon)
        return self.cfg.cfl / denominator

    def water_volume(self, U):
        """Water volume sum(h) * dx * dy of each member, as an FP64 tensor [B] on U's device.

        Summed on the device in FP64 in a fixed order (the same bits every
        call); the reference has no such diagnostic.  For conservation checks."""
        self._check_state(U, "U")
        bed, rough = self._fields(U)
        ws = self._workspace(U, bed, rough)
        state = U.detach().contiguous()
        out = torch.empty(U.shape[0], dtype=torch.float64, device=U.device)
        fn = getattr(_lib.load(), f"swe_water_volume_{_DTYPES[U.dtype]}")
        _lib.check(fn(ws["grid"], _lib.ptr(state), _lib.ptr(out), _lib.ptr(ws["buf"]), _lib.stream_ptr(U.device)),
                   "water_volume")
        return out

    # ------------------------------------------------------------------ time stepping
    @torch._dynamo.disable
    def step(self, U, model_time, max_dt, infiltration=None, boundary_level=None):
        """One adaptive midpoint step (src/solver.py:572-663); returns (U_new, dt as float).

        dt is detached exactly like the reference; with autograd enabled and U
        or the level requiring grad, the hand-written adjoint carries gradients.
        """
        self._check_state(U)
        level = self._prepare_boundary_level(
            self.cfg.constant_level if boundary_level is None else boundary_level, U)
        max_dt_value = self._scalar_float(max_dt, "max_dt") if torch.is_tensor(max_dt) else float(max_dt)
        self._check_source_grad(infiltration)
        forcing = None
        if self.rainfall is not None or infiltration is not None:
            # time_float = float(model_time) of a tensor in U's dtype (:608-611)
            t = model_time.detach() if torch.is_tensor(model_time) else U.new_tensor(model_time)
            forcing = (infiltration, float(t.cpu()))
        needs_grad = torch.is_grad_enabled() and (U.requires_grad or level.requires_grad or self._bed_trainable())
        if needs_grad:
            from .adjoint import differentiable_step
            return differentiable_step(self, U, level, max_dt_value, forcing)
        out, dt = self._step_raw(U.detach().contiguous(), level.detach(), max_dt_value, forcing)
        return out, dt

    def _step_raw(self, U, level, max_dt_value, forcing=None):
        bed, rough = self._fields(U)
        ws = self._workspace(U, bed, rough, fast=self._fast())
        lvl, stride = self._level_arg(level, U)
        lib = _lib.load()
        out = torch.empty_like(U)
        if forcing is not None:
            fz, keep = self._forcing(U, *forcing)
            fn = lib.swe_step_forced_f32 if U.dtype == torch.float32 else lib.swe_step_forced_f64
            _lib.check(fn(ws["grid"], ws["cfg"], ctypes.byref(fz), _lib.ptr(U), _lib.ptr(bed), _lib.ptr(rough),
                          _lib.ptr(lvl), stride, float(max_dt_value), ws["parity"], _lib.ptr(out), None, None,
                          _lib.ptr(ws["buf"]), self._rows(U), _lib.stream_ptr(U.device)), "step")
        else:
            fn = lib.swe_step_f32 if U.dtype == torch.float32 else lib.swe_step_f64
            _lib.check(fn(ws["grid"], ws["cfg"], _lib.ptr(U), _lib.ptr(bed), _lib.ptr(rough), _lib.ptr(lvl),
                          stride, float(max_dt_value), ws["parity"], _lib.ptr(out), _lib.ptr(ws["buf"]),
                          self._rows(U), _lib.stream_ptr(U.device)), "step")
        ws["parity"] ^= 1
        dt = _lib.read_control(ws["buf"]).last_dt
        if not math.isfinite(dt):
            raise FloatingPointError("The calculated timestep is non-finite.")
        if dt <= 0.0:
            raise FloatingPointError(f"Non-positive timestep: {dt}.")
        return out, dt

    # ------------------------------------------------------------------ forward integration
    @staticmethod
    def _scalar_float(value, name: str) -> float:
        if torch.is_tensor(value):
            if value.numel() != 1:
                raise ValueError(f"{name} must be scalar.")
            return float(value.detach().cpu())
        if isinstance(value, Real):
            return float(value)
        raise TypeError(f"{name} must be a real scalar or scalar tensor.")

    @torch._dynamo.disable
    def forward(self, U0, t_end, start_time=0.0, max_steps=1_000_000, infiltration=None, callback=None,
                boundary_level=None, max_dt_cap=None):
        """Advance for ``t_end`` seconds (src/solver.py:680-779); returns (U, hmax)."""
        if U0.ndim != 4 or U0.shape[1] != 3:
            raise ValueError(f"U0 must have shape [batch,3,ny,nx], got {tuple(U0.shape)}.")
        duration = self._scalar_float(t_end, "t_end")
        absolute_start = self._scalar_float(start_time, "start_time")
        if not isinstance(max_steps, Integral) or isinstance(max_steps, bool) or max_steps <= 0:
            raise ValueError(f"max_steps must be positive, got {max_steps!r}.")
        if duration < 0.0:
            raise ValueError(f"t_end duration cannot be negative: {duration}.")
        if duration == 0.0:
            return U0, U0[:, 0].clone()
        self._check_state(U0, "U0")
        cap = None if max_dt_cap is None else self._scalar_float(max_dt_cap, "max_dt_cap")
        self._check_source_grad(infiltration)
        level_req = boundary_level is not None and torch.is_tensor(boundary_level) and boundary_level.requires_grad
        dynamic = boundary_level is None and self.use_dynamic_bc
        series_req = dynamic and torch.is_tensor(self.bc_wls) and self.bc_wls.requires_grad
        needs_grad = torch.is_grad_enabled() and (U0.requires_grad or level_req or series_req
                                                  or self._bed_trainable())
        forced = self.rainfall is not None or infiltration is not None
        if needs_grad:
            # the per-step work of the loop (src/solver.py:765-801: gauges, callback, a
            # differentiable hmax) and forcing take the host loop of differentiable steps,
            # like the reference's; the rest runs device-driven with the hand-written adjoint
            if (callback is not None or self.track_gauges or forced or self.differentiable_hmax
                    or (dynamic and U0.shape[0] != 1)):
                return self._forward_host(U0, duration, absolute_start, max_steps, callback, boundary_level, cap,
                                          infiltration)
            if dynamic:
                from .chunked import dynamic_forward
                return dynamic_forward(self, U0, duration, absolute_start, max_steps, cap)
            from .adjoint import differentiable_forward
            return differentiable_forward(self, U0, duration, absolute_start, max_steps, boundary_level, cap)
        if callback is not None:
            return self._forward_host(U0, duration, absolute_start, max_steps, callback, boundary_level, cap,
                                      infiltration)
        return self._forward_device(U0.detach().contiguous(), duration, absolute_start, max_steps,
                                    boundary_level, cap, infiltration)

    def _forward_host(self, U0, duration, absolute_start, max_steps, callback, boundary_level, cap,
                      infiltration=None):
        U, hmax, elapsed = U0, U0[:, 0].clone(), 0.0
        tolerance = max(1e-12, 1e-12 * abs(duration))
        for iteration in range(max_steps):
            remaining = duration - elapsed
            if remaining <= tolerance:
                return U, hmax
            step_max_dt = remaining if cap is None else min(remaining, cap)
            current_time = absolute_start + elapsed
            level = self._resolve_boundary_level(current_time, U, boundary_level)
            U, dt_value = self.step(U, model_time=U.new_tensor(current_time),
                                    max_dt=float(U.new_tensor(step_max_dt)), infiltration=infiltration,
                                    boundary_level=level)
            hmax = torch.maximum(hmax, U[:, 0])
            elapsed += dt_value
            if elapsed > duration and elapsed - duration <= tolerance:
                elapsed = duration
            if self.track_gauges:
                self._record_gauges(U, absolute_start + elapsed, dt_value)
            if callback is not None:
                callback(iteration, absolute_start + elapsed, U)
        raise RuntimeError(
            f"max_steps={max_steps} reached after {elapsed:.8f} s; requested duration={duration:.8f} s.")

    def _forward_device(self, U0, duration, absolute_start, max_steps, boundary_level, cap, infiltration=None):
        """Device-driven loop: dt, elapsed time and stopping live on the GPU."""
        lib = _lib.load()
        bed, rough = self._fields(U0)
        ws = self._workspace(U0, bed, rough, fast=self._fast())
        dynamic = boundary_level is None and self.use_dynamic_bc
        if dynamic:
            # the series is interpolated on the device at every step's start time (src/solver.py:391-415)
            from .chunked import dynamic_arg, dynamic_series

            series = dynamic_series(self, U0)
            dyn = dynamic_arg(*series, absolute_start)
            lvl, stride = None, 0
        else:
            level = self._resolve_boundary_level(absolute_start, U0, boundary_level)
            lvl, stride = self._level_arg(level, U0)
        stream = _lib.stream_ptr(U0.device)
        Ua = U0.clone()
        Ub = torch.empty_like(U0)
        hmax = U0[:, 0].clone()
        names = list(self.gauge_indices) if self.track_gauges else []
        ny, nx = U0.shape[-2:]
        for n_ in names:
            iy, ix = self.gauge_indices[n_]
            if not (0 <= int(iy) < ny and 0 <= int(ix) < nx):
                raise IndexError(f"gauge {n_!r} at {(iy, ix)} lies outside the {ny}x{nx} grid")
        if len(names) > 1024:
            raise ValueError("at most 1024 gauges are tracked on the device")
        gauges = (torch.tensor([list(self.gauge_indices[n]) for n in names], dtype=torch.int32,
                               device=U0.device).contiguous() if names else None)
        capacity = 256
        dts = torch.empty(capacity, dtype=U0.dtype, device=U0.device)
        gd = torch.empty(capacity, max(len(names), 1), dtype=U0.dtype, device=U0.device)
        gl = torch.empty_like(gd)
        _lib.check(lib.swe_run_begin(_lib.ptr(ws["buf"]), float(duration), -1.0 if cap is None else float(cap),
                                     U0.element_size(), stream), "run_begin")
        fz, keep = self._forcing(U0, infiltration, absolute_start)
        if dynamic:
            adv = lib.swe_advance_dynamic_f32 if U0.dtype == torch.float32 else lib.swe_advance_dynamic_f64

            def fn(grid, cfg, bed_, rough_, _lvl, _stride, Ua_, Ub_, first, n, hmax_, dts_, cap_, gauges_, n_g, gd_,
                   gl_, _gcap, ws_, rows, stream_):
                return adv(grid, cfg, ctypes.byref(dyn), None if fz is None else ctypes.byref(fz), bed_, rough_, Ua_,
                           Ub_, None, None, 0, first, n, hmax_, dts_, cap_, gauges_, n_g, gd_, gl_, ws_, rows, stream_)
        elif fz is None:
            fn = lib.swe_advance_f32 if U0.dtype == torch.float32 else lib.swe_advance_f64
        else:
            forced = lib.swe_advance_forced_f32 if U0.dtype == torch.float32 else lib.swe_advance_forced_f64

            def fn(*args):
                return forced(args[0], args[1], ctypes.byref(fz), *args[2:])
        launched, chunk = 0, 8
        ctl = None
        while launched < max_steps:
            n = min(chunk, max_steps - launched)
            if launched + n > capacity:
                capacity = max(2 * capacity, launched + n)
                dts = torch.cat((dts, torch.empty(capacity - dts.numel(), dtype=dts.dtype, device=dts.device)))
                grow = torch.empty(capacity - gd.shape[0], gd.shape[1], dtype=gd.dtype, device=gd.device)
                gd, gl = torch.cat((gd, grow)), torch.cat((gl, grow.clone()))
            _lib.check(fn(ws["grid"], ws["cfg"], _lib.ptr(bed), _lib.ptr(rough), _lib.ptr(lvl), stride,
                          _lib.ptr(Ua), _lib.ptr(Ub), launched, n, _lib.ptr(hmax), _lib.ptr(dts), capacity,
                          _lib.ptr(gauges), len(names), _lib.ptr(gd) if names else None,
                          _lib.ptr(gl) if names else None, capacity, _lib.ptr(ws["buf"]), self._rows(U0),
                          stream), "advance")
            launched += n
            ctl = _lib.read_control(ws["buf"])
            if ctl.done:
                break
            # aim the next chunk at the remaining duration, from the latest dt
            remaining = duration - ctl.elapsed[ctl.steps & 1]
            est = int(remaining / ctl.last_dt) + 2 if ctl.last_dt > 0 else chunk
            chunk = max(4, min(1024, est))
        steps = ctl.steps
        sync_parity(ws, ctl)
        if ctl.status != 0:
            if ctl.status == 1:
                raise FloatingPointError("The calculated timestep is non-finite.")
            raise FloatingPointError(f"Non-positive timestep: {ctl.bad_dt}.")
        U = Ua if steps % 2 == 0 else Ub
        if self.track_gauges and steps > 0:
            dt_host = dts[:steps].cpu().tolist()
            depth = gd[:steps].cpu().tolist()
            lev = gl[:steps].cpu().tolist()
            elapsed = 0.0
            tol = max(1e-12, 1e-12 * abs(duration))
            for k in range(steps):
                elapsed += dt_host[k]
                if elapsed > duration and elapsed - duration <= tol:
                    elapsed = duration
                self.ts_time.append(absolute_start + elapsed)
                self.ts_dt.append(dt_host[k])
                for j, name in enumerate(names):
                    self.ts_h.setdefault(name, []).append(depth[k][j])
                    self.ts_z.setdefault(name, []).append(lev[k][j])
        self.last_dt_series = dts[:steps]
        if not ctl.done:
            elapsed = ctl.elapsed[steps & 1]
            raise RuntimeError(
                f"max_steps={max_steps} reached after {elapsed:.8f} s; requested duration={duration:.8f} s.")
        return U, hmax

    def _record_gauges(self, U, absolute_time: float, dt_value: float) -> None:
        """Per-step gauge records of member 0 (src/solver.py:781-801)."""
        self.ts_time.append(absolute_time)
        self.ts_dt.append(dt_value)
        bed = self.bed
        for name, (iy, ix) in self.gauge_indices.items():
            depth = float(U[0, 0, iy, ix].detach().cpu())
            z = float(bed[0, 0, iy, ix].detach().cpu())
            self.ts_h.setdefault(name, []).append(depth)
            self.ts_z.setdefault(name, []).append(depth + z)

    # ------------------------------------------------------------------ DEM regularisation
    def dem_regularization(self, prior_weight=1.0, slope_weight=0.0, curvature_weight=0.0):
        """Prior / slope / curvature penalty of the DEM correction (src/solver.py:807-836)."""
        bed = self.bed
        correction = bed - self.bed_reference
        gx, gy = bed[..., 1:] - bed[..., :-1], bed[..., 1:, :] - bed[..., :-1, :]
        prior = correction.square().mean()
        slope = gx.square().mean() + gy.square().mean()
        cx = (gx[..., 1:] - gx[..., :-1]).square().mean() if gx.shape[-1] >= 2 else bed.new_zeros(())
        cy = (gy[..., 1:, :] - gy[..., :-1, :]).square().mean() if gy.shape[-2] >= 2 else bed.new_zeros(())
        return (float(prior_weight) * prior + float(slope_weight) * slope
                + float(curvature_weight) * (cx + cy))
