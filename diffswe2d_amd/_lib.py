"""ctypes binding of libswe2d_hip.so (the C ABI declared in include/swe2d_hip.h).

The library is built in-tree by ``diffswe2d_amd.build``.  There is no fallback:
if the library is missing or fails to load, every call raises.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import torch

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libswe2d_hip.so"

SWE_ABI_VERSION = 5
BC_CODES = {"transmissive": 0, "periodic": 1, "constant": 2, "water_level": 3, "wall": 4, "absorbing": 5}
FRICTION_CODES = {"manning": 0}

ERR_DT_NONFINITE = -5
ERR_DT_NONPOSITIVE = -6


class SweGrid(ctypes.Structure):
    _fields_ = [("batch", ctypes.c_int32), ("ny", ctypes.c_int32), ("nx", ctypes.c_int32),
                ("bed_batch", ctypes.c_int32), ("rough_batch", ctypes.c_int32), ("pad", ctypes.c_int32),
                ("dx", ctypes.c_double), ("dy", ctypes.c_double)]


class SweConfigC(ctypes.Structure):
    _fields_ = [("gravity", ctypes.c_double), ("dry_depth", ctypes.c_double), ("epsilon", ctypes.c_double),
                ("cfl", ctypes.c_double), ("limiter_theta", ctypes.c_double),
                ("constant_value", ctypes.c_double), ("boundary", ctypes.c_int32 * 4),
                ("water_slope_reset", ctypes.c_int32), ("positivity_limiter", ctypes.c_int32),
                ("apply_friction", ctypes.c_int32), ("friction_model", ctypes.c_int32),
                ("grad_bed", ctypes.c_int32), ("arithmetic", ctypes.c_int32), ("fixed_dt", ctypes.c_int32)]


class SweScenario(ctypes.Structure):
    _fields_ = [("ny", ctypes.c_int32), ("nx", ctypes.c_int32), ("dx", ctypes.c_double), ("dy", ctypes.c_double),
                ("U_a", ctypes.c_void_p), ("U_b", ctypes.c_void_p), ("k1", ctypes.c_void_p),
                ("bed", ctypes.c_void_p), ("roughness", ctypes.c_void_p), ("levels", ctypes.c_void_p),
                ("n_chunks", ctypes.c_int32), ("chunk_seconds", ctypes.c_double), ("gauges", ctypes.c_void_p),
                ("n_gauges", ctypes.c_int32), ("gauge_out", ctypes.c_void_p), ("rows_hint", ctypes.c_int32)]


class SweForcing(ctypes.Structure):
    _fields_ = [("rain_times", ctypes.c_void_p), ("rain_rates", ctypes.c_void_p), ("rain_nt", ctypes.c_int32),
                ("pad", ctypes.c_int32), ("infiltration", ctypes.c_void_p), ("infil_bstride", ctypes.c_int64),
                ("time", ctypes.c_double)]


class SweDynamicLevel(ctypes.Structure):
    _fields_ = [("times", ctypes.c_void_p), ("levels", ctypes.c_void_p), ("n", ctypes.c_int32),
                ("pad", ctypes.c_int32), ("start_time", ctypes.c_double), ("step_level", ctypes.c_void_p),
                ("step_index", ctypes.c_void_p), ("step_weight", ctypes.c_void_p)]


class SweControl(ctypes.Structure):
    _fields_ = [("elapsed", ctypes.c_double * 2), ("duration", ctypes.c_double), ("tol", ctypes.c_double),
                ("cap", ctypes.c_double), ("bad_dt", ctypes.c_double), ("done", ctypes.c_int32),
                ("steps", ctypes.c_int32), ("status", ctypes.c_int32), ("pad", ctypes.c_int32),
                ("last_dt", ctypes.c_double), ("chunked", ctypes.c_int32), ("chunk", ctypes.c_int32),
                ("n_chunks", ctypes.c_int32), ("pad2", ctypes.c_int32), ("level_f64", ctypes.c_double),
                ("level_f32", ctypes.c_float), ("pad3", ctypes.c_float), ("levels", ctypes.c_void_p)]


_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_G = ctypes.POINTER(SweGrid)
_C = ctypes.POINTER(SweConfigC)
_F = ctypes.POINTER(SweForcing)
_DL = ctypes.POINTER(SweDynamicLevel)

_SIGNATURES = {
    "swe_abi_version": ([], _I),
    "swe_strerror": ([_I], ctypes.c_char_p),
    "swe_last_error_message": ([], ctypes.c_char_p),
    "swe_control_bytes": ([], ctypes.c_size_t),
    "swe_workspace_bytes": ([_G, _I], ctypes.c_size_t),
    "swe_workspace_init": ([_P, ctypes.c_size_t, _P], _I),
    "swe_spatial_operator_f32": ([_G, _C, _P, _P, _P, _I, _P, _P, _P, _P, _I, _P], _I),
    "swe_spatial_operator_f64": ([_G, _C, _P, _P, _P, _I, _P, _P, _P, _P, _I, _P], _I),
    "swe_step_f32": ([_G, _C, _P, _P, _P, _P, _I, _D, _I, _P, _P, _I, _P], _I),
    "swe_step_f64": ([_G, _C, _P, _P, _P, _P, _I, _D, _I, _P, _P, _I, _P], _I),
    "swe_step_stage_f32": ([_G, _C, _I, _P, _P, _P, _P, _I, _D, _I, _P, _P, _I, _P], _I),
    "swe_step_noalloc": ([_G, _C, _I, _P, _P, _P, _P, _I, _D, _P, _P, _I, _P], _I),
    "swe_step_record_f32": ([_G, _C, _P, _P, _P, _P, _I, _D, _I, _P, _P, _P, _P, _I, _P], _I),
    "swe_step_record_f64": ([_G, _C, _P, _P, _P, _P, _I, _D, _I, _P, _P, _P, _P, _I, _P], _I),
    "swe_run_begin": ([_P, _D, _D, _I, _P], _I),
    "swe_advance_f32": ([_G, _C, _P, _P, _P, _I, _P, _P, _I, _I, _P, _P, _I, _P, _I, _P, _P, _I, _P, _I, _P], _I),
    "swe_advance_f64": ([_G, _C, _P, _P, _P, _I, _P, _P, _I, _I, _P, _P, _I, _P, _I, _P, _P, _I, _P, _I, _P], _I),
    "swe_step_forced_f32": ([_G, _C, _F, _P, _P, _P, _P, _I, _D, _I, _P, _P, _P, _P, _I, _P], _I),
    "swe_step_forced_f64": ([_G, _C, _F, _P, _P, _P, _P, _I, _D, _I, _P, _P, _P, _P, _I, _P], _I),
    "swe_advance_forced_f32": ([_G, _C, _F, _P, _P, _P, _I, _P, _P, _I, _I, _P, _P, _I, _P, _I, _P, _P, _I, _P, _I,
                                _P], _I),
    "swe_advance_forced_f64": ([_G, _C, _F, _P, _P, _P, _I, _P, _P, _I, _I, _P, _P, _I, _P, _I, _P, _P, _I, _P, _I,
                                _P], _I),
    "swe_advance_dynamic_f32": ([_G, _C, _DL, _F, _P, _P, _P, _P, _P, _P, _I, _I, _I, _P, _P, _I, _P, _I, _P, _P,
                                 _P, _I, _P], _I),
    "swe_advance_dynamic_f64": ([_G, _C, _DL, _F, _P, _P, _P, _P, _P, _P, _I, _I, _I, _P, _P, _I, _P, _I, _P, _P,
                                 _P, _I, _P], _I),
    "swe_run_begin_chunked": ([_P, _D, _I, _P, _I, _P], _I),
    "swe_advance_chunked_f32": ([_G, _C, _P, _P, _P, _P, _P, _P, _I, _I, _I, _P, _P, _I, _P, _I, _P, _P, _I, _P], _I),
    "swe_advance_chunked_f64": ([_G, _C, _P, _P, _P, _P, _P, _P, _I, _I, _I, _P, _P, _I, _P, _I, _P, _P, _I, _P], _I),
    "swe_advance_tape_f32": ([_G, _C, _P, _P, _P, _I, _P, _P, _P, _I, _I, _I, _P, _P, _I, _P, _I, _P], _I),
    "swe_advance_tape_f64": ([_G, _C, _P, _P, _P, _I, _P, _P, _P, _I, _I, _I, _P, _P, _I, _P, _I, _P], _I),
    "swe_adjoint_workspace_bytes": ([_G, _I], ctypes.c_size_t),
    "swe_adjoint_bed_grad_offset": ([_G, _I], ctypes.c_size_t),
    "swe_adjoint_last_error_message": ([], ctypes.c_char_p),
    "swe_operator_adjoint_f32": ([_G, _C, _P, _P, _P, _I, _P, _P, _P, _P, _I, _P], _I),
    "swe_operator_adjoint_f64": ([_G, _C, _P, _P, _P, _I, _P, _P, _P, _P, _I, _P], _I),
    "swe_step_adjoint_f32": ([_G, _C, _P, _P, _P, _P, _P, _P, _I, _D, _P, _P, _P, _P, _I, _P], _I),
    "swe_step_adjoint_f64": ([_G, _C, _P, _P, _P, _P, _P, _P, _I, _D, _P, _P, _P, _P, _I, _P], _I),
    "swe_segment_tape_bytes": ([_G, _I, _I], ctypes.c_size_t),
    "swe_segment_adjoint_f32": ([_G, _C, _P, _P, _P, _P, _I, _P, _I, _I, _P, _P, _P, _P, _P, _P, _I, _P], _I),
    "swe_segment_adjoint_f64": ([_G, _C, _P, _P, _P, _P, _I, _P, _I, _I, _P, _P, _P, _P, _P, _P, _I, _P], _I),
    "swe_segment_reverse_f32": ([_G, _C, _P, _P, _P, _P, _P, _P, _I, _P, _I, _P, _P, _P, _P, _I, _P], _I),
    "swe_segment_reverse_f64": ([_G, _C, _P, _P, _P, _P, _P, _P, _I, _P, _I, _P, _P, _P, _P, _I, _P], _I),
    "swe_segment_reverse_chunked_f32": ([_G, _C, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I, _P, _I, _P, _P, _P, _P, _P,
                                         _I, _P], _I),
    "swe_segment_reverse_chunked_f64": ([_G, _C, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I, _P, _I, _P, _P, _P, _P, _P,
                                         _I, _P], _I),
    "swe_segment_adjoint_chunked_f32": ([_G, _C, _P, _P, _P, _P, _P, _P, _P, _I, _I, _P, _I, _P, _P, _P, _P, _P, _P,
                                         _P, _I, _P], _I),
    "swe_segment_adjoint_chunked_f64": ([_G, _C, _P, _P, _P, _P, _P, _P, _P, _I, _I, _P, _I, _P, _P, _P, _P, _P, _P,
                                         _P, _I, _P], _I),
    "swe_valu_probe": ([_P, _I, _I, _P], _I),
    "swe_valu_mix_probe": ([_P, _I, _I, _P], _I),
    "swe_division_probe": ([_P, _P, ctypes.c_longlong, _P, _P, _P, _P], _I),
    "swe_sqrt_probe": ([_P, ctypes.c_longlong, _P, _P, _P], _I),
    "swe_pow_probe": ([_P, ctypes.c_longlong, ctypes.c_float, _P, _P, _P], _I),
    "swe_lane_shift_probe": ([_P, ctypes.c_longlong, _P, _P, _P, _P, _P], _I),
    "swe_water_volume_f32": ([_G, _P, _P, _P, _P], _I),
    "swe_water_volume_f64": ([_G, _P, _P, _P, _P], _I),
    "swe_ensemble_workspace_bytes": ([_I], ctypes.c_size_t),
    "swe_ensemble_last_error_message": ([], ctypes.c_char_p),
    "swe_ensemble_init": ([_C, _I, _P, _P, ctypes.c_size_t, ctypes.POINTER(ctypes.c_longlong), _P], _I),
    "swe_ensemble_advance": ([_P, _I, ctypes.c_longlong, _I, _I, _I, _P], _I),
    "swe_ensemble_clock_offset": ([_I], ctypes.c_size_t),
    "swe_ensemble_clock_bytes": ([], ctypes.c_size_t),
    "sur_last_error_message": ([], ctypes.c_char_p),
    "sur_gemm_nt_bf16": ([_I, _I, _I, _P, ctypes.c_long, _P, ctypes.c_long, _P, ctypes.c_long, _I, _P, _P], _I),
    "sur_im2col": ([_P, _I, _I, _I, _I, _I, _I, _I, _P, _P], _I),
    "sur_col2im": ([_P, _I, _I, _I, _I, _I, _P, _P], _I),
    "sur_bias_act": ([_P, _P, _I, ctypes.c_long, _I, _P, _P, _P, _P], _I),
    "sur_act_bwd": ([_P, _P, _I, ctypes.c_long, _I, _P, _P, _P, _I, _P], _I),
    "sur_bias_grad": ([_P, _I, ctypes.c_long, _P, _P], _I),
    "sur_cast_bf16": ([_P, ctypes.c_long, _P, _P], _I),
    "sur_cast_pad_rows": ([_P, _I, _I, _I, _P, _P], _I),
    "sur_transpose_bf16": ([_P, _I, _I, _I, _P, _P], _I),
}

_lib = None


def exported_symbols():
    """Names every library build must export (the declarations of include/swe2d_hip.h)."""
    return tuple(_SIGNATURES)


def load(path: Path | None = None):
    """Load (once) and type the library; raises if it is missing or incompatible."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path is not None else LIB_PATH
    if not p.exists():
        raise RuntimeError(
            f"HIP library {p} is missing; build it with `python -m diffswe2d_amd.build`"
        )
    lib = ctypes.CDLL(str(p))
    for name, (args, res) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    if lib.swe_abi_version() != SWE_ABI_VERSION:
        raise RuntimeError(f"libswe2d_hip ABI {lib.swe_abi_version()} != expected {SWE_ABI_VERSION}")
    if lib.swe_control_bytes() != ctypes.sizeof(SweControl):
        raise RuntimeError("swe_control layout mismatch between header and binding")
    if path is None:
        _lib = lib
    return lib


def check(code: int, what: str, source: str = "swe_last_error_message"):
    """Raise RuntimeError for a non-zero status; source names the message function of the entry point's file."""
    if code != 0:
        lib = load()
        msg = getattr(lib, source)().decode(errors="replace")
        raise RuntimeError(f"{what} failed ({code}: {lib.swe_strerror(code).decode()}): {msg}")


def ptr(t):
    return ctypes.c_void_p(0) if t is None else ctypes.c_void_p(t.data_ptr())


def stream_ptr(device) -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def make_config(cfg) -> SweConfigC:
    """Translate an SWEConfig into the ABI struct (strings -> codes)."""
    c = SweConfigC()
    c.gravity, c.dry_depth, c.epsilon = float(cfg.gravity), float(cfg.dry_depth), float(cfg.epsilon)
    c.cfl, c.limiter_theta = float(cfg.cfl), float(cfg.limiter_theta)
    c.constant_value = float(cfg.constant_value)
    for i, side in enumerate((cfg.boundary_left, cfg.boundary_right, cfg.boundary_top, cfg.boundary_bottom)):
        if side not in BC_CODES:
            raise ValueError(
                f"Unsupported {('left', 'right', 'top', 'bottom')[i]} boundary type {side!r}. "
                f"Valid types are {sorted(BC_CODES)}."
            )
        c.boundary[i] = BC_CODES[side]
    c.water_slope_reset = int(bool(cfg.water_slope_reset))
    c.positivity_limiter = int(bool(cfg.positivity_limiter))
    c.apply_friction = int(bool(cfg.apply_friction))
    c.friction_model = 0 if cfg.frictionmodel == "manning" else 1
    return c


def make_grid(batch, ny, nx, bed_batch, rough_batch, dx, dy) -> SweGrid:
    g = SweGrid()
    g.batch, g.ny, g.nx = int(batch), int(ny), int(nx)
    g.bed_batch, g.rough_batch = int(bed_batch), int(rough_batch)
    g.dx, g.dy = float(dx), float(dy)
    return g


def read_control(ws: torch.Tensor) -> SweControl:
    raw = ws[: ctypes.sizeof(SweControl)].cpu().numpy().tobytes()
    return SweControl.from_buffer_copy(raw)
