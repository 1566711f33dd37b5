"""ctypes access to libhostcheck.so: the kernel math run on the host (tests only)."""
import ctypes
from pathlib import Path

import numpy as np
import torch

from diffswe2d_amd import _lib

HERE = Path(__file__).resolve().parent
LIB = HERE / "libhostcheck.so"


def build():
    import subprocess

    root = HERE.parent.parent
    cmd = ["/opt/rocm/bin/hipcc", "-x", "c++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
           "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", f"-I{root / 'include'}",
           f"-I{root / 'diffswe2d_amd' / 'csrc'}", str(HERE / "hostcheck.cpp"), "-o", str(LIB)]
    subprocess.run(cmd, check=True, capture_output=True)


_h = None


def lib():
    global _h
    if _h is None:
        root = HERE.parent.parent
        sources = [HERE / "hostcheck.cpp", *sorted((root / "diffswe2d_amd" / "csrc").glob("*.h"))]
        if not LIB.exists() or LIB.stat().st_mtime < max(p.stat().st_mtime for p in sources):
            build()
        _h = ctypes.CDLL(str(LIB))
    return _h


def _p(a):
    return ctypes.c_void_p(a.ctypes.data)


def _np(t, dtype):
    return np.ascontiguousarray(t.detach().cpu().numpy() if torch.is_tensor(t) else t, dtype=dtype)


def _level(level, B, dtype):
    lv = np.atleast_1d(_np(level, dtype)).reshape(-1)
    return (lv, 0) if lv.size == 1 else (lv, 1)


def _args(cfg, U, bed, dx, dy, dtype):
    B, _, ny, nx = U.shape
    return _lib.make_config(cfg), B, ny, nx, bed.shape[0], ctypes.c_double(dx), ctypes.c_double(dy)


def suffix(dtype):
    return "f64" if dtype == np.float64 else "f32"


def spatial_operator(cfg, U, bed, dx, dy, level):
    dtype = np.float64 if U.dtype == torch.float64 else np.float32
    Un, bn = _np(U, dtype), _np(bed, dtype)
    lv, ls = _level(level, Un.shape[0], dtype)
    c, B, ny, nx, bb, cdx, cdy = _args(cfg, Un, bn, dx, dy, dtype)
    rhs = np.zeros_like(Un)
    axay = np.zeros(2, dtype=dtype)
    getattr(lib(), f"hc_spatial_operator_{suffix(dtype)}")(
        ctypes.byref(c), B, ny, nx, bb, cdx, cdy, _p(Un), _p(bn), _p(lv), ls, _p(rhs), _p(axay))
    return rhs, axay


def spatial_operator_bwd(cfg, U, bed, dx, dy, level, lam):
    dtype = np.float64 if U.dtype == torch.float64 else np.float32
    Un, bn, ln = _np(U, dtype), _np(bed, dtype), _np(lam, dtype)
    lv, ls = _level(level, Un.shape[0], dtype)
    c, B, ny, nx, bb, cdx, cdy = _args(cfg, Un, bn, dx, dy, dtype)
    gU = np.zeros_like(Un)
    gl = np.zeros(B, dtype=dtype)
    getattr(lib(), f"hc_spatial_operator_bwd_{suffix(dtype)}")(
        ctypes.byref(c), B, ny, nx, bb, cdx, cdy, _p(Un), _p(bn), _p(lv), ls, _p(ln), _p(gU), _p(gl))
    return gU, gl


def step_fwd(cfg, U, bed, rough, dx, dy, level, dt):
    dtype = np.float64 if U.dtype == torch.float64 else np.float32
    Un, bn, rn = _np(U, dtype), _np(bed, dtype), _np(rough, dtype)
    lv, ls = _level(level, Un.shape[0], dtype)
    c, B, ny, nx, bb, cdx, cdy = _args(cfg, Un, bn, dx, dy, dtype)
    k1, k2, out = np.zeros_like(Un), np.zeros_like(Un), np.zeros_like(Un)
    cdt = ctypes.c_double(dt) if dtype == np.float64 else ctypes.c_float(dt)
    getattr(lib(), f"hc_step_fwd_{suffix(dtype)}")(
        ctypes.byref(c), B, ny, nx, bb, cdx, cdy, _p(Un), _p(bn), _p(rn), _p(lv), ls, cdt, _p(k1), _p(k2), _p(out))
    return out, k1, k2


def step_bwd(cfg, U, bed, rough, dx, dy, level, dt, lam):
    dtype = np.float64 if U.dtype == torch.float64 else np.float32
    Un, bn, rn, ln = _np(U, dtype), _np(bed, dtype), _np(rough, dtype), _np(lam, dtype)
    lv, ls = _level(level, Un.shape[0], dtype)
    c, B, ny, nx, bb, cdx, cdy = _args(cfg, Un, bn, dx, dy, dtype)
    gU = np.zeros_like(Un)
    gl = np.zeros(B, dtype=dtype)
    cdt = ctypes.c_double(dt) if dtype == np.float64 else ctypes.c_float(dt)
    getattr(lib(), f"hc_step_bwd_{suffix(dtype)}")(
        ctypes.byref(c), B, ny, nx, bb, cdx, cdy, _p(Un), _p(bn), _p(rn), _p(lv), ls, cdt, _p(ln), _p(gU), _p(gl))
    return gU, gl
