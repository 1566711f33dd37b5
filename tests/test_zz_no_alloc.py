"""swe_step_noalloc (the step entry point that allocates nothing) against swe_step_*.

Two calls in a row on one GPU share the fixed scratch; the second one is on a
smaller grid, so anything the first call left there must not leak into it.  The
output state and dt must equal swe_step_f32 / swe_step_f64's bit for bit (the
same kernels run on the same inputs).  Grids beyond the scratch's capacity are
rejected on the host before any launch.
"""
import ctypes

import numpy as np
import pytest
import torch

import diffswe2d_amd as swe
from diffswe2d_amd import _lib

pytestmark = pytest.mark.gpu

DT = {4: torch.float32, 8: torch.float64}


def case(seed, B, ny, nx, dtype, dev):
    g = torch.Generator().manual_seed(seed)
    h = 0.5 + torch.rand(B, 1, ny, nx, generator=g, dtype=torch.float64)
    h[:, :, : ny // 3, : nx // 4] = 0.0  # a dry corner: wet/dry faces
    hu = 0.3 * torch.randn(B, 1, ny, nx, generator=g, dtype=torch.float64)
    hv = 0.3 * torch.randn(B, 1, ny, nx, generator=g, dtype=torch.float64)
    U = torch.cat((h, hu, hv), 1)
    U[:, 1:] *= (h > 0)
    bed = 0.2 * torch.rand(1, 1, ny, nx, generator=g, dtype=torch.float64)
    rough = torch.full((1, 1, ny, nx), 0.02, dtype=torch.float64)
    return U.to(dtype).to(dev), bed.to(dtype).to(dev), rough.to(dtype).to(dev)


def step_with_workspace(cfg_c, grid, U, bed, rough, lvl, dev):
    lib = _lib.load()
    word = U.element_size()
    ws = torch.zeros(lib.swe_workspace_bytes(grid, word), dtype=torch.uint8, device=dev)
    _lib.check(lib.swe_workspace_init(_lib.ptr(ws), ws.numel(), _lib.stream_ptr(dev)), "workspace_init")
    out = torch.empty_like(U)
    fn = lib.swe_step_f32 if word == 4 else lib.swe_step_f64
    _lib.check(fn(ctypes.byref(grid), ctypes.byref(cfg_c), _lib.ptr(U), _lib.ptr(bed), _lib.ptr(rough), _lib.ptr(lvl),
                  0, 1e30, 0, _lib.ptr(out), _lib.ptr(ws), 0, _lib.stream_ptr(dev)), "step")
    return out, _lib.read_control(ws).last_dt


@pytest.mark.parametrize("word", [4, 8])
def test_noalloc_step_equals_step(word, cuda_device):
    lib = _lib.load()
    dtype = DT[word]
    cfg = swe.SWEConfig(boundary_left="water_level", boundary_right="wall", boundary_top="transmissive",
                        boundary_bottom="periodic", apply_friction=True)
    cfg_c = _lib.make_config(cfg)
    lvl = torch.tensor([0.9], dtype=dtype, device=cuda_device)
    # second call smaller than the first
    for seed, (B, ny, nx) in ((1, (2, 300, 410)), (2, (1, 37, 53))):
        U, bed, rough = case(seed, B, ny, nx, dtype, cuda_device)
        grid = _lib.make_grid(B, ny, nx, 1, 1, 0.05, 0.07)
        ref, dt_ref = step_with_workspace(cfg_c, grid, U, bed, rough, lvl, cuda_device)
        out = torch.full_like(U, float("nan"))
        dt = torch.full((1,), float("nan"), dtype=dtype, device=cuda_device)
        print(f"noalloc step: dtype bytes {word}, B={B}, {ny}x{nx}", flush=True)
        _lib.check(lib.swe_step_noalloc(ctypes.byref(grid), ctypes.byref(cfg_c), word, _lib.ptr(U), _lib.ptr(bed),
                                        _lib.ptr(rough), _lib.ptr(lvl), 0, 1e30, _lib.ptr(out), _lib.ptr(dt), 0,
                                        _lib.stream_ptr(cuda_device)), "step_noalloc")
        torch.cuda.synchronize()
        assert float(dt) == dt_ref
        assert torch.equal(out, ref), f"state differs: max {float((out - ref).abs().max()):.3e}"


def test_noalloc_rejects_grids_beyond_its_scratch(cuda_device):
    lib = _lib.load()
    cfg_c = _lib.make_config(swe.SWEConfig())
    dummy = torch.zeros(1, dtype=torch.float32, device=cuda_device)
    for word, cells_max in ((4, 5_592_128), (8, 2_796_064)):
        nx = 4096
        ny = cells_max // nx + 1
        grid = _lib.make_grid(1, ny, nx, 1, 1, 1.0, 1.0)
        rc = lib.swe_step_noalloc(ctypes.byref(grid), ctypes.byref(cfg_c), word, _lib.ptr(dummy), _lib.ptr(dummy),
                                  _lib.ptr(dummy), _lib.ptr(dummy), 0, 1e30, _lib.ptr(dummy), _lib.ptr(dummy), 0,
                                  _lib.stream_ptr(cuda_device))
        assert rc == -7
        assert b"fixed scratch holds" in lib.swe_last_error_message()
    # a non-positive dt leaves U_out untouched and is reported through dt_out
    U, bed, rough = case(3, 1, 20, 24, torch.float32, cuda_device)
    grid = _lib.make_grid(1, 20, 24, 1, 1, 0.1, 0.1)
    out = torch.full_like(U, 7.0)
    dt = torch.full((1,), float("nan"), dtype=torch.float32, device=cuda_device)
    _lib.check(lib.swe_step_noalloc(ctypes.byref(grid), ctypes.byref(cfg_c), 4, _lib.ptr(U), _lib.ptr(bed),
                                    _lib.ptr(rough), _lib.ptr(dummy), 0, 0.0, _lib.ptr(out), _lib.ptr(dt), 0,
                                    _lib.stream_ptr(cuda_device)), "step_noalloc")
    torch.cuda.synchronize()
    assert float(dt) == 0.0
    assert bool((out == 7.0).all())
