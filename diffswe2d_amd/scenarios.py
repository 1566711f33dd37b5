"""Seeded synthetic workloads for the shallow-water time step.

Each builder returns a plain dict of float64 numpy arrays plus the scalar
settings of the run, so that the HIP path and the oracle start from identical
inputs.  Grids use cell-centre coordinates x_i = (i + 0.5) * dx, the model-grid
convention of the reference (src/io_netcdf.py:63-64).

Workloads
---------
* ``dam_break``      circular dam break, 256 x 256, dx = 0.1 m, walls, n = 0.
* ``coastal_runup``  1:30 beach with 12 Gaussian bumps and a dry crest,
                     544 x 344 at dx = 0.01 m, Manning n = 0.012, west water
                     level from a sum of 3 Gaussian pulses, 6 wet gauges.
* ``tsunami_basin``  8192 x 8192 at dx = 100 m, fractal bathymetry
                     (spectral slope -2) mapped to 10..4000 m depth, ~15 % land
                     on the east coast, 2 m Gaussian hump (sigma 20 km),
                     n = 0.025, transmissive boundaries.
"""
from __future__ import annotations


import numpy as np


def dam_break(n: int = 256, dx: float = 0.1, radius: float = 5.0, h_in: float = 2.0,
              h_out: float = 0.5):
    """Circular dam break on a flat bed centred in the square domain."""
    x = (np.arange(n) + 0.5) * dx
    X, Y = np.meshgrid(x, x)
    c = 0.5 * n * dx
    h = np.where((X - c) ** 2 + (Y - c) ** 2 <= radius * radius, h_in, h_out)
    U = np.stack([h, np.zeros_like(h), np.zeros_like(h)])[None]
    return dict(U=U, bed=np.zeros((1, 1, n, n)), rough=np.zeros((1, 1, n, n)), dx=dx, dy=dx,
                boundaries=("wall", "wall", "wall", "wall"), steps=500)


def gaussian_pulses(rng, count=3, amp_max=0.02, width=(0.5, 2.0), centre=(2.0, 12.0)):
    """(amplitude, centre time, width) triples of a seeded pulse train."""
    amp = rng.uniform(0.25 * amp_max, amp_max, count)
    t_c = rng.uniform(centre[0], centre[1], count)
    w = rng.uniform(width[0], width[1], count)
    return np.stack([amp, t_c, w], axis=1)


def pulse_level(pulses, t):
    """Water level sum_k A_k exp(-((t - t_k) / w_k)^2) at times t."""
    t = np.asarray(t, dtype=np.float64)[..., None]
    return np.sum(pulses[:, 0] * np.exp(-((t - pulses[:, 1]) / pulses[:, 2]) ** 2), axis=-1)


def runup_bed(nx, ny, dx, rng, crest=0.03, bumps=12):
    """1:30 beach from 0.13 m depth in the west up to a flat dry crest, plus bumps."""
    x = (np.arange(nx) + 0.5) * dx
    y = (np.arange(ny) + 0.5) * dx
    X, Y = np.meshgrid(x, y)
    z = np.minimum(-0.13 + X / 30.0, crest)
    Lx, Ly = nx * dx, ny * dx
    for _ in range(bumps):
        height = rng.uniform(0.0, 0.05)
        sigma = rng.uniform(0.05, 0.3)
        cx, cy = rng.uniform(0.0, Lx), rng.uniform(0.0, Ly)
        z = z + height * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / (2.0 * sigma * sigma))
    return z


def coastal_runup(seed: int = 2, nx: int = 544, ny: int = 344, dx: float = 0.01,
                  duration: float = 22.5, sample_dt: float = 0.05, n_gauges: int = 6):
    """Coastal run-up forward problem; the domain shape is kept when nx/ny shrink.

    Reduced variants (e.g. nx=136, ny=86, dx=0.04) cover the same 5.44 x 3.44 m
    basin with the same bathymetry recipe and seed.
    """
    rng = np.random.default_rng(seed)
    z = runup_bed(nx, ny, dx, rng)
    pulses = gaussian_pulses(rng)
    h = np.maximum(-z, 0.0)
    wet = np.argwhere(h > 0.01)
    pick = rng.choice(len(wet), size=n_gauges, replace=False)
    gauges = [tuple(int(v) for v in wet[i]) for i in sorted(pick)]
    U = np.stack([h, np.zeros_like(h), np.zeros_like(h)])[None]
    n_chunks = int(round(duration / sample_dt))
    levels = pulse_level(pulses, np.arange(n_chunks) * sample_dt)
    return dict(U=U, bed=z[None, None], rough=np.full((1, 1, ny, nx), 0.012), dx=dx, dy=dx,
                boundaries=("water_level", "wall", "wall", "wall"), pulses=pulses,
                chunk_levels=levels, sample_dt=sample_dt, duration=duration, gauges=gauges)


def fractal_field(n: int, rng, slope: float = -2.0):
    """Real field with isotropic power spectrum ~ k^slope, normalised to [0, 1]."""
    noise = rng.standard_normal((n, n)).astype(np.float32)
    spec = np.fft.rfft2(noise)
    ky = np.fft.fftfreq(n)[:, None]
    kx = np.fft.rfftfreq(n)[None, :]
    k = np.sqrt(kx * kx + ky * ky)
    k[0, 0] = 1.0
    spec *= (k ** (0.5 * slope)).astype(np.complex64)
    spec[0, 0] = 0.0
    f = np.fft.irfft2(spec, s=(n, n)).astype(np.float64)
    f -= f.min()
    f /= max(f.max(), 1e-30)
    return f


def tsunami_basin(seed: int = 3, n: int = 8192, dx: float = 100.0, land_fraction: float = 0.15,
                  hump: float = 2.0, sigma: float = 20e3):
    """Large open basin; float32 arrays to keep the 8192^2 case at ~1.3 GB of host memory."""
    rng = np.random.default_rng(seed)
    f = fractal_field(n, rng)
    depth = 10.0 + (4000.0 - 10.0) * f
    z = -depth
    # noisy east coastline: land where column index exceeds a wandering edge
    base = int(round((1.0 - land_fraction) * n))
    wander = np.cumsum(rng.standard_normal(n)) * (n / 2048.0)
    wander -= wander.mean()
    edge = np.clip(base + wander.astype(np.int64), n // 2, n - 4)
    cols = np.arange(n)[None, :]
    land = cols >= edge[:, None]
    z = np.where(land, 1.0 + 49.0 * f, z)
    x = (np.arange(n) + 0.5) * dx
    cx = rng.uniform(0.2, 0.6) * n * dx
    cy = rng.uniform(0.3, 0.7) * n * dx
    gx = np.exp(-((x - cx) ** 2) / (2 * sigma * sigma))
    gy = np.exp(-((x - cy) ** 2) / (2 * sigma * sigma))
    eta = hump * gy[:, None] * gx[None, :]
    h = np.where(land, 0.0, np.maximum(eta - z, 0.0)).astype(np.float32)
    U = np.zeros((1, 3, n, n), dtype=np.float32)
    U[0, 0] = h
    return dict(U=U, bed=z.astype(np.float32)[None, None],
                rough=np.full((1, 1, n, n), 0.025, dtype=np.float32), dx=dx, dy=dx,
                boundaries=("transmissive",) * 4, steps=2000, land=float(land.mean()))


# ---------------------------------------------------------------------------
# boundary-control inversion and ensemble workloads
# ---------------------------------------------------------------------------
RUNUP_GRIDS = ((172, 272, 0.02), (344, 544, 0.01), (688, 1088, 0.005))


def physical_gauges(gauges, dx_from, dx_to):
    """Map gauge cells between resolutions of the same basin by their centre position."""
    out = []
    for iy, ix in gauges:
        y, x = (iy + 0.5) * dx_from, (ix + 0.5) * dx_from
        out.append((int(y // dx_to), int(x // dx_to)))
    return out


def spline_times(n_chunks=450, sample_dt=0.05):
    return np.arange(n_chunks) * sample_dt


def greville_points(t0, t1, n_controls, degree=3):
    """Greville abscissae of the clamped B-spline basis on [t0, t1]."""
    n_inner = n_controls - degree - 1
    inner = np.linspace(t0, t1, n_inner + 2)[1:-1] if n_inner > 0 else np.empty(0)
    knots = np.concatenate([[t0] * (degree + 1), inner, [t1] * (degree + 1)])
    return np.array([knots[i + 1:i + degree + 1].mean() for i in range(n_controls)])


def random_waveform_controls(rng, n_controls, n_chunks=450, sample_dt=0.05, pulses=3):
    """Seeded spline controls: a random Gaussian-pulse train sampled at the Greville points."""
    t_end = (n_chunks - 1) * sample_dt
    p = gaussian_pulses(rng, count=pulses, centre=(0.1 * t_end, 0.6 * t_end))
    return pulse_level(p, greville_points(0.0, t_end, n_controls)), p


def inversion_problem(seed=4, nx=544, ny=344, dx=0.01, n_controls=64, n_chunks=450, sample_dt=0.05):
    """Boundary-control inversion on the coastal run-up domain: truth waveform and setup."""
    base = coastal_runup(nx=nx, ny=ny, dx=dx)
    # the gauges are physical points (the 544 x 344 grid's picks), the same
    # points the ensemble scenarios sample at every resolution
    base["gauges"] = physical_gauges(coastal_runup()["gauges"], 0.01, dx)
    rng = np.random.default_rng(seed)
    t_end = (n_chunks - 1) * sample_dt
    truth_pulses = gaussian_pulses(rng, centre=(0.1 * t_end, 0.6 * t_end))
    times = spline_times(n_chunks, sample_dt)
    return dict(base=base, truth_pulses=truth_pulses, truth_levels=pulse_level(truth_pulses, times),
                times=times, n_controls=n_controls, n_chunks=n_chunks, sample_dt=sample_dt)


def ensemble_scenarios(n=256, seed=5, n_controls=64, n_chunks=450, sample_dt=0.05, grids=RUNUP_GRIDS):
    """Seeded scenarios cycling through the run-up grids, each with its own spline waveform.

    Returns (scenarios, controls [n, n_controls]); the basis that maps controls
    to the per-chunk levels is ``inversion.clamped_bspline_basis(times, n_controls)``.
    """
    import torch

    from .inversion import clamped_bspline_basis

    rng = np.random.default_rng(seed)
    ref = coastal_runup()  # 544 x 344 gauges define the physical gauge points
    times = torch.tensor(spline_times(n_chunks, sample_dt))
    basis = clamped_bspline_basis(times, n_controls).numpy()
    bases = {}
    out, ctrls = [], []
    for i in range(n):
        ny, nx, dx = grids[i % len(grids)]
        if (ny, nx) not in bases:
            b = coastal_runup(nx=nx, ny=ny, dx=dx)
            bases[(ny, nx)] = (b, physical_gauges(ref["gauges"], 0.01, dx))
        b, g = bases[(ny, nx)]
        c, _ = random_waveform_controls(rng, n_controls, n_chunks, sample_dt)
        ctrls.append(c)
        out.append(dict(bed=b["bed"][0, 0], rough=b["rough"][0, 0], dx=dx, dy=dx, U0=b["U"][0],
                        levels=basis @ c, chunk=sample_dt, gauges=g))
    return out, np.stack(ctrls)
