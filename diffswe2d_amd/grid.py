"""Uniform square-cell model grid (src/io_netcdf.py:12-60, ``ModelGrid``).

x and y hold cell-centre coordinates; dx == dy.  ``from_bounds`` takes the
outer edges of the domain and one resolution, and refuses an extent that is
not a whole number of cells rather than change the cell size.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch


@dataclass(frozen=True)
class ModelGrid:
    x: torch.Tensor
    y: torch.Tensor
    dx: float
    dy: float

    @property
    def nx(self) -> int:
        return int(self.x.numel())

    @property
    def ny(self) -> int:
        return int(self.y.numel())

    @classmethod
    def from_bounds(cls, xmin, xmax, ymin, ymax, resolution, dtype=torch.float64, device=None):
        """Cells of width ``resolution`` between the outer edges (io_netcdf.py:46-60)."""
        if resolution <= 0:
            raise ValueError("resolution must be positive")
        counts = []
        for lo, hi in ((xmin, xmax), (ymin, ymax)):
            exact = (hi - lo) / resolution
            n = int(round(exact))
            if n < 2:
                raise ValueError("Use at least two cells in each direction")
            if not np.isclose(exact, n, rtol=1e-10, atol=1e-10):
                raise ValueError("Model bounds must be exactly divisible by resolution")
            counts.append(n)
        x = xmin + (torch.arange(counts[0], dtype=dtype, device=device) + 0.5) * resolution
        y = ymin + (torch.arange(counts[1], dtype=dtype, device=device) + 0.5) * resolution
        return cls(x=x, y=y, dx=float(resolution), dy=float(resolution))

    @classmethod
    def from_cell_centres(cls, x, y, dtype=torch.float64, device=None):
        """Grid from strictly increasing, uniform centre vectors with dx == dy (io_netcdf.py:30-44)."""
        x = torch.as_tensor(x, dtype=dtype, device=device)
        y = torch.as_tensor(y, dtype=dtype, device=device)
        if x.ndim != 1 or y.ndim != 1 or x.numel() < 2 or y.numel() < 2:
            raise ValueError("x and y must be 1-D cell-centre arrays with at least two entries")
        steps = []
        for name, c in (("x", x), ("y", y)):
            d = torch.diff(c)
            if not bool(torch.all(d > 0)):
                raise ValueError("Model-grid coordinates must be strictly increasing")
            h = float(d.median())
            if not bool(torch.allclose(d, torch.full_like(d, h), rtol=1e-5, atol=max(1e-10, h * 1e-8))):
                raise ValueError(f"The SWE solver requires a uniform model {name} grid")
            steps.append(h)
        dx, dy = steps
        if not np.isclose(dx, dy, rtol=1e-8, atol=max(1e-10, dx * 1e-10)):
            raise ValueError("The simplified model grid requires dx == dy")
        return cls(x=x, y=y, dx=dx, dy=dy)
