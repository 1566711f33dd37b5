"""Rainfall record on the model grid -- the reference's ``RainfallForcing``
(src/forcing.py:5-38), same constructor checks and the same ``at`` sampling.

``SWE2D(..., rainfall=RainfallForcing(times, rates))`` hands the record to the
step kernels, which sample it on the device at each stage's model time (the
same searchsorted / linear interpolation / end-hold rule as ``at``)."""
from __future__ import annotations

import numpy as np
import torch
import torch.nn as nn


class RainfallForcing(nn.Module):
    """times [nt] in seconds, rates [nt, 1, ny, nx] in m/s; linear in time, held at both ends."""

    def __init__(self, times: torch.Tensor, rates: torch.Tensor):
        super().__init__()
        if times.ndim != 1 or rates.ndim != 4 or rates.shape[0] != times.numel():
            raise ValueError("Expected times=[nt] and rates=[nt,1,ny,nx]")
        if times.numel() < 1:
            raise ValueError("Rainfall forcing must contain at least one time")
        if times.numel() > 1 and not torch.all(times[1:] > times[:-1]):
            raise ValueError("Rainfall times must be strictly increasing")
        self.register_buffer("times", times)
        self.register_buffer("rates", rates)

    def at(self, time) -> torch.Tensor:
        """Rainfall [1, 1, ny, nx] at model time `time` (seconds)."""
        if self.times.numel() == 1:
            return self.rates[0:1]
        t = torch.as_tensor(time, dtype=self.times.dtype, device=self.times.device)
        right = torch.searchsorted(self.times, t).clamp(1, self.times.numel() - 1)
        left = right - 1
        width = torch.clamp(self.times[right] - self.times[left], min=1.0e-12)
        weight = (t - self.times[left]) / width
        value = (1.0 - weight) * self.rates[left] + weight * self.rates[right]
        value = torch.where(t <= self.times[0], self.rates[0], value)
        value = torch.where(t >= self.times[-1], self.rates[-1], value)
        return value.unsqueeze(0)


def load_rainfall_txt(filepath):
    """Text rainfall record (column 1: time in s, column 2: rain in mm/h) -> (times, rates in m/s),
    as the reference's loader (src/io_ascii.py:1543-1550)."""
    data = np.loadtxt(filepath)
    return data[:, 0], data[:, 1] / (1000.0 * 3600.0)
