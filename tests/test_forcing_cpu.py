"""RainfallForcing: the reference's constructor checks and time sampling (src/forcing.py:13-38)."""
import pytest
import torch

import diffswe2d_amd as swe


def test_rainfall_forcing_checks():
    with pytest.raises(ValueError):
        swe.RainfallForcing(torch.tensor([0.0, 1.0]), torch.zeros(3, 1, 2, 2))
    with pytest.raises(ValueError):
        swe.RainfallForcing(torch.tensor([1.0, 0.0]), torch.zeros(2, 1, 2, 2))
    rf = swe.RainfallForcing(torch.tensor([0.0, 2.0]), torch.stack([torch.zeros(1, 2, 2), torch.ones(1, 2, 2)]))
    assert float(rf.at(1.0)[0, 0, 0, 0]) == 0.5
    assert float(rf.at(-1.0)[0, 0, 0, 0]) == 0.0 and float(rf.at(5.0)[0, 0, 0, 0]) == 1.0
