"""Spline basis and loss pieces of the inversion driver (CPU)."""
import numpy as np
import torch

from diffswe2d_amd import inversion


def test_bspline_basis_matches_reference_golden(golden):
    d = golden("gradient")
    for name, dt in (("f64", torch.float64), ("f32", torch.float32)):
        times = torch.arange(10, dtype=dt) * 0.2
        B = inversion.clamped_bspline_basis(times, 64)
        np.testing.assert_allclose(B.numpy(), d[f"basis_{name}"], rtol=0, atol=1e-6 if name == "f32" else 1e-14)


def test_bspline_partition_of_unity_and_interpolation():
    t = torch.linspace(0.0, 22.45, 450, dtype=torch.float64)
    B = inversion.clamped_bspline_basis(t, 64)
    assert torch.allclose(B.sum(1), torch.ones(450, dtype=torch.float64), atol=1e-12)
    assert torch.all(B >= -1e-15)
    # clamped ends: the first and last samples take the first and last control exactly
    assert abs(float(B[0, 0]) - 1.0) < 1e-15 and abs(float(B[-1, -1]) - 1.0) < 1e-15
    # cubic splines reproduce linear functions through Greville abscissae
    c = torch.linspace(-1, 3, 64, dtype=torch.float64)
    s = B @ c
    assert torch.all(s[1:] >= s[:-1] - 1e-12)


def test_misfit_is_mean_of_per_gauge_mse():
    obs = torch.tensor([[1.0, 2.0], [3.0, 5.0]])
    tgt = torch.zeros_like(obs)
    per = ((obs - tgt) ** 2).mean(0)
    m = inversion.GaugeProblem.misfit(obs, tgt)
    assert m.dtype == torch.float64  # accumulated in FP64
    assert float(m) == float(per.mean()) == (1.0 + 9.0) / 4 + (4.0 + 25.0) / 4
