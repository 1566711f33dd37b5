"""BF16 MFMA surrogate kernels against CPU torch (FP32 accumulation over bf16-rounded operands)."""
import pytest
import torch
import torch.nn.functional as F

from diffswe2d_amd import surrogate as S

pytestmark = pytest.mark.gpu


def bf(x):
    return x.to(torch.bfloat16).float()


@pytest.mark.parametrize("M,N,K,splits", [(64, 64, 16, 1), (100, 70, 48, 1), (6, 900, 160, 1), (160, 800, 28800, 16),
                                          (33, 17, 800, 4)])
def test_gemm_nt_matches_fp32_matmul(cuda_device, M, N, K, splits):
    g = torch.Generator().manual_seed(M * 7 + N)
    A = torch.randn(M, K, generator=g)
    B = torch.randn(N, K, generator=g)
    Kn = S._Kernels(cuda_device)
    C = Kn.gemm(M, N, K, A.to(torch.bfloat16).to(cuda_device), K, B.to(torch.bfloat16).to(cuda_device), K,
                splits=splits)
    ref = bf(A) @ bf(B).T
    err = float((C.cpu() - ref).abs().max() / ref.abs().max())
    print(f"gemm {M}x{N}x{K} splits {splits}: max rel err {err:.2e}")
    assert err < 1e-5


def torch_net(specs, params, x):
    """CPU reference: each layer on bf16-rounded input and weights, FP32 math."""
    h = x[:, None, :] if x.dim() == 2 else x
    for i, sp in enumerate(specs):
        W, b = params[2 * i], params[2 * i + 1]
        y = F.conv1d(bf(h), bf(W), b, padding=sp.dil * (sp.kw - 1) // 2, dilation=sp.dil)
        if sp.relu:
            y = torch.relu(y)
        h = y + h if sp.residual else y
    return h


@pytest.mark.parametrize("n_cond", [0, 1])
def test_network_forward_and_gradients(cuda_device, n_cond):
    specs = S.default_specs(channels=32, dilations=(1, 4, 16), n_out=6, n_cond=n_cond)
    net = S.Surrogate(specs, device=cuda_device, seed=3)
    Bn, L = 8, 450
    g = torch.Generator().manual_seed(1)
    x = torch.randn(Bn, L, generator=g) * 0.02
    tgt = torch.randn(Bn, 6, L, generator=g) * 0.02
    cond = torch.arange(Bn, dtype=torch.float32)[:, None] * 0.5 - 1.0 if n_cond else None
    xg = x.to(cuda_device).requires_grad_(True)
    y = net(xg, None if cond is None else cond.to(cuda_device))
    loss = ((y - tgt.to(cuda_device)) / 0.02).square().mean()
    loss.backward()
    params = [p.detach().cpu().clone().requires_grad_(True) for p in net.weights]
    xr = (x / 0.02).clone().requires_grad_(True)
    xin = xr if cond is None else torch.cat((xr[:, None, :], cond[:, :, None].expand(-1, -1, L)), 1)
    yr = torch_net(specs, params, xin) * 0.02
    lr = ((yr - tgt) / 0.02).square().mean()
    lr.backward()
    ef = float((y.detach().cpu() - yr.detach()).abs().max() / yr.detach().abs().max())
    print(f"forward max rel err {ef:.2e}; loss {float(loss.detach()):.6f} vs {float(lr.detach()):.6f}")
    # bf16 activations: a last-bit FP32 difference before a bf16 rounding flips that
    # activation by one bf16 ulp (0.4 %), so outputs agree to ~1e-3, not FP32 precision
    assert ef < 5e-3
    for i, (p, q) in enumerate(zip(net.weights, params)):
        e = float((p.grad.cpu() - q.grad).norm() / max(float(q.grad.norm()), 1e-30))
        print(f"param {i}: grad rel err {e:.2e}")
        assert e < 2e-2  # the MFMA path rounds the backward operands to bf16 as well
    ex = float((xg.grad.cpu() * 0.02 - xr.grad).norm() / xr.grad.norm())
    print(f"input grad rel err {ex:.2e}")
    assert ex < 2e-2


def test_training_reduces_loss(cuda_device):
    specs = S.default_specs(channels=32, dilations=(1, 2, 4, 8), n_out=2)
    net = S.Surrogate(specs, device=cuda_device, seed=0)
    g = torch.Generator().manual_seed(0)
    n, L = 64, 128
    x = torch.randn(n, L, generator=g).cumsum(1) * 0.002
    y = torch.stack([torch.roll(x, 5, 1), 0.5 * torch.roll(x, 11, 1)], 1)
    hist = S.train(net, x.to(cuda_device), y.to(cuda_device), epochs=40, batch=16, lr=3e-3)
    print(f"loss {hist[0][2]:.4e} -> {hist[-1][2]:.4e}")
    assert hist[-1][2] < 0.2 * hist[0][2]


def test_condition_shape_is_checked(cuda_device):
    net = S.Surrogate(S.default_specs(channels=16, dilations=(1,), n_cond=1), device=cuda_device)
    with pytest.raises(ValueError):
        net(torch.zeros(2, 32, device=cuda_device))
    plain = S.Surrogate(S.default_specs(channels=16, dilations=(1,)), device=cuda_device)
    with pytest.raises(ValueError):
        plain(torch.zeros(2, 32, device=cuda_device), torch.zeros(2, 1, device=cuda_device))


def test_parameter_count_default():
    net_specs = S.default_specs()
    n = sum(sp.cout * sp.cin * sp.kw + sp.cout for sp in net_specs)
    assert 0.9e6 < n < 1.2e6
