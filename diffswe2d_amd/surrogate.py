"""1-D convolution surrogate: boundary-level series -> gauge time series.

No reference counterpart: this learned part warm-starts the boundary-control
inversion.  The network maps the dense boundary series (basis @ 64 spline
controls, one value per 0.05 s output chunk) to the 6 gauge series of the same
length.  It is a stack of 'same' 1-D convolutions (kernel 7, then 8 dilated
kernel-5 layers with residual connections, then a 1x1 output layer; 160
channels, ~1.03 M parameters), whose receptive field (about 520 samples)
covers the whole 450-sample record.

Every convolution runs on the project's BF16 MFMA kernels (csrc/surrogate.hip):
weights and activations are rounded to bf16, products accumulate in FP32.
The backward pass (input, weight and bias gradients) uses the same kernels.
Master weights and the Adam update stay in FP32 (torch.optim).
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass

import torch

from . import _lib


def _ceil16(v):
    return (v + 15) // 16 * 16


@dataclass
class ConvSpec:
    cin: int
    cout: int
    kw: int
    dil: int
    relu: bool
    residual: bool


def _chk(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed: {_lib.load().sur_last_error_message().decode()}")


class _Kernels:
    """Thin wrappers over the surrogate C ABI on torch tensors."""

    def __init__(self, device):
        self.lib = _lib.load()
        self.device = device

    def s(self):
        return _lib.stream_ptr(self.device)

    def gemm(self, M, N, K, A, lda, B, ldb, splits=1):
        C = torch.empty(M, N, dtype=torch.float32, device=self.device)
        part = torch.empty(splits, M, N, dtype=torch.float32, device=self.device) if splits > 1 else None
        _chk(self.lib.sur_gemm_nt_bf16(M, N, K, _lib.ptr(A), lda, _lib.ptr(B), ldb, _lib.ptr(C), N, splits,
                                       _lib.ptr(part), self.s()), "gemm")
        return C

    def im2col(self, X, cin, Bn, L, kw, dil, kpad, transpose):
        shape = (kpad, Bn * L) if transpose else (Bn * L, kpad)
        col = torch.empty(shape, dtype=torch.bfloat16, device=self.device)
        _chk(self.lib.sur_im2col(_lib.ptr(X), cin, Bn, L, kw, dil, kpad, int(transpose), _lib.ptr(col), self.s()),
             "im2col")
        return col

    def col2im(self, dcol, cin, Bn, L, kw, dil):
        dX = torch.empty(cin, Bn * L, dtype=torch.float32, device=self.device)
        _chk(self.lib.sur_col2im(_lib.ptr(dcol), cin, Bn, L, kw, dil, _lib.ptr(dX), self.s()), "col2im")
        return dX

    def bias_act(self, C, bias, relu):
        M, N = C.shape
        pre = torch.empty_like(C)
        out = torch.empty_like(C)
        _chk(self.lib.sur_bias_act(_lib.ptr(C), _lib.ptr(bias), M, N, int(relu), _lib.ptr(pre), None, _lib.ptr(out),
                                   self.s()), "bias_act")
        return pre, out

    def act_bwd(self, g, pre, relu, ld_t):
        M, N = g.shape
        g_pre = torch.empty_like(g)
        g_bf = torch.empty(M, N, dtype=torch.bfloat16, device=self.device)
        g_t = torch.zeros(N, ld_t, dtype=torch.bfloat16, device=self.device)
        _chk(self.lib.sur_act_bwd(_lib.ptr(g), _lib.ptr(pre), M, N, int(relu), _lib.ptr(g_pre), _lib.ptr(g_bf),
                                  _lib.ptr(g_t), ld_t, self.s()), "act_bwd")
        return g_pre, g_bf, g_t

    def bias_grad(self, g):
        M, N = g.shape
        db = torch.empty(M, dtype=torch.float32, device=self.device)
        _chk(self.lib.sur_bias_grad(_lib.ptr(g), M, N, _lib.ptr(db), self.s()), "bias_grad")
        return db

    def cast(self, x):
        y = torch.empty(x.shape, dtype=torch.bfloat16, device=self.device)
        _chk(self.lib.sur_cast_bf16(_lib.ptr(x), x.numel(), _lib.ptr(y), self.s()), "cast")
        return y

    def cast_pad_rows(self, x, cpad):
        R, Cc = x.shape
        y = torch.empty(R, cpad, dtype=torch.bfloat16, device=self.device)
        _chk(self.lib.sur_cast_pad_rows(_lib.ptr(x), R, Cc, cpad, _lib.ptr(y), self.s()), "cast_pad_rows")
        return y

    def transpose(self, x, rpad):
        R, Cc = x.shape
        y = torch.empty(Cc, rpad, dtype=torch.bfloat16, device=self.device)
        _chk(self.lib.sur_transpose_bf16(_lib.ptr(x), R, Cc, rpad, _lib.ptr(y), self.s()), "transpose")
        return y


def default_specs(channels=160, dilations=(1, 2, 4, 8, 16, 32, 64, 1), n_out=6, n_cond=0):
    """Layer list; the first layer reads the series plus n_cond constant condition channels."""
    specs = [ConvSpec(1 + n_cond, channels, 7, 1, True, False)]
    specs += [ConvSpec(channels, channels, 5, d, True, True) for d in dilations]
    specs.append(ConvSpec(channels, n_out, 1, 1, False, False))
    return specs


def _conv_forward(K, spec, x, W, b, Bn, L):
    """x fp32 [cin][B*L] -> (y fp32 [cout][B*L], saved) with bf16 operands."""
    kpad = _ceil16(spec.cin * spec.kw)
    xb = K.cast(x.contiguous())
    col = K.im2col(xb, spec.cin, Bn, L, spec.kw, spec.dil, kpad, False)
    Wp = K.cast_pad_rows(W.detach().reshape(spec.cout, spec.cin * spec.kw).contiguous(), kpad)
    C = K.gemm(spec.cout, Bn * L, kpad, Wp, kpad, col, kpad)
    pre, y = K.bias_act(C, b.detach().contiguous(), spec.relu)
    if spec.residual:
        y = y + x
    return y, (xb, pre)


def _conv_backward(K, spec, saved, W, g_y, Bn, L, need_dx, splits, need_w=True):
    xb, pre = saved
    BL = Bn * L
    cpad = _ceil16(spec.cout)
    g_pre, g_bf, g_t = K.act_bwd(g_y.contiguous(), pre, spec.relu, cpad)
    kpad = _ceil16(spec.cin * spec.kw)
    nk = spec.cin * spec.kw
    dW = db = None
    if need_w:
        if BL % 16:
            raise ValueError(f"weight gradients need batch * length divisible by 16, got {BL}")
        colT = K.im2col(xb, spec.cin, Bn, L, spec.kw, spec.dil, kpad, True)
        dW = K.gemm(spec.cout, nk, BL, g_bf, BL, colT, BL, splits=splits).reshape(spec.cout, spec.cin, spec.kw)
        db = K.bias_grad(g_pre)
    dx = None
    if need_dx:
        Wt = K.transpose(W.detach().reshape(spec.cout, nk).contiguous(), cpad)
        dcol = K.gemm(nk, BL, cpad, Wt, cpad, g_t, cpad)
        dx = K.col2im(dcol, spec.cin, Bn, L, spec.kw, spec.dil)
        if spec.residual:
            dx = dx + g_y
    return dx, dW, db


class _NetFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, net, *params):
        Bn, cin, L = x.shape
        K = net.K
        h = x.detach().permute(1, 0, 2).reshape(cin, Bn * L).contiguous()
        saved = []
        for i, spec in enumerate(net.specs):
            h, s = _conv_forward(K, spec, h, params[2 * i], params[2 * i + 1], Bn, L)
            saved.append(s)
        ctx.net, ctx.saved, ctx.shape = net, saved, (Bn, cin, L)
        ctx.params = params
        n_out = net.specs[-1].cout
        return h.reshape(n_out, Bn, L).permute(1, 0, 2).contiguous()

    @staticmethod
    def backward(ctx, g):
        net, (Bn, cin, L) = ctx.net, ctx.shape
        n_out = net.specs[-1].cout
        gy = g.permute(1, 0, 2).reshape(n_out, Bn * L).contiguous().float()
        grads = [None] * (2 * len(net.specs))
        need_x = ctx.needs_input_grad[0]
        splits = max(1, min(32, (Bn * L) // 1024))
        for i in range(len(net.specs) - 1, -1, -1):
            spec = net.specs[i]
            need_dx = i > 0 or need_x
            need_w = ctx.needs_input_grad[2 + 2 * i] or ctx.needs_input_grad[3 + 2 * i]
            gy, dW, db = _conv_backward(net.K, spec, ctx.saved[i], ctx.params[2 * i], gy, Bn, L, need_dx, splits,
                                        need_w)
            grads[2 * i], grads[2 * i + 1] = dW, db
        gx = gy.reshape(cin, Bn, L).permute(1, 0, 2) if need_x else None
        return (gx, None, *grads)


class Surrogate(torch.nn.Module):
    """Boundary series [B, L] (+ per-sample conditions [B, n_cond]) -> gauge series [B, n_out, L].

    Conditions (e.g. the grid resolution of the run that produced a pair) enter
    as constant input channels, so one network serves every resolution of an
    ensemble."""

    def __init__(self, specs=None, device="cuda", seed=0, in_scale=0.02, out_scale=0.02):
        super().__init__()
        self.specs = specs or default_specs()
        self.K = _Kernels(torch.device(device))
        g = torch.Generator().manual_seed(seed)
        self.weights = torch.nn.ParameterList()
        for spec in self.specs:
            fan_in = spec.cin * spec.kw
            std = math.sqrt(2.0 / fan_in) * (0.5 if spec.residual else 1.0)
            W = torch.randn(spec.cout, spec.cin, spec.kw, generator=g) * std
            self.weights.append(torch.nn.Parameter(W.to(device)))
            self.weights.append(torch.nn.Parameter(torch.zeros(spec.cout, device=device)))
        self.in_scale, self.out_scale = in_scale, out_scale
        self.n_cond = self.specs[0].cin - 1

    def n_params(self):
        return sum(p.numel() for p in self.weights)

    def forward(self, series, cond=None):
        """series: boundary levels [B, L] (metres) -> gauge h + z [B, n_out, L] (metres)."""
        x = (series / self.in_scale).float()[:, None, :]
        if self.n_cond:
            if cond is None or cond.shape != (series.shape[0], self.n_cond):
                raise ValueError(f"this surrogate needs conditions of shape ({series.shape[0]}, {self.n_cond})")
            x = torch.cat((x, cond.float()[:, :, None].expand(-1, -1, series.shape[1])), 1)
        elif cond is not None:
            raise ValueError("this surrogate takes no conditions")
        y = _NetFn.apply(x.contiguous(), self, *self.weights)
        return y * self.out_scale


def train(net, series, targets, epochs=60, batch=64, lr=2e-3, log=None, time_budget=None, seed=0, cond=None):
    """Adam on FP32 master weights; the MFMA kernels compute every forward and backward."""
    opt = torch.optim.Adam(net.weights, lr=lr)
    n = series.shape[0]
    g = torch.Generator(device="cpu").manual_seed(seed)
    hist = []
    t0 = time.perf_counter()
    steps_per_epoch = max(1, n // batch)
    sched = torch.optim.lr_scheduler.CosineAnnealingLR(opt, T_max=epochs * steps_per_epoch)
    for ep in range(epochs):
        perm = torch.randperm(n, generator=g).to(series.device)
        tot = 0.0
        for s in range(steps_per_epoch):
            idx = perm[s * batch:(s + 1) * batch]
            opt.zero_grad()
            pred = net(series[idx], None if cond is None else cond[idx])
            loss = ((pred - targets[idx]) / net.out_scale).square().mean()
            loss.backward()
            opt.step()
            sched.step()
            tot += float(loss.detach())
        hist.append((ep, time.perf_counter() - t0, tot / steps_per_epoch))
        if log:
            log(f"surrogate epoch {ep}: loss {tot / steps_per_epoch:.4e}")
        if time_budget is not None and time.perf_counter() - t0 > time_budget:
            break
    return hist


if __name__ == "__main__":
    import argparse
    import json

    ap = argparse.ArgumentParser(description="Time surrogate training on synthetic pairs of the warm start's shapes "
                                             "(boundary series [L] + a condition channel -> 6 gauge series [L])")
    ap.add_argument("--pairs", type=int, default=256)
    ap.add_argument("--length", type=int, default=450)
    ap.add_argument("--epochs", type=int, default=3)
    ap.add_argument("--batch", type=int, default=32)
    a = ap.parse_args()
    dev = torch.device("cuda:0")
    g = torch.Generator().manual_seed(0)
    series = (0.01 * torch.randn(a.pairs, a.length, generator=g)).to(dev)
    targets = (0.01 * torch.randn(a.pairs, 6, a.length, generator=g)).to(dev)
    cond = torch.zeros(a.pairs, 1, device=dev)
    net = Surrogate(default_specs(n_cond=1), device=dev)
    train(net, series, targets, epochs=1, batch=a.batch, cond=cond)  # first-use kernel loads
    torch.cuda.synchronize()
    hist = train(net, series, targets, epochs=a.epochs, batch=a.batch, cond=cond)
    steps = a.epochs * max(1, a.pairs // a.batch)
    sec = hist[-1][1]
    # forward GEMM FLOPs of one sample, x3 for a training step (forward, input and weight gradients)
    flops = sum(2.0 * s.cout * s.cin * s.kw * a.length for s in net.specs) * a.batch * 3
    print(json.dumps({"params": net.n_params(), "batch": a.batch, "steps": steps, "ms_per_step": 1e3 * sec / steps,
                      "gemm_tflop_per_step": flops / 1e12, "tflops_effective": flops * steps / sec / 1e12}))
