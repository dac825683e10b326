"""Pins of the emulator-network oracle (oracle/mlp_oracle.py), independent of it: torch float64 autograd
for the forward pass and the gradient, central finite differences, torch.optim.Adam with a LambdaLR
cosine schedule for the update, and hand-picked BF16 roundings."""
import math

import numpy as np
import pytest
import torch

import mlp_oracle as mo


def torch_model(p):
    d = {k: torch.tensor(v, dtype=torch.float64, requires_grad=True) for k, v in mo.unpack(p).items()}
    return d


def torch_loss(d, x, t):
    h = torch.nn.functional.silu(x @ d["W0"] + d["b0"])
    for k in range(1, mo.NMID + 1):
        h = torch.nn.functional.silu(h @ d[f"W{k}"] + d[f"b{k}"])
    y = h @ d["w5"] + d["b5"][0]
    return ((y - t) ** 2).mean(), y


def test_bf16_rounding():
    vals = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 2 ** -9, -2.5, 0.0, 3.0e38])
    got = mo.to_bf16(vals)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7 -> even (1); 1 + 3 2^-8 -> 1 + 2^-6 (even); 2^-9 rounds down
    assert list(got[:6]) == [1.0, 1.0, 1 + 2 ** -6, 1.0, -2.5, 0.0]
    assert math.isclose(got[6], 3.0e38, rel_tol=2 ** -8)


def test_forward_and_gradient_vs_torch():
    p = mo.init_params(3)
    rng = np.random.default_rng(4)
    x = rng.normal(size=(64, 8))
    t = rng.normal(size=64)
    loss, g = mo.loss_grad(p, x, t, bf16=False)
    d = torch_model(p)
    tl, ty = torch_loss(d, torch.tensor(x), torch.tensor(t))
    tl.backward()
    y, _, _ = mo.forward(p, x, bf16=False)
    assert np.allclose(y, ty.detach().numpy(), rtol=1e-12, atol=1e-12)
    assert math.isclose(loss, tl.item(), rel_tol=1e-12)
    tg = mo.pack({k: v.grad.numpy() for k, v in d.items()})
    assert np.abs(g - tg).max() < 1e-12 * np.abs(tg).max()


def test_gradient_finite_differences():
    p = mo.init_params(5).astype(np.float64)
    rng = np.random.default_rng(6)
    x = rng.normal(size=(64, 8))
    t = rng.normal(size=64)
    _, g = mo.loss_grad(p, x, t, bf16=False)
    lay, _ = mo.layout()
    for name, off, shape in lay:
        i = off + int(rng.integers(0, int(np.prod(shape))))
        h = 1e-5
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (mo.loss_grad(pp, x, t, bf16=False)[0] - mo.loss_grad(pm, x, t, bf16=False)[0]) / (2 * h)
        assert abs(fd - g[i]) < 1e-6 * max(1.0, abs(g[i])), (name, fd, g[i])


def test_adam_vs_torch():
    rng = np.random.default_rng(7)
    p0 = rng.normal(size=50)
    w = torch.tensor(p0.copy(), requires_grad=True)
    total, lr0 = 7, 0.01
    opt = torch.optim.Adam([w], lr=lr0, betas=(0.9, 0.999), eps=1e-8)
    sch = torch.optim.lr_scheduler.LambdaLR(opt, lambda s: 0.5 * (1 + math.cos(math.pi * min(s / total, 1.0))))
    p, m, v = p0.copy(), np.zeros(50), np.zeros(50)
    for step in range(5):
        g = rng.normal(size=50)
        w.grad = torch.tensor(g)
        opt.step()
        sch.step()
        p, m, v = mo.adam_step(p, m, v, g, step, lr0, total)
        assert np.allclose(p, w.detach().numpy(), rtol=0, atol=1e-12)


def test_fold_standardisation():
    """The emulator folds its input/output standardisation into the first and last layers: the folded
    network on raw inputs equals the original on standardised inputs, de-standardised."""
    from crime_amd.emulator import fold_standardisation

    p = mo.init_params(11)
    rng = np.random.default_rng(12)
    x = rng.normal(size=(100, 8)) * 3 + 5
    xm, xs = x.mean(axis=0), x.std(axis=0)
    ym, ys = 1.7, 0.3
    q = fold_standardisation(p, xm, xs, ym, ys)
    y0, _, _ = mo.forward(p, (x - xm) / xs, bf16=False)
    y1, _, _ = mo.forward(q, x, bf16=False)
    assert np.allclose(y1, y0 * ys + ym, rtol=1e-5, atol=1e-5)
