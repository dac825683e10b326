"""The emulator network on the GPU (include/crime_mlp.h) against the FP64 oracle with the GPU's BF16
rounding points emulated (oracle/mlp_oracle.py).

Bars: outputs to 1e-6 of their rms in 90% of the rows (BF16 ties flipped by FP32 accumulation order move
the rest by up to 5e-2); loss to 1e-4 relative; every parameter block's gradient to 1e-2 of
its norm (the remaining differences are FP32 accumulation order and BF16 ties flipped by it); Adam
steps equal to the oracle's Adam (given the step's own gradient) within 1e-3 lr; split training runs give
the same parameters bit for bit.
"""
import numpy as np
import pytest

import mlp_oracle as mo
import oracle
from crime_amd.emulator import Mlp

pytestmark = pytest.mark.gpu


def data(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(n, 8)).astype(np.float32)
    t = (np.sin(x[:, 0]) + 0.5 * x[:, 1] * x[:, 2] - 0.3 * x[:, 7] ** 2).astype(np.float32)
    return x, t


def test_forward_vs_oracle():
    p = mo.init_params(1)
    x, _ = data(1000, 2)   # not a multiple of the 64-row tiles
    with Mlp(max_batch=512) as net:
        net.set_params(p)
        y = net.forward(x)
        assert np.array_equal(net.params(), p)
    ref, _, _ = mo.forward(p, x, bf16=True)
    err = np.abs(y - ref) / np.sqrt(np.mean(ref ** 2))
    q = np.percentile(err, [90, 99, 100])
    print(f"forward: err / rms at 90/99/100%: {q}")
    # most rows agree to FP32 accumulation order; a BF16 rounding tie decided differently by FP32 and
    # FP64 accumulation moves an activation by one BF16 step in about 1% of the rows
    assert q[0] < 1e-6 and q[1] < 5e-3 and q[2] < 5e-2


def test_forward_mean_of_networks():
    """crime_mlp_forward_mean (inputs sent once, networks of different max_batch) against the mean of each
    network's own forward: the same rows in the same kernels, so equal up to the FP32 rounding of the mean."""
    x, _ = data(3000, 5)   # several batches of the smaller max_batch, the last one partial
    nets = [Mlp(max_batch=1024), Mlp(max_batch=512), Mlp(max_batch=1024)]
    try:
        for k, net in enumerate(nets):
            net.set_params(mo.init_params(10 + k))
        each = np.mean([net.forward(x) for net in nets], axis=0)
        mean = Mlp.forward_mean(nets, x)
        assert np.allclose(mean, each, rtol=1e-6, atol=1e-6 * np.abs(each).max())
        assert np.array_equal(Mlp.forward_mean(nets[:1], x), nets[0].forward(x))
    finally:
        for net in nets:
            net.close()


def test_loss_and_gradient_vs_oracle():
    p = mo.init_params(3)
    x, t = data(1024, 4)
    with Mlp(max_batch=1024) as net:
        net.set_params(p)
        loss, g = net.loss_grad(x, t)
    rl, rg = mo.loss_grad(p, x, t, bf16=True)
    assert abs(loss - rl) < 1e-4 * rl
    lay, _ = mo.layout()
    for name, off, shape in lay:
        n = int(np.prod(shape))
        a, b = g[off:off + n], rg[off:off + n]
        err = np.linalg.norm(a - b) / np.linalg.norm(b)
        print(f"grad {name}: rel err {err:.2e}")
        assert err < 1e-2, name


def gathered(xall, tall, step, batch, seed):
    """The minibatch the GPU draws at a step: Philox-4x32-10 of (step, row, 0, 0), key (seed, 'MLP1')."""
    idx = []
    for b in range(batch):
        r = oracle.philox([step, b, 0, 0], seed, 0x4D4C5031)
        idx.append(((r[0] << 32) | r[1]) % len(xall))
    idx = np.array(idx)
    return xall[idx], tall[idx]


def test_adam_steps_vs_oracle():
    """Steps one Adam update at a time: the minibatch the device drew (regathered here with the same Philox
    draw) gives, through crime_mlp_loss_grad, the gradient the step used -- bit for bit, as both run the same
    deterministic kernels -- and the oracle's Adam with that gradient must reproduce the update."""
    p0 = mo.init_params(5)
    xall, tall = data(5000, 6)
    lr, total, seed, batch, steps = 1e-3, 50, 77, 128, 4
    m, v = np.zeros(p0.size), np.zeros(p0.size)
    with Mlp(max_batch=256) as net:
        net.set_params(p0)
        net.set_data(xall, tall)
        for s in range(steps):
            p_before = net.params().astype(np.float64)
            xb, tb = gathered(xall, tall, s, batch, seed)
            l, g = net.loss_grad(xb, tb)
            rl, rg = mo.loss_grad(p_before.astype(np.float32), xb, tb, bf16=True)
            assert abs(l - rl) < 1e-4 * rl
            assert np.linalg.norm(g - rg) < 1e-2 * np.linalg.norm(rg)
            loss = net.train(1, batch, lr, total, seed=seed)
            assert abs(loss[0] - l) < 1e-5 * l      # the step saw the same minibatch
            ref, m, v = mo.adam_step(p_before, m, v, g.astype(np.float64), s, lr, total)
            d = np.abs(net.params() - ref)
            print(f"Adam step {s}: max |dp| = {d.max():.2e}")
            assert d.max() < 1e-3 * lr + 1e-6 * np.abs(ref).max()


def test_training_converges_and_replays_deterministically():
    xall, tall = data(20000, 8)
    p0 = mo.init_params(9)
    with Mlp(max_batch=1024) as a, Mlp(max_batch=1024) as b:
        for net in (a, b):
            net.set_params(p0)
            net.set_data(xall, tall)
        la = a.train(600, 1024, 2e-3, 600, seed=3)
        lb = np.concatenate([b.train(250, 1024, 2e-3, 600, seed=3), b.train(350, 1024, 2e-3, 600, seed=3)])
        pa, pb = a.params(), b.params()
        xt, tt = data(4096, 10)
        yt = a.forward(xt)
    assert np.array_equal(pa, pb)
    assert np.allclose(la, lb, rtol=1e-5, atol=0)   # the loss is summed with atomics
    print(f"training: loss {la[:10].mean():.3e} -> {la[-50:].mean():.3e}; test mse {np.mean((yt - tt) ** 2):.3e}")
    assert la[-50:].mean() < 0.02 * la[:10].mean()
    assert np.mean((yt - tt) ** 2) < 0.02 * np.var(tt)


def test_eager_training_equals_graph_replay():
    """Steps enqueued one by one (crime_mlp_set_graph(h, 0)) run the same kernels as the replayed graph:
    bit-identical parameters."""
    xall, tall = data(8192, 12)
    p0 = mo.init_params(4)
    with Mlp(max_batch=1024) as a, Mlp(max_batch=1024) as b:
        for net in (a, b):
            net.set_params(p0)
            net.set_data(xall, tall)
        b.set_graph(False)
        a.train(40, 1024, 2e-3, 40, seed=5)
        b.train(40, 1024, 2e-3, 40, seed=5)
        assert np.array_equal(a.params(), b.params())
