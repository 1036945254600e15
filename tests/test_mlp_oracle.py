"""Pins of the FP64 MLP oracle (oracle/mlp.c) used to check the HIP
learned-tail kernels: torch (CPU, float64) nn.Linear + SiLU forward and
autograd gradients as an independent implementation, central finite
differences of the loss, and closed forms of silu."""
import math

import numpy as np
import pytest


def _rand_net(rng, dims):
    import oracle as orc

    n = orc.mlp_nparam(dims)
    p = np.empty(n)
    o = 0
    for l in range(len(dims) - 1):
        din, dout = dims[l], dims[l + 1]
        p[o:o + dout * din] = rng.standard_normal(dout * din) / math.sqrt(din)
        o += dout * din
        p[o:o + dout] = 0.1 * rng.standard_normal(dout)
        o += dout
    assert o == n
    return p


def _torch_net(dims, p):
    import torch

    layers, o = [], 0
    for l in range(len(dims) - 1):
        din, dout = dims[l], dims[l + 1]
        lin = torch.nn.Linear(din, dout).double()
        with torch.no_grad():
            lin.weight.copy_(torch.from_numpy(p[o:o + dout * din].reshape(dout, din)))
            o += dout * din
            lin.bias.copy_(torch.from_numpy(p[o:o + dout]))
            o += dout
        layers.append(lin)
        if l < len(dims) - 2:
            layers.append(torch.nn.SiLU())
    return torch.nn.Sequential(*layers)


def test_silu_closed_forms(orc):
    assert orc.silu(0.0) == 0.0
    assert orc.dsilu(0.0) == pytest.approx(0.5, abs=1e-15)
    for z in (-3.0, -0.5, 0.7, 2.5):
        s = 1.0 / (1.0 + math.exp(-z))
        assert orc.silu(z) == pytest.approx(z * s, rel=1e-14)
        h = 1e-6
        fd = (orc.silu(z + h) - orc.silu(z - h)) / (2 * h)
        assert orc.dsilu(z) == pytest.approx(fd, rel=1e-8)


@pytest.mark.parametrize("dims", [[5, 3], [21, 16, 16, 20], [7, 33, 9, 4]])
def test_mlp_forward_and_grad_vs_torch(orc, dims):
    import torch

    rng = np.random.default_rng(len(dims) * 100 + dims[0])
    p = _rand_net(rng, dims)
    B = 13
    x = rng.standard_normal((B, dims[0]))
    t = rng.standard_normal((B, dims[-1]))
    net = _torch_net(dims, p)
    xt = torch.from_numpy(x)
    yt = net(xt)
    y = orc.mlp_forward(dims, p, x)
    assert np.allclose(y, yt.detach().numpy(), rtol=1e-12, atol=1e-12)
    loss_t = ((yt - torch.from_numpy(t)) ** 2).mean()
    loss_t.backward()
    gt = np.concatenate([torch.cat([m.weight.grad.reshape(-1), m.bias.grad]).numpy()
                         for m in net if isinstance(m, torch.nn.Linear)])
    loss, g = orc.mlp_loss_grad(dims, p, x, t)
    assert loss == pytest.approx(loss_t.item(), rel=1e-12)
    assert np.allclose(g, gt, rtol=1e-10, atol=1e-13)


def test_mlp_grad_finite_differences(orc):
    dims = [6, 8, 5, 3]
    rng = np.random.default_rng(5)
    p = _rand_net(rng, dims)
    x = rng.standard_normal((9, dims[0]))
    t = rng.standard_normal((9, dims[-1]))
    _, g = orc.mlp_loss_grad(dims, p, x, t)
    h = 1e-6
    for i in rng.choice(p.size, 25, replace=False):
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (orc.mlp_loss_grad(dims, pp, x, t, False)[0] - orc.mlp_loss_grad(dims, pm, x, t, False)[0]) / (2 * h)
        assert g[i] == pytest.approx(fd, rel=1e-6, abs=1e-10)
