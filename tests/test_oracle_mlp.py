"""Pins of the prefilter network's CPU restatement (oracle/pbw_mlp_ref.py).

The exact-FP64 mode (no BF16 rounding) is checked against torch's fp64 modules
and autograd, and the gradient against central finite differences of the loss;
the BF16 rounding against torch's bfloat16 conversion; Adam against
torch.optim.Adam.  A transposed weight, a dropped bias or ReLU mask, a wrong
sign or a missing pos_weight fails one of them.
"""
import numpy as np
import pytest
import torch

from oracle import pbw_mlp_ref as R


def rand_theta(rng, scale=0.2):
    th = rng.normal(0.0, scale, R.NP)
    th[R.OFF_B1:R.OFF_W2] = rng.normal(0.0, 0.1, R.H)
    th[R.OFF_B2:R.OFF_W3] = rng.normal(0.0, 0.1, R.H)
    th[R.OFF_B3] = 0.1
    return th


def torch_net(theta):
    t = torch.tensor(theta, dtype=torch.float64)
    net = torch.nn.Sequential(torch.nn.Linear(R.IN, R.H), torch.nn.ReLU(), torch.nn.Linear(R.H, R.H),
                              torch.nn.ReLU(), torch.nn.Linear(R.H, 1)).double()
    with torch.no_grad():
        # the layout stores W1[in][out], W2[in][out]; torch keeps weight[out][in]
        net[0].weight.copy_(t[R.OFF_W1:R.OFF_B1].reshape(R.IN, R.H).T)
        net[0].bias.copy_(t[R.OFF_B1:R.OFF_W2])
        net[2].weight.copy_(t[R.OFF_W2:R.OFF_B2].reshape(R.H, R.H).T)
        net[2].bias.copy_(t[R.OFF_B2:R.OFF_W3])
        net[4].weight.copy_(t[R.OFF_W3:R.OFF_B3].reshape(1, R.H))
        net[4].bias.copy_(t[R.OFF_B3:R.NP])
    return net


def torch_grad(net):
    g = np.zeros(R.NP)
    g[R.OFF_W1:R.OFF_B1] = net[0].weight.grad.T.numpy().ravel()
    g[R.OFF_B1:R.OFF_W2] = net[0].bias.grad.numpy()
    g[R.OFF_W2:R.OFF_B2] = net[2].weight.grad.T.numpy().ravel()
    g[R.OFF_B2:R.OFF_W3] = net[2].bias.grad.numpy()
    g[R.OFF_W3:R.OFF_B3] = net[4].weight.grad.numpy().ravel()
    g[R.OFF_B3] = net[4].bias.grad.numpy()[0]
    return g


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(1)
    a = np.concatenate([rng.normal(0, 1, 5000), rng.normal(0, 1e-30, 100), [0.0, -0.0, 1.0, 1.00390625,
                                                                           1.01171875, 3.0e38]]).astype(np.float32)
    ref = torch.tensor(a).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(R.bf16(a), ref)


def test_exact_forward_matches_torch_modules():
    rng = np.random.default_rng(2)
    th = rand_theta(rng)
    x = rng.normal(0, 1, (300, R.IN))
    z, _ = R.forward(th, x, rounding=False)
    zt = torch_net(th)(torch.tensor(x)).detach().numpy()[:, 0]
    assert np.allclose(z, zt, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("pos_weight", [1.0, 3.5])
def test_exact_gradient_matches_torch_autograd(pos_weight):
    rng = np.random.default_rng(3)
    th = rand_theta(rng)
    x = rng.normal(0, 1, (200, R.IN))
    y = (rng.random(200) < 0.4).astype(np.float64)
    g, loss = R.gradient(th, x, y, pos_weight, rounding=False)
    net = torch_net(th)
    crit = torch.nn.BCEWithLogitsLoss(pos_weight=torch.tensor(pos_weight, dtype=torch.float64))
    lt = crit(net(torch.tensor(x))[:, 0], torch.tensor(y))
    lt.backward()
    assert abs(loss - lt.item()) < 1e-12
    gt = torch_grad(net)
    assert np.allclose(g, gt, rtol=1e-9, atol=1e-13)


def test_exact_gradient_matches_finite_differences():
    rng = np.random.default_rng(4)
    th = rand_theta(rng)
    x = rng.normal(0, 1, (64, R.IN))
    y = (rng.random(64) < 0.5).astype(np.float64)
    g, _ = R.gradient(th, x, y, 2.0, rounding=False)
    idx = np.concatenate([rng.choice(R.OFF_B1, 12, replace=False), R.OFF_B1 + rng.choice(R.H, 4, replace=False),
                          R.OFF_W2 + rng.choice(R.H * R.H, 12, replace=False),
                          R.OFF_B2 + rng.choice(R.H, 4, replace=False), R.OFF_W3 + rng.choice(R.H, 4, replace=False),
                          [R.OFF_B3]])
    h = 1e-6
    for i in idx:
        tp, tm = th.copy(), th.copy()
        tp[i] += h
        tm[i] -= h
        fd = (R.loss_of(tp, x, y, 2.0, rounding=False) - R.loss_of(tm, x, y, 2.0, rounding=False)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])) + 1e-8, (i, fd, g[i])


def test_rounded_gradient_stays_close_to_exact():
    """The BF16 rounding points perturb the gradient by ~2^-8 relative, never more."""
    rng = np.random.default_rng(5)
    th = rand_theta(rng)
    x = R.bf16(rng.normal(0, 1, (256, R.IN)))
    y = (rng.random(256) < 0.3).astype(np.float64)
    g, _ = R.gradient(th, x, y)
    ge, _ = R.gradient(R.bf16(th.astype(np.float32)), x, y, rounding=False)
    assert np.linalg.norm(g - ge) < 0.02 * np.linalg.norm(ge)


def test_adam_matches_torch_optimizer():
    rng = np.random.default_rng(6)
    th = rng.normal(0, 1, 500)
    p = torch.nn.Parameter(torch.tensor(th.copy()))
    opt = torch.optim.Adam([p], lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    m = np.zeros(500)
    v = np.zeros(500)
    cur = th.copy()
    for t in range(1, 6):
        g = rng.normal(0, 1, 500) * (10.0 ** rng.integers(-6, 2, 500))
        p.grad = torch.tensor(g.copy())
        opt.step()
        cur, m, v = R.adam_step(cur, m, v, g, t)
        assert np.allclose(cur, p.detach().numpy(), rtol=1e-12, atol=1e-15), t
