"""The oracle's learned-preconditioner network (3 -> 64 -> 64 -> 1, GELU in
its tanh form) pinned against independent implementations: torch's GELU
(approximate='tanh') and a torch network built from the same parameter
layout, whose autograd gradient and central finite differences of the
oracle's loss must both match the oracle's backpropagation."""
import numpy as np
import pytest
import torch

from oracle import oracle as orc

P = orc.MLP_NPARAM


def _params(seed):
    rng = np.random.default_rng(seed)
    p = np.zeros(P)
    p[0:192] = rng.standard_normal(192) * 0.8
    p[192:256] = rng.standard_normal(64) * 0.3
    p[256:4352] = rng.standard_normal(4096) / 8.0
    p[4352:4416] = rng.standard_normal(64) * 0.3
    p[4416:4480] = rng.standard_normal(64) / 8.0
    p[4480] = 0.1
    return p


def _torch_net(p):
    t = torch.tensor(p, dtype=torch.float64)
    W1 = t[0:192].reshape(64, 3)
    b1 = t[192:256]
    W2 = t[256:4352].reshape(64, 64)
    b2 = t[4352:4416]
    W3 = t[4416:4480]
    b3 = t[4480]

    def f(x):
        g = torch.nn.functional.gelu
        h1 = g(x @ W1.T + b1, approximate="tanh")
        h2 = g(h1 @ W2.T + b2, approximate="tanh")
        return h2 @ W3 + b3
    return f


def test_gelu_matches_torch_and_identities():
    xs = np.array([-6.0, -2.5, -1.0, -0.3, 0.0, 1e-3, 0.7, 1.0, 3.0, 8.0])
    ours = np.array([orc.gelu(x) for x in xs])
    ref = torch.nn.functional.gelu(torch.tensor(xs), approximate="tanh").numpy()
    assert np.abs(ours - ref).max() <= 1e-15
    # gelu(x) - gelu(-x) = x for the tanh form, and it tracks x * Phi(x)
    neg = np.array([orc.gelu(-x) for x in xs])
    assert np.abs(ours - neg - xs).max() <= 1e-14
    from scipy.special import ndtr
    assert np.abs(ours - xs * ndtr(xs)).max() <= 5e-4


@pytest.mark.parametrize("seed", [1, 2])
def test_mlp_forward_matches_torch(seed):
    p = _params(seed)
    rng = np.random.default_rng(100 + seed)
    x = rng.standard_normal((257, 3)) * np.array([2.0, 0.5, 0.3]) + np.array([0.0, 0.0, 0.5])
    y = orc.mlp_forward(p, x.ravel())
    ref = _torch_net(p)(torch.tensor(x)).numpy()
    assert np.abs(y - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("seed", [3, 4])
def test_mlp_gradient_matches_autograd(seed):
    p = _params(seed)
    rng = np.random.default_rng(200 + seed)
    m = 97
    x = rng.standard_normal((m, 3))
    q = rng.standard_normal(m)
    t = rng.standard_normal(m)
    w = rng.uniform(0.1, 2.0, m)
    g, loss = orc.mlp_grad(p, x.ravel(), q, t, w)
    pt = torch.tensor(p, dtype=torch.float64, requires_grad=True)
    W1, b1 = pt[0:192].reshape(64, 3), pt[192:256]
    W2, b2 = pt[256:4352].reshape(64, 64), pt[4352:4416]
    W3, b3 = pt[4416:4480], pt[4480]
    gl = torch.nn.functional.gelu
    xt = torch.tensor(x)
    y = gl(gl(xt @ W1.T + b1, approximate="tanh") @ W2.T + b2, approximate="tanh") @ W3 + b3
    lt = (torch.tensor(w) * (torch.tensor(q) * y - torch.tensor(t)) ** 2).sum()
    lt.backward()
    ref = pt.grad.numpy()
    assert abs(loss - lt.item()) <= 1e-12 * abs(lt.item())
    assert np.abs(g - ref).max() <= 1e-11 * np.abs(ref).max()
    assert abs(loss - orc.mlp_loss(p, x.ravel(), q, t, w)) <= 1e-12 * loss


def test_mlp_gradient_finite_differences():
    """Central differences of the loss on a sample of parameters from every
    layer (weights and biases)."""
    p = _params(5)
    rng = np.random.default_rng(7)
    m = 31
    x = rng.standard_normal(3 * m)
    q = rng.standard_normal(m)
    t = rng.standard_normal(m)
    w = np.ones(m)
    g, _ = orc.mlp_grad(p, x, q, t, w)
    picks = [0, 1, 2, 100, 191, 192, 230, 255, 256, 300, 2000, 4351, 4352, 4400, 4415, 4416, 4450, 4479, 4480]
    h = 1e-6
    for i in picks:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (orc.mlp_loss(pp, x, q, t, w) - orc.mlp_loss(pm, x, q, t, w)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])
