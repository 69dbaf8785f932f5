"""The oracle's MLP (16-64-64-3) forward and backward, pinned by torch
autograd in FP64 (an independent implementation of the same network and
loss) and by central finite differences."""
import numpy as np
import pytest

IN, H, OUT = 16, 64, 3


def unpack(p):
    import torch
    p = torch.as_tensor(p, dtype=torch.float64)
    o = 0
    W1 = p[o:o + IN * H].reshape(IN, H); o += IN * H
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    W3 = p[o:o + H * OUT].reshape(H, OUT); o += H * OUT
    b3 = p[o:o + OUT]; o += OUT
    return W1, b1, W2, b2, W3, b3


def torch_loss(params, X, y, t, m):
    import torch
    p = torch.tensor(params, dtype=torch.float64, requires_grad=True)
    W1, b1, W2, b2, W3, b3 = unpack(p)
    X = torch.as_tensor(X, dtype=torch.float64)
    h1 = torch.relu(X @ W1 + b1)
    h2 = torch.relu(h1 @ W2 + b2)
    z = h2 @ W3 + b3
    ce = torch.nn.functional.cross_entropy(z[:, :2], torch.as_tensor(y, dtype=torch.long), reduction="none")
    mse = 0.5 * torch.as_tensor(m, dtype=torch.float64) * (z[:, 2] - torch.as_tensor(t, dtype=torch.float64)) ** 2
    loss = (ce + mse).mean()
    loss.backward()
    return float(loss), p.grad.numpy(), (h1.detach().numpy(), h2.detach().numpy(), z.detach().numpy())


def problem(seed, B=37):
    rng = np.random.default_rng(seed)
    params = (rng.standard_normal(16 * 64 + 64 + 64 * 64 + 64 + 64 * 3 + 3) * 0.2).astype(np.float32)
    X = rng.standard_normal((B, IN)).astype(np.float32)
    y = rng.integers(0, 2, B).astype(np.int32)
    t = rng.uniform(0, 3, B).astype(np.float32)
    m = (rng.uniform(0, 1, B) < 0.5).astype(np.float32)
    return params, X, y, t, m


@pytest.mark.parametrize("seed", [0, 1])
def test_forward_matches_torch(orc, seed):
    params, X, y, t, m = problem(seed)
    H1, H2, Z = orc.mlp_forward(params, X)
    _, _, (h1, h2, z) = torch_loss(params, X, y, t, m)
    assert np.max(np.abs(H1 - h1)) <= 1e-5 * max(1, np.abs(h1).max())
    assert np.max(np.abs(H2 - h2)) <= 1e-5 * max(1, np.abs(h2).max())
    assert np.max(np.abs(Z - z)) <= 1e-5 * max(1, np.abs(z).max())


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_backward_matches_torch_autograd(orc, seed):
    params, X, y, t, m = problem(seed)
    loss, g = orc.mlp_backward(params, X, y, t, m)
    tl, tg, _ = torch_loss(params, X, y, t, m)
    assert abs(loss - tl) <= 1e-5 * max(1, abs(tl))
    assert np.max(np.abs(g - tg)) <= 1e-5 * max(1e-3, np.abs(tg).max())


def test_backward_matches_finite_differences(orc):
    params, X, y, t, m = problem(5, B=9)
    _, g = orc.mlp_backward(params, X, y, t, m)
    rng = np.random.default_rng(9)
    idx = rng.choice(params.size, 40, replace=False)
    p64 = params.astype(np.float64)
    for i in idx:
        h = 1e-4
        pp = p64.copy(); pp[i] += h
        pm = p64.copy(); pm[i] -= h
        fd = (torch_loss(pp, X, y, t, m)[0] - torch_loss(pm, X, y, t, m)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 2e-3 * max(1e-2, abs(fd)), (i, fd, g[i])


def test_a_dropped_relu_mask_is_caught(orc):
    """A plausible backward mistake (no ReLU mask on the hidden gradient),
    planted in a numpy restatement, disagrees with the oracle."""
    params, X, y, t, m = problem(3)
    _, g = orc.mlp_backward(params, X, y, t, m)
    W1, b1, W2, b2, W3, b3 = [a.numpy() for a in unpack(params)]
    h1 = np.maximum(X @ W1 + b1, 0)
    h2 = np.maximum(h1 @ W2 + b2, 0)
    z = h2 @ W3 + b3
    e = np.exp(z[:, :2] - z[:, :2].max(1, keepdims=True))
    p = e / e.sum(1, keepdims=True)
    dz = np.zeros_like(z)
    dz[:, :2] = p - np.eye(2)[y]
    dz[:, 2] = m * (z[:, 2] - t)
    dz /= X.shape[0]
    d2 = dz @ W3.T  # mask dropped
    gW2_wrong = h1.T @ d2
    o = 16 * 64 + 64
    assert not np.allclose(gW2_wrong.ravel(), g[o:o + 64 * 64], atol=1e-6)
