"""Closure network kernels against the FP64 oracle: the BF16 copy bit for
bit, forward outputs, the batch gradient (an FP64 analytic backward pinned
by finite differences), the Adam update and a training run."""
import numpy as np
import pytest

from oracle import mlp as M

pytestmark = pytest.mark.gpu


def _net(seed=3):
    from boltzmann_amd.closure import ClosureNet
    net = ClosureNet(seed=seed)
    net.norm = {"in_shift": np.array([1.5, 6.0, 16.0]), "in_scale": np.array([0.9, 0.12, 0.6]),
                "out_shift": np.array([-17.0, -16.0, 24.0]),
                "out_scale": np.array([2.0, 1.5, 0.4])}
    return net


def _data(n, seed=0):
    rng = np.random.default_rng(seed)
    x = np.stack([rng.uniform(0, 3, n), rng.uniform(0.3, 20, n), rng.uniform(13, 19, n)], 1)
    t = np.stack([-16 - 2 * np.cos(x[:, 0]), -15.5 - 0.05 * x[:, 1], 24 + 0.2 * x[:, 0]], 1)
    return x.astype(np.float32), t.astype(np.float32)


def _theta(net):
    return M.bf16_value(net.bf16_bits())


def test_bf16_copy_bit_exact(gpu):
    net = _net()
    np.testing.assert_array_equal(net.bf16_bits(), M.bf16_bits(net.parameters()))
    net.close()


@pytest.mark.parametrize("n", [1, 77, 4096])
def test_forward_matches_oracle(gpu, n):
    net = _net()
    x, _ = _data(n)
    got = net.predict(x)
    want = M.forward(_theta(net), net.norm, x.astype(np.float64))
    tol = 2e-5 * net.norm["out_scale"] + 1e-6 * np.abs(want)
    assert np.all(np.abs(got - want) <= tol)
    net.close()


def test_forward_empty(gpu):
    import torch
    net = _net()
    y = net.forward(torch.zeros(0, 3, device="cuda"))
    assert y.shape == (0, 3)
    net.close()


@pytest.mark.parametrize("n", [64, 301])
def test_gradient_matches_oracle(gpu, n):
    net = _net()
    x, t = _data(n, seed=n)
    import torch
    g, loss = net.grad(torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda())
    g = g.cpu().numpy().astype(np.float64)
    loss = float(loss.item())
    want_loss, want = M.loss_and_grad(_theta(net), net.norm, x.astype(np.float64),
                                      t.astype(np.float64))
    assert abs(loss - want_loss) <= 1e-5 * want_loss
    scale = np.abs(want).max()
    assert np.all(np.abs(g - want) <= 2e-5 * scale + 1e-4 * np.abs(want)), \
        np.abs(g - want).max() / scale
    # the sum runs in a fixed order: a second call gives the same bits
    g2, _ = net.grad(torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda())
    np.testing.assert_array_equal(g2.cpu().numpy(), g.astype(np.float32))
    net.close()


def test_adam_matches_oracle(gpu):
    import torch
    net = _net()
    x, t = _data(500, seed=5)
    w0 = net.parameters()
    m0 = np.zeros_like(w0)
    v0 = np.zeros_like(w0)
    for step in (1, 2):
        g, _ = net.grad(torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda())
        gh = g.cpu().numpy()
        net.adam_step(1e-3)
        w1, m1, v1 = M.adam(w0, m0, v0, gh, 1e-3, 0.9, 0.999, 1e-8, step)
        got = net.parameters()
        np.testing.assert_allclose(got, w1, rtol=1e-6, atol=1e-9)
        np.testing.assert_allclose(net.m.cpu().numpy(), m1, rtol=1e-6, atol=1e-12)
        np.testing.assert_allclose(net.v.cpu().numpy(), v1, rtol=1e-6, atol=1e-15)
        np.testing.assert_array_equal(net.bf16_bits(), M.bf16_bits(got))
        w0, m0, v0 = got, m1, v1
    net.close()


def test_training_reduces_loss(gpu):
    net = _net(seed=11)
    x, t = _data(40_000, seed=9)
    net.fit_normalisation(x, t)
    before = net.loss(x, t)
    hist = net.train(x, t, epochs=15, batch=2048, lr=3e-3)
    after = net.loss(x, t)
    assert hist[-1] < hist[0]
    assert after < 0.02 * before, (before, after)
    net.close()


def test_errors(gpu):
    import torch
    from boltzmann_amd import _lib
    net = _net()
    with pytest.raises(_lib.BzError):
        net.grad(torch.zeros(0, 3, device="cuda"), torch.zeros(0, 3, device="cuda"))
    net.norm["out_scale"] = np.array([1.0, 0.0, 1.0])
    with pytest.raises(_lib.BzError):
        net.forward(torch.zeros(4, 3, device="cuda"))
    net.close()
