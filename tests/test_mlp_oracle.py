"""Pins of the closure oracle: BF16 rounding known answers, the parameter
layout, and the analytic gradient against central finite differences."""
import numpy as np

from oracle import mlp as M


def test_layout_size():
    assert M.N_PARAM == 8771
    assert M.OFFSETS["b4"][0] == 8768


def test_bf16_rounding_known_answers():
    vals = np.array([1.0, 1.00390625, 1.01171875, -2.0, 3.0e38, 1e-40, 0.0], dtype=np.float32)
    bits = M.bf16_bits(vals)
    # 1.00390625 = 0x3F808000 is a tie -> even (0x3F80); 1.01171875 =
    # 0x3F818000 is a tie -> even (0x3F82); 3e38 rounds up within range
    assert list(bits[:4]) == [0x3F80, 0x3F80, 0x3F82, 0xC000]
    assert bits[4] == 0x7F62
    assert M.bf16_value(bits[:4]).tolist() == [1.0, 1.0, 1.015625, -2.0]
    nan = np.array([np.nan], dtype=np.float32)
    assert M.bf16_bits(nan)[0] == 0x7FC0
    # round to nearest: |bf16(x) - x| <= half an ulp = 2^(floor(log2|x|) - 8)
    x = np.random.default_rng(0).normal(size=10000).astype(np.float32)
    err = np.abs(M.bf16_value(M.bf16_bits(x)) - x.astype(np.float64))
    half_ulp = 2.0 ** (np.floor(np.log2(np.abs(x.astype(np.float64)))) - 8)
    assert np.all(err <= half_ulp)
    assert np.any(err > 0.9 * half_ulp)  # and the bound is tight


def _norm():
    return {"in_shift": np.array([1.5, 5.0, 15.0]), "in_scale": np.array([0.8, 0.1, 0.5]),
            "out_shift": np.array([-16.0, -15.0, 24.0]), "out_scale": np.array([2.0, 1.5, 0.5])}


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    theta = rng.normal(0, 0.3, M.N_PARAM)
    norm = _norm()
    x = np.stack([rng.uniform(0, 3, 16), rng.uniform(0.5, 15, 16), rng.uniform(13, 17, 16)], 1)
    t = rng.normal(0, 1, (16, 3)) * norm["out_scale"] + norm["out_shift"]
    loss, g = M.loss_and_grad(theta, norm, x, t)
    # the loss matches the forward pass
    yn = (M.forward(theta, norm, x) - norm["out_shift"]) / norm["out_scale"]
    tn = (t - norm["out_shift"]) / norm["out_scale"]
    assert abs(loss - 0.5 * np.sum((yn - tn) ** 2) / 16) <= 1e-12 * loss
    idx = np.concatenate([rng.choice(M.N_PARAM, 60, replace=False),
                          [M.OFFSETS[k][0] for k in ("W1", "b1", "W2", "b3", "W4", "b4")]])
    h = 1e-6
    for i in idx:
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h
        tm[i] -= h
        fd = (M.loss_and_grad(tp, norm, x, t)[0] - M.loss_and_grad(tm, norm, x, t)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(abs(g).max(), 1e-12), (i, fd, g[i])


def test_adam_first_step_is_signed_lr():
    """With m = v = 0 and bias correction, the first Adam step moves every
    parameter by lr * sign(g) (up to eps)."""
    rng = np.random.default_rng(2)
    w = rng.normal(size=100).astype(np.float32)
    g = rng.normal(size=100).astype(np.float32)
    z = np.zeros(100, np.float32)
    w2, m, v = M.adam(w, z, z, g, 1e-3, 0.9, 0.999, 1e-12, 1)
    np.testing.assert_allclose(w - w2, 1e-3 * np.sign(g), rtol=1e-4)
