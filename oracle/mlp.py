"""CPU oracle of the learned closure (test infrastructure only).

FP64 restatement of the perceptron the HIP kernels run: inputs normalised
as x_n = (x - in_shift) * in_scale, three hidden layers of 64 SiLU units,
a linear output, y = y_n * out_scale + out_shift.  The weights are the BF16
copy of an FP32 master (round to nearest even), here widened exactly to
FP64, so the oracle and the kernels see identical parameters and differ
only by FP32 vs FP64 arithmetic.  The loss is L = (1/2n) sum |y_n - t_n|^2
in normalised output units.  The analytic backward pass is pinned by
central finite differences in FP64 (tests/test_mlp_oracle.py).

The closure has no counterpart in the reference; the layer sizes, the
activation and the BF16/FP32 split follow the task's closure description.
"""
from __future__ import annotations

import numpy as np

N_IN, N_H, N_OUT = 3, 64, 3
# flat layout, row-major W[out][in]
SHAPES = [("W1", (N_H, N_IN)), ("b1", (N_H,)), ("W2", (N_H, N_H)), ("b2", (N_H,)),
          ("W3", (N_H, N_H)), ("b3", (N_H,)), ("W4", (N_OUT, N_H)), ("b4", (N_OUT,))]
OFFSETS = {}
_o = 0
for _name, _shape in SHAPES:
    OFFSETS[_name] = (_o, _shape)
    _o += int(np.prod(_shape))
N_PARAM = _o  # 8771


def bf16_bits(master: np.ndarray) -> np.ndarray:
    """Round FP32 values to BF16 (nearest, ties to even); returns uint16 bits."""
    u = np.ascontiguousarray(master, dtype=np.float32).view(np.uint32).astype(np.uint64)
    nan = ((u & 0x7F800000) == 0x7F800000) & ((u & 0x7FFFFF) != 0)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    r[nan] = 0x7FC0
    return r


def bf16_value(bits: np.ndarray) -> np.ndarray:
    """Exact FP64 value of BF16 bit patterns."""
    return (np.asarray(bits, dtype=np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(theta: np.ndarray) -> dict:
    return {k: theta[o:o + int(np.prod(sh))].reshape(sh) for k, (o, sh) in OFFSETS.items()}


def _silu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return z * s, s * (1.0 + z * (1.0 - s))


def forward(theta, norm, x):
    """y[n][3] in physical units; theta are the (BF16-valued) parameters."""
    P = unpack(np.asarray(theta, dtype=np.float64))
    xn = (np.asarray(x, np.float64) - norm["in_shift"]) * norm["in_scale"]
    h = xn
    for k in (1, 2, 3):
        h, _ = _silu(h @ P[f"W{k}"].T + P[f"b{k}"])
    yn = h @ P["W4"].T + P["b4"]
    return yn * norm["out_scale"] + norm["out_shift"]


def loss_and_grad(theta, norm, x, t):
    """L and dL/dtheta (flat, FP64) for a batch."""
    P = unpack(np.asarray(theta, dtype=np.float64))
    n = x.shape[0]
    xn = (np.asarray(x, np.float64) - norm["in_shift"]) * norm["in_scale"]
    tn = (np.asarray(t, np.float64) - norm["out_shift"]) / norm["out_scale"]
    acts, ders = [xn], []
    h = xn
    for k in (1, 2, 3):
        h, d = _silu(h @ P[f"W{k}"].T + P[f"b{k}"])
        acts.append(h)
        ders.append(d)
    yn = h @ P["W4"].T + P["b4"]
    e = yn - tn
    loss = 0.5 * np.sum(e * e) / n
    g = {}
    delta = e / n
    g["W4"] = delta.T @ acts[3]
    g["b4"] = delta.sum(axis=0)
    back = delta @ P["W4"]
    for k in (3, 2, 1):
        delta = back * ders[k - 1]
        g[f"W{k}"] = delta.T @ acts[k - 1]
        g[f"b{k}"] = delta.sum(axis=0)
        back = delta @ P[f"W{k}"]
    flat = np.zeros(N_PARAM)
    for k, (o, sh) in OFFSETS.items():
        flat[o:o + int(np.prod(sh))] = g[k].ravel()
    return loss, flat


def adam(master, m, v, grad, lr, beta1, beta2, eps, step):
    """One Adam update in FP32, the order the kernel uses."""
    f = np.float32
    g = grad.astype(f)
    m2 = f(beta1) * m + (f(1) - f(beta1)) * g
    v2 = f(beta2) * v + (f(1) - f(beta2)) * g * g
    b1 = f(1.0 - beta1 ** step)
    b2 = f(1.0 - beta2 ** step)
    w = master - f(lr) * (m2 / b1) / (np.sqrt(v2 / b2) + f(eps))
    return w.astype(f), m2.astype(f), v2.astype(f)
