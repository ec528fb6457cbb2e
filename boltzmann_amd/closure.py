"""Learned closure: a small perceptron for the ionization / excitation rate
coefficients and the mobility as functions of the local reduced field, the
mean electron energy and the electron density, trained on batched-swarm
data with the library's own forward / backward / Adam kernels.

    inputs   [log10 E/N (Td), mean energy (eV), log10 n_e (m^-3)]
    outputs  [log10 k_ion (m^3/s), log10 k_exc (m^3/s), log10 mu N (1/(V m s))]

Every pass reads the BF16 copy of an FP32 master; the FP32 master, the Adam
moments and the normalisation live here.  There is no CPU path: the network
runs only in the HIP kernels (bz_mlp_*).
"""
from __future__ import annotations

import ctypes as C
import math
import sys
from typing import Optional

import numpy as np
import torch

from . import _lib

N_IN, N_H, N_OUT = 3, 64, 3
N_PARAM = _lib.BZ_MLP_NPARAM
# flat layout (row-major W[out][in]); matches include/bz.h
LAYOUT = [("W1", (N_H, N_IN)), ("b1", (N_H,)), ("W2", (N_H, N_H)), ("b2", (N_H,)),
          ("W3", (N_H, N_H)), ("b3", (N_H,)), ("W4", (N_OUT, N_H)), ("b4", (N_OUT,))]
RATE_FLOOR = 1e-22   # m^3/s: log targets use max(k, floor)
MOB_FLOOR = 1e20     # 1/(V m s)


def _p(t: Optional[torch.Tensor]):
    return None if t is None else C.c_void_p(t.data_ptr())


def init_parameters(seed: int = 0) -> np.ndarray:
    """FP32 master: uniform(+-sqrt(6 / fan_in)) weights (He), zero biases."""
    rng = np.random.default_rng(seed)
    out = []
    for name, shape in LAYOUT:
        if name.startswith("W"):
            lim = math.sqrt(6.0 / shape[1]) * (0.5 if name == "W4" else 1.0)
            out.append(rng.uniform(-lim, lim, shape).ravel())
        else:
            out.append(np.zeros(shape))
    theta = np.concatenate(out).astype(np.float32)
    assert theta.size == N_PARAM
    return theta


class ClosureNet:
    """The network on one GPU.  ctx: an existing library context (its stream
    is used), or None to create a small one."""

    def __init__(self, device: int = 0, seed: int = 0, ctx=None):
        if not torch.cuda.is_available():
            raise _lib.BzError("no GPU visible: the closure runs only in the HIP kernels")
        self.lib = _lib.load()
        self.dev = torch.device("cuda", device)
        torch.cuda.set_device(self.dev)
        self.own_ctx = ctx is None
        if ctx is None:
            ctx = C.c_void_p()
            _lib.check(self.lib.bz_ctx_create(C.byref(ctx), device, _lib.BZ_F32, 1024, 2, 1),
                       "bz_ctx_create")
            self.stream = torch.cuda.current_stream(self.dev)
            _lib.check(self.lib.bz_set_stream(ctx, C.c_void_p(self.stream.cuda_stream)))
        self.ctx = ctx
        f32 = dict(dtype=torch.float32, device=self.dev)
        self.master = torch.from_numpy(init_parameters(seed)).to(self.dev)
        self.m = torch.zeros(N_PARAM, **f32)
        self.v = torch.zeros(N_PARAM, **f32)
        self.w = torch.zeros(N_PARAM, dtype=torch.int16, device=self.dev)  # BF16 bits
        self.grad_buf = torch.zeros(N_PARAM, **f32)
        self.loss_buf = torch.zeros(1, **f32)
        self.ws = None
        self.norm = {"in_shift": np.zeros(3), "in_scale": np.ones(3),
                     "out_shift": np.zeros(3), "out_scale": np.ones(3),
                     "in_lo": np.zeros(3), "in_hi": np.zeros(3)}
        self.t = 0
        self.refresh()

    # ------------------------------------------------------------------
    def refresh(self):
        """BF16 copy of the master."""
        _lib.check(self.lib.bz_mlp_round(self.ctx, _p(self.master), _p(self.w)), "bz_mlp_round")

    def struct(self) -> _lib.Mlp:
        s = _lib.Mlp()
        s.w = self.w.data_ptr()
        for k in ("in_shift", "in_scale", "out_shift", "out_scale"):
            getattr(s, k)[:] = [float(v) for v in self.norm[k]]
        for k in ("in_lo", "in_hi"):
            getattr(s, k)[:] = [float(v) for v in self.norm.get(k, np.zeros(3))]
        return s

    def fit_normalisation(self, x: np.ndarray, t: np.ndarray):
        """Standardise inputs and targets with the training set statistics."""
        xs = np.std(x, axis=0)
        ts = np.std(t, axis=0)
        self.norm = {"in_shift": np.mean(x, axis=0), "in_scale": 1.0 / np.where(xs > 0, xs, 1.0),
                     "out_shift": np.mean(t, axis=0), "out_scale": np.where(ts > 0, ts, 1.0),
                     "in_lo": np.min(x, axis=0).astype(np.float64),
                     "in_hi": np.max(x, axis=0).astype(np.float64)}

    def set_parameters(self, theta: np.ndarray):
        self.master.copy_(torch.from_numpy(np.asarray(theta, np.float32)))
        self.refresh()

    def parameters(self) -> np.ndarray:
        torch.cuda.synchronize(self.dev)
        return self.master.cpu().numpy()

    def bf16_bits(self) -> np.ndarray:
        torch.cuda.synchronize(self.dev)
        return self.w.cpu().numpy().view(np.uint16)

    # ------------------------------------------------------------------
    def forward(self, x: torch.Tensor, out: Optional[torch.Tensor] = None) -> torch.Tensor:
        """x: [n, 3] FP32 device tensor in physical units -> [n, 3]."""
        x = x.contiguous()
        n = x.shape[0]
        if out is None:
            out = torch.empty(n, N_OUT, dtype=torch.float32, device=self.dev)
        net = self.struct()
        _lib.check(self.lib.bz_mlp_forward(self.ctx, C.byref(net), _p(x), _p(out), n),
                   "bz_mlp_forward")
        return out

    def grad(self, x: torch.Tensor, t: torch.Tensor):
        """Gradient of the batch loss (device tensor) and the loss (device scalar)."""
        n = x.shape[0]
        need = int(self.lib.bz_mlp_workspace_floats(n))
        if self.ws is None or self.ws.numel() < need:
            self.ws = torch.empty(need, dtype=torch.float32, device=self.dev)
        net = self.struct()
        _lib.check(self.lib.bz_mlp_grad(self.ctx, C.byref(net), _p(x.contiguous()),
                                        _p(t.contiguous()), n, _p(self.ws), _p(self.grad_buf),
                                        _p(self.loss_buf)), "bz_mlp_grad")
        return self.grad_buf, self.loss_buf

    def adam_step(self, lr: float, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8):
        self.t += 1
        opt = _lib.Adam(lr, beta1, beta2, eps, 1.0 - beta1 ** self.t, 1.0 - beta2 ** self.t)
        _lib.check(self.lib.bz_mlp_adam(self.ctx, C.byref(opt), _p(self.master), _p(self.m),
                                        _p(self.v), _p(self.grad_buf), _p(self.w)),
                   "bz_mlp_adam")

    def train(self, x: np.ndarray, t: np.ndarray, epochs: int = 40, batch: int = 8192,
              lr: float = 3e-3, lr_final: float = 1e-4, seed: int = 0, log_every: int = 0):
        """Mini-batch Adam on (x, t) with a cosine learning-rate decay.
        Returns the per-epoch mean training loss."""
        xd = torch.from_numpy(np.ascontiguousarray(x, np.float32)).to(self.dev)
        td = torch.from_numpy(np.ascontiguousarray(t, np.float32)).to(self.dev)
        n = xd.shape[0]
        g = torch.Generator(device=self.dev)
        g.manual_seed(seed)
        steps_per_epoch = max(1, n // batch)
        total = epochs * steps_per_epoch
        hist = []
        k = 0
        for ep in range(epochs):
            perm = torch.randperm(n, device=self.dev, generator=g)
            acc = torch.zeros(1, dtype=torch.float32, device=self.dev)
            for s in range(steps_per_epoch):
                idx = perm[s * batch:(s + 1) * batch]
                xb = xd.index_select(0, idx)
                tb = td.index_select(0, idx)
                self.grad(xb, tb)
                acc += self.loss_buf
                c = 0.5 * (1 + math.cos(math.pi * k / max(total - 1, 1)))
                self.adam_step(lr_final + (lr - lr_final) * c)
                k += 1
            hist.append(float(acc.item()) / steps_per_epoch)
            if log_every and (ep % log_every == 0 or ep == epochs - 1):
                print(f"epoch {ep}: loss {hist[-1]:.4e}", file=sys.stderr, flush=True)
        return hist

    def loss(self, x: np.ndarray, t: np.ndarray) -> float:
        xd = torch.from_numpy(np.ascontiguousarray(x, np.float32)).to(self.dev)
        td = torch.from_numpy(np.ascontiguousarray(t, np.float32)).to(self.dev)
        self.grad(xd, td)
        return float(self.loss_buf.item())

    def predict(self, x: np.ndarray) -> np.ndarray:
        xd = torch.from_numpy(np.ascontiguousarray(x, np.float32)).to(self.dev)
        return self.forward(xd).cpu().numpy()

    # ------------------------------------------------------------------
    def save(self, path):
        np.savez(path, master=self.parameters(), **{k: np.asarray(v) for k, v in self.norm.items()})

    def load(self, path):
        with np.load(path, allow_pickle=False) as z:
            self.norm = {k: np.array(z[k], dtype=np.float64)
                         for k in ("in_shift", "in_scale", "out_shift", "out_scale")}
            for k in ("in_lo", "in_hi"):
                self.norm[k] = np.array(z[k], dtype=np.float64) if k in z else np.zeros(3)
            self.set_parameters(z["master"])

    def close(self):
        if self.own_ctx and self.ctx:
            self.lib.bz_ctx_destroy(self.ctx)
            self.ctx = C.c_void_p()


def features(en_td, mean_energy_ev, ne_m3) -> np.ndarray:
    """Network inputs from physical quantities."""
    en = np.maximum(np.abs(np.asarray(en_td, np.float64)), 1e-3)
    ne = np.maximum(np.asarray(ne_m3, np.float64), 1.0)
    return np.stack([np.log10(en), np.asarray(mean_energy_ev, np.float64), np.log10(ne)], axis=-1)


def targets(k_ion, k_exc, mobility_n) -> np.ndarray:
    """Network targets from rates (m^3/s) and reduced mobility (1/(V m s))."""
    return np.stack([np.log10(np.maximum(k_ion, RATE_FLOOR)),
                     np.log10(np.maximum(k_exc, RATE_FLOOR)),
                     np.log10(np.maximum(mobility_n, MOB_FLOOR))], axis=-1)

