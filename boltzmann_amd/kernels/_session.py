"""Contexts and scratch buffers shared by the reference-surface stage functions.

The stage functions take the reference's arrays (AoS rows on the GPU as
torch tensors, grid arrays on the host as numpy) and hand SoA copies to the
HIP kernels of libbz.so.  One libbz context is kept per (device, precision,
capacity, nodes) and reused across calls.
"""
from __future__ import annotations

import ctypes as C
from typing import Dict, Tuple

import numpy as np
import torch

from .. import _lib


class Session:
    def __init__(self, device: int, precision: int, capacity: int, max_nodes: int,
                 max_keys: int):
        if not torch.cuda.is_available():
            raise _lib.BzError("no GPU visible: the stage functions run only in the HIP kernels")
        self.lib = _lib.load()
        self.dev = torch.device("cuda", device)
        self.precision = precision
        self.dtype = torch.float64 if precision == _lib.BZ_F64 else torch.float32
        self.capacity = capacity
        self.ctx = C.c_void_p()
        _lib.check(self.lib.bz_ctx_create(C.byref(self.ctx), device, precision, capacity,
                                          max(max_nodes, 2), max(max_keys, 1)), "bz_ctx_create")
        self.stream = torch.cuda.current_stream(self.dev)
        _lib.check(self.lib.bz_set_stream(self.ctx, C.c_void_p(self.stream.cuda_stream)))
        z = dict(dtype=self.dtype, device=self.dev)
        self.soa = [[torch.zeros(capacity, **z) for _ in range(5)] for _ in range(2)]
        self.structs = [_lib.State(*[a.data_ptr() for a in bufs]) for bufs in self.soa]
        self.R = torch.zeros((7, capacity), **z)
        self.G = torch.zeros((3, capacity), **z)
        self.moments = torch.zeros(3 * max(max_nodes, 1), dtype=torch.float64, device=self.dev)
        self.step = 0
        # moments of the last push, keyed by the AoS buffer and its row count
        self.cached: Dict[Tuple[int, int], Tuple[torch.Tensor, int]] = {}

    def clock_set(self, n: int, t: float = 0.0):
        _lib.check(self.lib.bz_clock_set(self.ctx, n, self.step, t), "bz_clock_set")

    def count(self) -> int:
        n = C.c_int64()
        _lib.check(self.lib.bz_clock_get(self.ctx, C.byref(n), None, None, None))
        return n.value


_sessions: Dict[tuple, Session] = {}


def session(device: int, precision: int, capacity: int, max_nodes: int,
            max_keys: int = 1) -> Session:
    for key, s in _sessions.items():
        if (key[0] == device and key[1] == precision and s.capacity >= capacity
                and key[3] >= max_nodes and key[4] >= max_keys):
            return s
    key = (device, precision, capacity, max_nodes, max_keys)
    s = Session(device, precision, capacity, max_nodes, max_keys)
    _sessions[key] = s
    return s


def ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def as_gpu(a, name: str = "array") -> torch.Tensor:
    """A device array of the caller as a torch tensor over the same memory
    (no copy): a torch tensor, any object exposing __cuda_array_interface__
    (CuPy and numba device arrays), or one exposing DLPack."""
    if isinstance(a, torch.Tensor):
        t = a
    elif hasattr(a, "__cuda_array_interface__"):
        t = torch.as_tensor(a, device="cuda")
    elif hasattr(a, "__dlpack__"):
        t = torch.from_dlpack(a)
    else:
        raise TypeError(f"{name} must be a device array (torch tensor, __cuda_array_interface__ "
                        f"or DLPack), got {type(a).__name__}")
    if not t.is_cuda:
        raise TypeError(f"{name} must live in device memory")
    return t


def field_array(a):
    """A grid array of the caller: a device array (as_gpu) or host numpy."""
    if isinstance(a, torch.Tensor) or hasattr(a, "__cuda_array_interface__") or \
            hasattr(a, "__dlpack__") and not isinstance(a, np.ndarray):
        return as_gpu(a)
    return np.asarray(a)


def as_device(a, dev, dtype=torch.float64) -> torch.Tensor:
    if isinstance(a, torch.Tensor):
        return a.to(device=dev, dtype=dtype)
    return torch.as_tensor(np.asarray(a), dtype=dtype, device=dev)
