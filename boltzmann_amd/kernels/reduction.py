"""The reference's native reduction entry (kernels/reduction/__init__.py:1-11,
kernels/reduction/cuda_reduction/__init__.py:11-15) on bz_reduction.

reduction(r, p, M, RM, dw, dv, dx, ptcl): dw (M, r) FP64 device array
receives in rows 0..RM-1 the sums of the value rows dv (ptcl, r) over the
coarse rows int(dx*M) // (M/RM); the remaining rows are zeroed.  init_reduction
keeps the backend switch of the reference; there is one backend.
"""
from __future__ import annotations

import torch

from .. import _lib
from . import _session as S

reduction = None


def _reduction(r: int, p: int, M: int, RM: int, dw, dv, dx, ptcl: int):
    for name, t in (("dw", dw), ("dv", dv), ("dx", dx)):
        if not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype != torch.float64:
            raise TypeError(f"{name} must be a float64 CUDA (HIP) tensor")
        if not t.is_contiguous():
            raise ValueError(f"{name} must be contiguous")
    dev = dw.device.index or 0
    # a context whose node count covers RM rows of r columns (its partial
    # and level-2 buffers are sized from it)
    s = S.session(dev, _lib.BZ_F64, max(int(ptcl), 1), (int(RM) * int(r) + 3) // 4 + 1, 1)
    _lib.check(s.lib.bz_reduction(s.ctx, int(r), int(M), int(RM), S.ptr(dw), S.ptr(dv),
                                  S.ptr(dx), int(ptcl)), "bz_reduction")


def init_reduction(pyk: bool = False):
    """Select the reduction backend (the reference's -p flag picks pykokkos;
    this library has the HIP backend only)."""
    global reduction
    reduction = _reduction
    return reduction


init_reduction(False)
