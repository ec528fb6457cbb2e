"""bin_particles with the reference's 1/3/5-bin signatures
(1bin/kernels/bin_particles.py:24-31, 3bin :71-78, 5bin :130-140).

The reference appends row indices to per-(bin, cell) lists with atomic
slots, from tosum[:,3] (the energy post_processing left) and curr_xbins.
Here libbz's bz_bin_lists reads the same arrays in place (torch tensors,
__cuda_array_interface__ or DLPack objects, no copy): the keys come from
tosum[:,3] with the reference's compares, a stable radix sort orders them,
and a kernel fills every list in ascending row order up to its max_ppc and
writes the full counts to write_inds, as the reference's atomics do.
"""
from __future__ import annotations

import ctypes as C

from .. import _lib
from . import _session as S

_THRESHOLDS = {1: (), 3: (10.0, 15.76), 5: (5.0, 10.0, 15.76, 20.0)}


def _bin(d_tosum_ar, lists, inds, d_curr_xbins, Nc, N, maxes):
    """lists / inds / maxes ordered from the lowest energy bin upwards."""
    nb = len(lists)
    Nc, N = int(Nc), int(N)
    tosum = S.as_gpu(d_tosum_ar, "d_tosum_ar")
    cx = S.as_gpu(d_curr_xbins, "d_curr_xbins")
    lists = [S.as_gpu(a, "cell list") for a in lists]
    inds = [S.as_gpu(a, "write_inds") for a in inds]
    import torch
    if tosum.dtype != torch.float64 or tosum.dim() != 2 or not tosum.is_contiguous():
        raise TypeError("d_tosum_ar must be a contiguous float64 (Nmax, 4) device array")
    if cx.dtype != torch.int64 or any(a.dtype != torch.int64 for a in lists + inds):
        raise TypeError("curr_xbins, the cell lists and write_inds must be int64")
    for a in lists:
        if a.dim() != 2 or a.shape[1] != N + 1 or not a.is_contiguous():
            raise TypeError("cell lists must be contiguous (max_ppc, N+1) arrays")
    cap = max(int(tosum.shape[0]), 1)
    dev = tosum.device.index or 0
    s = S.session(dev, _lib.BZ_F64, cap, N + 1, (N + 1) * nb)
    thr = (C.c_double * 8)(*_THRESHOLDS[nb])
    lp = (C.c_void_p * nb)(*[a.data_ptr() for a in lists])
    ip = (C.c_void_p * nb)(*[a.data_ptr() for a in inds])
    mp = (C.c_int64 * nb)(*[min(int(m), int(a.shape[0])) for m, a in zip(maxes, lists)])
    _lib.check(s.lib.bz_bin_lists(s.ctx, C.c_void_p(tosum.data_ptr()), tosum.shape[1],
                                  C.c_void_p(cx.data_ptr()), Nc, N, nb, thr, lp, ip, mp),
               "bz_bin_lists")


def bin_particles(*args):
    """1bin: (tosum, LE_list, LE_inds, curr_xbins, Nc, N, max_LE, tpb)
    3bin: (tosum, HE, ME, LE lists, HE, ME, LE inds, curr_xbins, Nc, N, max_HE, max_ME,
           max_LE, tpb)
    5bin: (tosum, HE, ME1, ME2, ME3, LE lists, HE, ME1, ME2, ME3, LE inds, curr_xbins, Nc, N,
           max_HE, max_ME1, max_ME2, max_ME3, max_LE, tpb)"""
    if len(args) == 8:
        tosum, le, le_i, cx, Nc, N, mle, _ = args
        return _bin(tosum, [le], [le_i], cx, Nc, N, [mle])
    if len(args) == 14:
        tosum, he, me, le, he_i, me_i, le_i, cx, Nc, N, mhe, mme, mle, _ = args
        return _bin(tosum, [le, me, he], [le_i, me_i, he_i], cx, Nc, N, [mle, mme, mhe])
    if len(args) == 20:
        (tosum, he, m1, m2, m3, le, he_i, m1_i, m2_i, m3_i, le_i, cx, Nc, N, mhe, mm1, mm2, mm3,
         mle, _) = args
        return _bin(tosum, [le, m1, m2, m3, he], [le_i, m1_i, m2_i, m3_i, he_i], cx, Nc, N,
                    [mle, mm1, mm2, mm3, mhe])
    raise TypeError(f"bin_particles takes 8, 14 or 20 arguments ({len(args)} given)")


