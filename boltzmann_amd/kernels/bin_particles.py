"""bin_particles with the reference's 1/3/5-bin signatures
(1bin/kernels/bin_particles.py:24-31, 3bin :71-78, 5bin :130-140).

The reference appends particle indices to per-(bin, cell) lists with atomic
slots.  Here the keys (cell from curr_xbins, energy bin from tosum[:,3]) are
sorted by the stable HIP radix sort (bz_bin_sort), so every list holds its
particles in ascending index order; write_inds receive the full counts and
lists are filled up to max_ppc entries, as in the reference.
"""
from __future__ import annotations

import ctypes as C

import torch

from .. import _lib
from .. import physics as ph
from . import _session as S

_THRESHOLDS = {1: (), 3: (10.0, 15.76), 5: (5.0, 10.0, 15.76, 20.0)}


def _bin(d_tosum_ar, lists, inds, d_curr_xbins, Nc, N, maxes):
    """lists / inds / maxes ordered from the lowest energy bin upwards."""
    nb = len(lists)
    Nc, N = int(Nc), int(N)
    cap = max(int(d_tosum_ar.shape[0]), 1)
    dev = d_tosum_ar.device.index or 0
    s = S.session(dev, _lib.BZ_F64, cap, N + 1, (N + 1) * nb)
    a = s.soa[0]
    kev = 0.5 * ph.M_E * ph.J_EV
    # a state whose cell is exactly curr_xbins (dx = 1) and whose energy is tosum[:,3]
    a[0][:Nc].copy_(d_curr_xbins[:Nc].to(torch.float64))
    a[1][:Nc].copy_(torch.sqrt(torch.clamp(d_tosum_ar[:Nc, 3], min=0.0) / kev))
    a[2][:Nc].zero_()
    a[3][:Nc].zero_()
    a[4][:Nc].copy_(d_tosum_ar[:Nc, 0])
    p = _lib.Params()
    p.dx = 1.0
    p.L = float(N + 1)
    p.kev = kev
    p.dim = 1
    p.xs_mode = _lib.BZ_XS_TABLE
    p.n_cells = N
    p.n_nodes = N + 1
    p.bnf = 1
    s.clock_set(Nc)
    n_keys = (N + 1) * nb
    counts = torch.zeros(n_keys + 1, dtype=torch.int32, device=s.dev)
    offsets = torch.zeros(n_keys + 2, dtype=torch.int32, device=s.dev)
    perm = torch.zeros(max(Nc, 1), dtype=torch.int32, device=s.dev)
    thr = (C.c_double * 8)(*_THRESHOLDS[nb])
    _lib.check(s.lib.bz_bin_sort(s.ctx, C.byref(s.structs[0]), C.byref(p), thr, nb,
                                 S.ptr(counts), S.ptr(offsets), S.ptr(perm)), "bz_bin_sort")
    cnt = counts[:n_keys].view(N + 1, nb)
    off = offsets[:n_keys].view(N + 1, nb)
    for b in range(nb):
        inds[b][:N + 1] = cnt[:, b].to(inds[b].dtype)
        cap_b = min(int(maxes[b]), int(lists[b].shape[0]))
        if cap_b <= 0:
            continue
        j = torch.arange(cap_b, device=s.dev).view(-1, 1)
        valid = j < cnt[:, b].view(1, -1).to(torch.int64)
        src = (off[:, b].view(1, -1).to(torch.int64) + j).clamp(max=max(Nc - 1, 0))
        vals = perm.to(torch.int64)[src]
        lists[b][:cap_b, :N + 1] = torch.where(valid, vals, lists[b][:cap_b, :N + 1].to(
            torch.int64)).to(lists[b].dtype)


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


