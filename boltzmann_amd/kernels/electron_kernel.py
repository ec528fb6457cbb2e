"""electron_kernel with the reference's signature (kernels/electron_kernel_numba.py:291-303).

The reference refreshes cp_R_vec with uniforms, runs collisions + push on the
AoS rows in place, appends each ionization's secondary at Nc + slot and
leaves the number of births in d_curr_count[0].  Here the same stage runs in
the fused HIP kernel (reference argon fits, direct collision probability),
which leaves the rows the same way: cp_R_vec / d_R_vec receive the uniforms
used (Philox stream), rows [0, Nc) are pushed in place (an absorbed row gets
w = 0, x = 1e-8 and v = 0, electron_kernel_numba.py:266-277), the secondaries
fill rows [Nc, Nc + births) in parent order (the reference's atomic slots in
one admissible order), and d_curr_count[0] is the number of births.
post_processing then compacts the rows as the reference does.  The node
moments of the pushed state are kept for post_processing.
"""
from __future__ import annotations

import ctypes as C
import math

import torch

from .. import _lib
from .. import physics as ph
from . import _session as S

L_REF = 2.54  # electron_kernel_numba.py:13
_deposit_bins = {"bnf": 4}


def set_deposit_bins(bnf: int) -> None:
    """Fine bins per node used for the moments post_processing returns (-bnf flag)."""
    _deposit_bins["bnf"] = int(bnf)


def _params(N: int, nn: float, dt: float, stride: int, seed: int) -> _lib.Params:
    p = _lib.Params()
    p.dt = dt
    p.xfac = 100 * dt
    p.afac = dt * (ph.Q_E / ph.M_E)
    p.kev = 0.5 * ph.M_E * ph.J_EV
    p.jm = ph.J_EV * ph.M_E
    p.fel = math.sqrt(1 - (2 * ph.M_E / ph.M_N_AR))
    p.L = L_REF
    p.dx = L_REF / N
    p.half_dx = 0.5 * p.dx
    p.L_dx = L_REF + p.dx
    p.dep_M = float(_deposit_bins["bnf"] * (N + 1))
    p.nn = nn
    p.e_ion = ph.E_ION
    p.e_exc = ph.E_EXC
    p.seed = seed
    p.dim = 1
    p.xs_mode = _lib.BZ_XS_FIT
    p.rng_mode = _lib.BZ_RNG_STREAM
    p.n_cells = N
    p.n_nodes = N + 1
    p.bnf = _deposit_bins["bnf"]
    p.r_stride = stride
    return p


def _rows_from_state(s, src, m, d_data_ar, d_tosum_ar, kev):
    """The pushed SoA state into the reference's AoS rows [0, m): x and v in
    data columns 0, 3, 4, 5, w and the energy in tosum columns 0 and 3; an
    absorbed row gets x = 1e-8, v = 0 (electron_kernel_numba.py:266-277)."""
    a = s.soa[src]
    d_data_ar[:m, 0].copy_(a[0][:m])
    d_data_ar[:m, 1:3].zero_()
    for k, col in ((1, 3), (2, 4), (3, 5)):
        d_data_ar[:m, col].copy_(a[k][:m])
    d_tosum_ar[:m, 0].copy_(a[4][:m])
    d_tosum_ar[:m, 1:3].zero_()
    dead = a[4][:m] <= 0
    d_data_ar[:m, 0].masked_fill_(dead, 1e-8)
    for col in (3, 4, 5):
        d_data_ar[:m, col].masked_fill_(dead, 0.0)
    v2 = d_data_ar[:m, 3] ** 2 + d_data_ar[:m, 4] ** 2 + d_data_ar[:m, 5] ** 2
    d_tosum_ar[:m, 3].copy_(kev * v2)


def electron_kernel(Nc, d_data_ar, d_tosum_ar, d_collct_ar, d_E_ar, cp_R_vec, d_R_vec,
                    d_currxbins, d_forE_xbins, nn, dt_ratio, dt, d_curr_count, d_g0_ar, d_g2_ar,
                    d_gas_v_ar, num_blocks=None, threads_per_block=None, seed: int = 12345):
    """The reference's arrays are used in place: torch tensors or any object
    exposing __cuda_array_interface__ (CuPy, numba device arrays) or DLPack.
    The collision counters of the push are kept for post_processing (the
    reference's d_collct_ar rows are not written: post_processing returns
    the counters this push deposited, counter_g0..g2)."""
    d_data_ar = S.as_gpu(d_data_ar, "d_data_ar")
    d_tosum_ar = S.as_gpu(d_tosum_ar, "d_tosum_ar")
    if d_data_ar.dtype != torch.float64 or d_data_ar.dim() != 2 or d_data_ar.shape[1] < 6:
        raise TypeError("d_data_ar must be a float64 (Nmax, 6) device array")
    cp_R_vec = None if cp_R_vec is None else S.as_gpu(cp_R_vec, "cp_R_vec")
    d_R_vec = None if d_R_vec is None else S.as_gpu(d_R_vec, "d_R_vec")
    d_gas_v_ar = None if d_gas_v_ar is None else S.as_gpu(d_gas_v_ar, "d_gas_v_ar")
    Nc = int(Nc)
    cap = int(d_data_ar.shape[0])
    N = int(S.field_array(d_E_ar).shape[0])
    dev = d_data_ar.device.index or 0
    s = S.session(dev, _lib.BZ_F64, cap, N + 1)
    _lib.check(s.lib.bz_set_collision_counts(s.ctx, 1))
    kev = 0.5 * ph.M_E * ph.J_EV
    a = s.soa[0]
    n = Nc
    a[0][:n].copy_(d_data_ar[:n, 0])
    for k, col in ((1, 3), (2, 4), (3, 5)):
        a[k][:n].copy_(d_data_ar[:n, col])
    a[4][:n].copy_(d_tosum_ar[:n, 0])
    cur = 0
    births_total = 0
    steps = max(int(dt_ratio), 1)
    E = S.as_device(S.field_array(d_E_ar), s.dev)
    for sub in range(steps):
        if sub > 0:
            # the next sub-step pushes the compacted rows (post_processing's filter)
            s.clock_set(n)
            _lib.check(s.lib.bz_compact(s.ctx, C.byref(s.structs[cur]),
                                        C.byref(s.structs[1 - cur])), "bz_compact")
            cur = 1 - cur
            n = s.count()
        _lib.check(s.lib.bz_fill_uniforms(s.ctx, S.ptr(s.R), s.capacity, n, seed, s.step))
        if sub == steps - 1:
            if cp_R_vec is not None and n:
                cp_R_vec[:n, :].copy_(s.R[:, :n].t())
            if d_R_vec is not None and d_R_vec is not cp_R_vec and n:
                d_R_vec[:n, :].copy_(s.R[:, :n].t())
        G = None
        if d_gas_v_ar is not None:
            s.G[:, :n].copy_(d_gas_v_ar[:n, :3].t())
            G = s.G
        p = _params(N, float(nn), float(dt) / steps, s.capacity, seed)
        s.clock_set(n)
        before = C.c_uint64 * _lib.BZ_NCOUNTERS
        c0 = before()
        _lib.check(s.lib.bz_clock_get(s.ctx, None, None, None, c0))
        # in place in buffer `cur`; the other buffer stages the births
        _lib.check(s.lib.bz_push_mcc(s.ctx, C.byref(s.structs[cur]), C.byref(s.structs[1 - cur]),
                                     S.ptr(E), C.byref(p), None, S.ptr(s.R), S.ptr(G),
                                     S.ptr(s.moments)), "bz_push_mcc")
        c1 = before()
        nout = C.c_int64()
        _lib.check(s.lib.bz_clock_get(s.ctx, C.byref(nout), None, None, c1))
        births_total += int(c1[_lib.COUNTER_NAMES.index("births")] -
                            c0[_lib.COUNTER_NAMES.index("births")])
        n = nout.value  # rows in place, births after them (capped at the capacity)
        s.step += 1
    # rows [0, Nc) in place, births at [Nc, m), in the reference layout
    _rows_from_state(s, cur, n, d_data_ar, d_tosum_ar, kev)
    if d_curr_count is not None:
        S.as_gpu(d_curr_count, "d_curr_count")[0] = births_total
    s.cached[(d_data_ar.data_ptr(), Nc + births_total)] = (s.moments.clone(), n, cur)
    return None
