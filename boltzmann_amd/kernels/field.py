"""EF_kernel and heavies_kernel_fluid with the reference's signatures
(kernels/EF_kernel_1D.py:8-44, kernels/heavies_kernel_1D.py:5-64).

Both take and return host numpy arrays like the reference's numba versions;
the solve itself runs in the FP64 field kernel of libbz.so (bz_field_step
with the Poisson or the fluid part switched on alone).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from .. import _lib
from . import _session as S


def _run(N, dx, t, dt, fp_over, ne, ni, ns, E, chol):
    s = S.session(0, _lib.BZ_F64, 1, N + 1)
    dev = s.dev
    f64 = dict(dtype=torch.float64, device=dev)
    mom = torch.zeros(3 * (N + 1), **f64)
    if ne is not None:
        mom[:N + 1] = torch.as_tensor(np.asarray(ne[:N + 1]), **f64)
    d_ni = torch.as_tensor(np.asarray(ni, dtype=np.float64), **f64).clone()
    d_ns = torch.as_tensor(np.asarray(ns, dtype=np.float64), **f64).clone() if ns is not None \
        else torch.zeros(N + 1, **f64)
    d_E = torch.as_tensor(np.asarray(E, dtype=np.float64), **f64).clone()
    d_Et = torch.zeros(N, **f64)
    Te = torch.zeros(N + 1, **f64)
    g2 = torch.zeros(N + 1, **f64)
    V = torch.zeros(max(N - 1, 1), **f64)
    ch = torch.as_tensor(np.ascontiguousarray(chol, dtype=np.float64).ravel(), **f64) \
        if chol is not None else torch.zeros(2 * max(N - 1, 1), **f64)
    fp = _lib.FieldParams()
    fp.dx = dx
    fp.dt = dt
    fp.n = N
    fp.skip_moments = 1
    for k, v in fp_over.items():
        setattr(fp, k, v)
    s.clock_set(0, t)
    _lib.check(s.lib.bz_field_step(s.ctx, C.byref(fp), S.ptr(mom), S.ptr(d_ni), S.ptr(d_ns),
                                   S.ptr(Te), S.ptr(g2), S.ptr(V), S.ptr(d_E), S.ptr(d_Et),
                                   S.ptr(ch), 0), "bz_field_step")
    return d_ni.cpu().numpy(), d_ns.cpu().numpy(), V.cpu().numpy(), d_E.cpu().numpy()


def EF_kernel(E_ar, ne_ar, ni_ar, curr_t, V_ar, V_rhs, Vc_diag, Vc_lower_diag, V_tempy, N, dx,
              epsilon, V0, freq):
    N = int(N)
    chol = np.stack([np.asarray(Vc_diag[:N - 1]), np.asarray(Vc_lower_diag[:N - 1])])
    _, _, V, E = _run(N, float(dx), float(curr_t), 0.0,
                      dict(eps=float(epsilon), v0=float(V0), freq=float(freq), poisson_on=1,
                           fluid_on=0),
                      ne_ar, ni_ar, None, E_ar, chol)
    V_ar[:N - 1] = V[:N - 1]
    E_ar[:N] = E
    return (V_ar, E_ar)


def heavies_kernel_fluid(break_flag, E_ar, ni_ar, ni_rhs, Ji_ar, ns_ar, ns_rhs, Js_ar, dx, dt,
                         mu_i, D_i, D_s, N):
    N = int(N)
    old = np.array(ni_ar, dtype=np.float64, copy=True)
    ni, ns, _, _ = _run(N, float(dx), 0.0, float(dt),
                        dict(mu_i=float(mu_i), D_i=float(D_i), D_s=float(D_s), fluid_on=1,
                             poisson_on=0),
                        None, ni_ar, ns_ar, E_ar, None)
    # heavies_kernel_1D.py:43-46: an interior density that turns negative
    if np.any((ni[1:N] < 0) & (old[1:N] > 0)):
        break_flag = 1
    ni_ar[:] = ni
    ns_ar[:] = ns
    return (ni_ar, ns_ar, break_flag)
