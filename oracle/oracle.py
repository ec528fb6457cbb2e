"""numpy front end of the CPU oracle (test infrastructure only).

The constants the oracle needs are derived here from the physical inputs,
written out from the reference's formulas (electron_kernel_numba.py:14-17,
:63, :206, :226, :258, :264; post_processing_1D.py:44; boltzmann.py:66-71),
not taken from the product's set-up code, so a mistake there shows up as a
parity failure.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import build as _build

_lib = None


class OrcParams(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "dt", "xfac", "afac", "kev", "jm", "fel", "L", "dx", "half_dx", "L_dx", "dep_M", "nn",
        "e_ion", "e_exc", "pc", "e_field0", "nu_max")] + [
        (n, C.c_int32) for n in ("dim", "xs_mode", "n_cells", "n_nodes", "bnf", "nthreads")] + [
        ("seed", C.c_uint64), ("step", C.c_uint64), ("push_order", C.c_int32),
        ("births_last", C.c_int32), ("ref_flags", C.c_int32), ("pad", C.c_int32)]


class OrcTable(C.Structure):
    _fields_ = [("energy", C.c_void_p), ("sig", C.c_void_p), ("slope", C.c_void_p),
                ("n", C.c_int32), ("pad", C.c_int32)]


def load():
    global _lib
    if _lib is None:
        path = _build.ensure_built()
        _lib = C.CDLL(str(path))
        _lib.orc_philox.restype = None
        _lib.orc_resample_uniform.restype = C.c_double
        _lib.orc_resample_uniform.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32]
        _lib.orc_collision_uniforms.restype = None
        _lib.orc_collision_uniforms.argtypes = [C.c_uint64, C.c_uint64, C.c_int64, C.c_int,
                                                C.c_void_p, C.c_int64]
        for nm in ("orc_push_step_f64", "orc_push_step_f32"):
            getattr(_lib, nm).restype = C.c_int64
        _lib.orc_field_step.restype = None
    return _lib


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def philox(ctr, key):
    lib = load()
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib.orc_philox(c, k, o)
    return tuple(o)


# Physical constants as the reference writes them.
Q_E = 1.602 * 10 ** (-19)
M_E = 9.10938 * 10 ** (-31)
M_N = 6.6 * 10 ** (-26)
J_EV = 6.241509 * (10 ** 18)


@dataclass
class Setup:
    """Physical inputs of one oracle run."""
    dim: int
    dt: float
    xs_mode: int = 1           # 0 reference fits, 1 table, 2 table with free flights
    n_cells: int = 1
    length_cm: float = 2.54
    bnf: int = 4
    gas_density_m3: float = 3.22e21
    mass_ratio: float = 1.36e-5
    e_ion: float = 15.76
    e_exc: float = 11.55
    nu_max: float = 0.0
    e_field0: float = 0.0
    nthreads: int = 1
    seed: int = 12345          # free-flight mode draws its own Philox stream
    step: int = 0
    push_order: int = 0        # 0 reference (x then v), 1 leapfrog (v then x)
    # secondaries after all survivors (the 1-D layout) or right after their
    # parents (the 0-D swarm layout); None: by dimension
    births_last: Optional[bool] = None
    # reference behaviours: 1 the dead ionizer's flag at node 0, 2 the
    # (e*w)/w binning energy
    ref_flags: int = 0

    def params(self) -> OrcParams:
        p = OrcParams()
        p.dt = self.dt
        p.xfac = 100 * self.dt
        p.afac = self.dt * (Q_E / M_E)
        p.kev = 0.5 * M_E * J_EV
        p.jm = J_EV * M_E
        p.fel = math.sqrt(1 - (2 * M_E / M_N)) if self.xs_mode == 0 else \
            math.sqrt(1 - 2 * self.mass_ratio)
        n = self.n_cells if self.dim == 1 else 1
        p.L = self.length_cm
        p.dx = self.length_cm / n
        p.half_dx = 0.5 * p.dx
        p.L_dx = self.length_cm + p.dx
        p.dep_M = float(self.bnf * (n + 1))
        p.nn = self.gas_density_m3
        p.e_ion = self.e_ion
        p.e_exc = self.e_exc
        if self.nu_max > 0:
            p_null = 1.0 - math.exp(-self.nu_max * self.dt)
            p.pc = p_null / self.nu_max
        p.e_field0 = self.e_field0
        p.dim = self.dim
        p.xs_mode = self.xs_mode
        p.n_cells = n
        p.n_nodes = n + 1 if self.dim == 1 else 1
        p.bnf = self.bnf
        p.nthreads = self.nthreads
        p.nu_max = self.nu_max
        p.seed = self.seed
        p.step = self.step
        p.push_order = self.push_order
        bl = self.births_last if self.births_last is not None else self.dim == 1
        p.births_last = 1 if bl else 0
        p.ref_flags = self.ref_flags
        return p


def collision_uniforms(seed: int, step: int, n: int, dtype) -> np.ndarray:
    """[7, n] uniforms of the Philox stream (columns 0,1,2,3,5 used)."""
    lib = load()
    R = np.zeros((7, max(n, 1)), dtype=dtype)
    lib.orc_collision_uniforms(C.c_uint64(seed), C.c_uint64(step), n,
                               8 if dtype == np.float64 else 4, _ptr(R), R.shape[1])
    return R


def gas_velocities(seed: int, step: int, n: int, dtype, sigma: float) -> np.ndarray:
    """[3, n] gas velocities of the Philox stream (purpose 1): Box-Muller with
    std `sigma` per component, as the device draws them for a warm gas."""
    lib = load()
    G = np.zeros((3, max(n, 1)), dtype=dtype)
    lib.orc_gas_velocities(C.c_uint64(seed), C.c_uint64(step), C.c_int64(n),
                           C.c_int(8 if dtype == np.float64 else 4), C.c_double(sigma), _ptr(G),
                           C.c_int64(G.shape[1]))
    return G


def fit_sigma(setup: "Setup", e: np.ndarray):
    """(sigma_el * nn, sigma_ion * nn) of the reference argon fits at energies e."""
    lib = load()
    e = np.ascontiguousarray(e)
    p = setup.params()
    a, b = np.zeros_like(e), np.zeros_like(e)
    fn = lib.orc_fit_sigma_f64 if e.dtype == np.float64 else lib.orc_fit_sigma_f32
    fn(C.byref(p), C.c_int64(e.size), _ptr(e), _ptr(a), _ptr(b))
    return a, b


def fill_holes(state):
    """Compaction in place by hole filling (the order this build fixes for
    the reference's atomic compaction, kernels/post_processing_1D.py:26-38):
    L = the number of slots with w > 0; the k-th slot below L with w <= 0 (in
    slot order) receives the k-th slot at or above L with w > 0 (in slot
    order).  Returns the first L slots of every array."""
    arrs = [np.array(a, copy=True) for a in state]
    w = arrs[4]
    L = int(np.count_nonzero(w > 0))
    holes = np.flatnonzero(~(w[:L] > 0))
    movers = L + np.flatnonzero(w[L:] > 0)
    assert holes.size == movers.size
    for a in arrs:
        a[holes] = a[movers]
    return tuple(a[:L] for a in arrs)


def expm1_exp(v: np.ndarray):
    """The oracle's (e^v - 1, e^v), the exponential of the collision
    probabilities, in v's precision."""
    lib = load()
    v = np.ascontiguousarray(v)
    a, b = np.zeros_like(v), np.zeros_like(v)
    fn = lib.orc_expm1_f64 if v.dtype == np.float64 else lib.orc_expm1_f32
    fn(C.c_int64(v.size), _ptr(v), _ptr(a), _ptr(b))
    return a, b


def resample_uniform(seed: int, step: int, key: int, m: int) -> float:
    return load().orc_resample_uniform(seed, step, key, m)


def push_step(setup: Setup, state, E, R, G=None, table=None, capacity=None,
              collisions: bool = False):
    """One collision + push + wall + deposit + compaction step.

    state: (x, vx, vy, vz, w) arrays of one dtype; R: [7, stride] uniforms
    (None in free-flight mode, which draws Philox(seed, step) itself);
    G: optional [3, stride] gas velocities; table: (energy, sig[3,n], slope[3,n]).
    Returns (new_state, moments[3, n_nodes], counters[8]), and with
    collisions=True also the collision counters [3, n_nodes] (sum w of the
    elastic and of the excitation collisions, number of ionizations)."""
    lib = load()
    x, vx, vy, vz, w = [np.ascontiguousarray(a) for a in state]
    dt = x.dtype
    n = x.size
    cap = capacity if capacity is not None else 2 * n + 16
    out = [np.zeros(cap, dtype=dt) for _ in range(5)]
    p = setup.params()
    nn = p.n_nodes
    mom = np.zeros(3 * nn)
    cnt = np.zeros(8, dtype=np.uint64)
    tab = None
    keep = []
    if table is not None:
        e, s, sl = [np.ascontiguousarray(a, dtype=dt) for a in table]
        keep += [e, s, sl]
        tab = OrcTable(energy=_ptr(e), sig=_ptr(s), slope=_ptr(sl), n=e.size)
    Ea = None if E is None else np.ascontiguousarray(E, dtype=dt)
    # R None: the oracle draws the Philox uniforms of (seed, step, slot) itself,
    # as free-flight mode (xs_mode 2) always does
    Ra = None if R is None else np.ascontiguousarray(R, dtype=dt)
    Ga = None if G is None else np.ascontiguousarray(G, dtype=dt)
    rs = Ra.shape[1] if Ra is not None else (Ga.shape[1] if Ga is not None else 1)
    if Ga is not None and Ga.shape[1] != rs:
        raise ValueError("R and G must share the column stride")
    coll = np.zeros(3 * nn) if collisions else None
    fn = lib.orc_push_step_f64 if dt == np.float64 else lib.orc_push_step_f32
    m = fn(C.byref(p), C.c_int64(n), *[_ptr(a) for a in (x, vx, vy, vz, w)], _ptr(Ea), _ptr(Ra),
           C.c_int64(rs), _ptr(Ga), C.byref(tab) if tab is not None else None,
           *[_ptr(a) for a in out], C.c_int64(cap), _ptr(mom), _ptr(cnt), _ptr(coll))
    if collisions:
        return tuple(a[:m] for a in out), mom.reshape(3, nn), cnt, coll.reshape(3, nn)
    return tuple(a[:m] for a in out), mom.reshape(3, nn), cnt


def deposit(setup: Setup, state):
    lib = load()
    x, vx, vy, vz, w = [np.ascontiguousarray(a) for a in state]
    p = setup.params()
    mom = np.zeros(3 * p.n_nodes)
    fn = lib.orc_deposit_f64 if x.dtype == np.float64 else lib.orc_deposit_f32
    fn(C.byref(p), C.c_int64(x.size), *[_ptr(a) for a in (x, vx, vy, vz, w)], _ptr(mom))
    return mom.reshape(3, p.n_nodes)


def field_step(N, dx, dt, eps, v0, freq, mu_i, D_i, D_s, fluid_on, poisson_on, t, mom, ni, ns,
               Te, g2, V, E, chol, ion_scheme=0, max_substeps=0):
    """FP64 grid step on copies of the arrays; returns (ni, ns, Te, g2, V, E)."""
    lib = load()
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (mom, ni, ns, Te, g2, V, E, chol)]
    lib.orc_field_step(C.c_int(N), *[C.c_double(v) for v in (dx, dt, eps, v0, freq, mu_i, D_i,
                                                               D_s)],
                       C.c_int(fluid_on), C.c_int(poisson_on), C.c_int(ion_scheme),
                       C.c_int(max_substeps), C.c_double(t),
                       *[_ptr(a) for a in arrs])
    return arrs[1:7]


class OrcPolicy(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("split_below", "merge_at", "merge_keep", "trigger_at",
                                          "min_wt", "split_to")] + [
        (n, C.c_int32) for n in ("ref_trigger", "max_ppc", "ref_rescale", "pad")]


def policy(split_below=0.5, merge_at=3.0, merge_keep=0.6, trigger_at=3.0, min_wt=0.1,
           split_to=2.0, ref_trigger=0, max_ppc=0, ref_rescale=0) -> OrcPolicy:
    return OrcPolicy(split_below, merge_at, merge_keep, trigger_at, min_wt, split_to,
                     ref_trigger, max_ppc, ref_rescale, 0)


def bin_sort(setup: Setup, nb: int, thresholds, state):
    """Keys, counts[n_keys+1], offsets[n_keys+2] and the stable permutation."""
    lib = load()
    x, vx, vy, vz, w = [np.ascontiguousarray(a) for a in state]
    p = setup.params()
    n_keys = p.n_nodes * nb
    thr = np.ascontiguousarray(np.asarray(thresholds, dtype=np.float64).reshape(-1))
    thr = np.concatenate([thr, np.zeros(1)])
    counts = np.zeros(n_keys + 1, dtype=np.int32)
    offsets = np.zeros(n_keys + 2, dtype=np.int32)
    perm = np.zeros(max(x.size, 1), dtype=np.uint32)
    fn = lib.orc_bin_sort_f64 if x.dtype == np.float64 else lib.orc_bin_sort_f32
    fn(C.byref(p), C.c_int(nb), _ptr(thr), C.c_int64(x.size),
       *[_ptr(a) for a in (x, vx, vy, vz, w)], _ptr(counts), _ptr(offsets), _ptr(perm))
    return counts, offsets, perm[:x.size]


def resample(setup: Setup, pol: OrcPolicy, nb: int, thresholds, targets, seed: int, step: int,
             state, capacity=None, uniforms=None):
    """Binning + sort + ES-WPNC split/merge; output in key order.
    uniforms: optional [n_keys, stride] caller stream (row k: segment k's
    draws).  Returns (new_state, counters[16])."""
    lib = load()
    x, vx, vy, vz, w = [np.ascontiguousarray(a) for a in state]
    dt = x.dtype
    p = setup.params()
    thr = np.concatenate([np.asarray(thresholds, dtype=np.float64).reshape(-1), np.zeros(1)])
    tg = np.ascontiguousarray(np.asarray(targets, dtype=np.int32))
    cap = capacity if capacity is not None else 4 * x.size + 1024
    out = [np.zeros(cap, dtype=dt) for _ in range(5)]
    cnt = np.zeros(16, dtype=np.uint64)
    U, us = _uniforms(uniforms)
    fn = lib.orc_resample_f64 if dt == np.float64 else lib.orc_resample_f32
    fn.restype = C.c_int64
    m = fn(C.byref(p), C.byref(pol), C.c_int(nb), _ptr(thr), _ptr(tg), C.c_uint64(seed),
           C.c_uint64(step), C.c_int64(x.size), *[_ptr(a) for a in (x, vx, vy, vz, w)],
           *[_ptr(a) for a in out], C.c_int64(cap), _ptr(cnt), _ptr(U), C.c_int64(us))
    return tuple(a[:m] for a in out), cnt


def _uniforms(uniforms):
    if uniforms is None:
        return None, 0
    U = np.ascontiguousarray(uniforms, dtype=np.float64)
    if U.ndim != 2:
        raise ValueError("uniforms: [n_keys, stride]")
    return U, U.shape[1]


def resample_inplace(setup: Setup, pol: OrcPolicy, nb: int, thresholds, targets, seed: int,
                     step: int, state, capacity=None, uniforms=None):
    """The reference's in-place layout: split particles keep their slot with
    the divided weight, merged members become holes (w = 0), clones and
    merged particles are appended after the slots in key order.  Returns
    (slots [0, n'), holes included, counters[16])."""
    lib = load()
    dt = np.asarray(state[0]).dtype
    n = int(np.asarray(state[0]).size)
    cap = capacity if capacity is not None else 4 * n + 1024
    arrs = []
    for a in state:
        b = np.zeros(cap, dtype=dt)
        b[:n] = a
        arrs.append(b)
    p = setup.params()
    thr = np.concatenate([np.asarray(thresholds, dtype=np.float64).reshape(-1), np.zeros(1)])
    tg = np.ascontiguousarray(np.asarray(targets, dtype=np.int32))
    cnt = np.zeros(16, dtype=np.uint64)
    U, us = _uniforms(uniforms)
    fn = lib.orc_resample_inplace_f64 if dt == np.float64 else lib.orc_resample_inplace_f32
    fn.restype = C.c_int64
    m = fn(C.byref(p), C.byref(pol), C.c_int(nb), _ptr(thr), _ptr(tg), C.c_uint64(seed),
           C.c_uint64(step), C.c_int64(n), *[_ptr(a) for a in arrs], C.c_int64(cap), _ptr(cnt),
           _ptr(U), C.c_int64(us))
    return tuple(a[:m] for a in arrs), cnt
