"""ctypes binding of libbz.so (the C ABI declared in include/bz.h).

The library is built in-tree by ``boltzmann_amd.build``.  Loading fails
loudly when it is missing: there is no CPU fallback for any stage.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libbz.so"

BZ_F32, BZ_F64 = 4, 8
BZ_XS_FIT, BZ_XS_TABLE, BZ_XS_FLIGHT, BZ_XS_DIRECT = 0, 1, 2, 3
BZ_RNG_PHILOX, BZ_RNG_STREAM = 0, 1
BZ_PUSH_REF, BZ_PUSH_LEAPFROG = 0, 1
BZ_REF_DEAD_IONIZER, BZ_REF_ENERGY_ROUNDTRIP = 1, 2
BZ_NCOUNTERS = 16
COUNTER_NAMES = ("births", "absorbed", "elastic", "excitation", "ionization", "null",
                 "overflow", "numax", "split", "merge", "resample", "particle_steps",
                 "copied", "sorted", "dep_range")


class BzError(RuntimeError):
    pass


class State(C.Structure):
    _fields_ = [("x", C.c_void_p), ("vx", C.c_void_p), ("vy", C.c_void_p),
                ("vz", C.c_void_p), ("w", C.c_void_p)]


class Params(C.Structure):
    _fields_ = [
        ("dt", C.c_double), ("xfac", C.c_double), ("afac", C.c_double), ("kev", C.c_double),
        ("jm", C.c_double), ("fel", C.c_double), ("L", C.c_double), ("dx", C.c_double),
        ("half_dx", C.c_double), ("L_dx", C.c_double), ("dep_M", C.c_double),
        ("nn", C.c_double), ("e_ion", C.c_double), ("e_exc", C.c_double),
        ("p_null", C.c_double), ("pc", C.c_double), ("nu_max", C.c_double),
        ("gas_sigma", C.c_double), ("e_field0", C.c_double),
        ("seed", C.c_uint64),
        ("dim", C.c_int32), ("xs_mode", C.c_int32), ("rng_mode", C.c_int32),
        ("n_cells", C.c_int32), ("n_nodes", C.c_int32), ("bnf", C.c_int32),
        ("r_stride", C.c_int64), ("push_order", C.c_int32), ("ref_flags", C.c_int32),
    ]


class XSec(C.Structure):
    _fields_ = [("energy", C.c_void_p), ("sig", C.c_void_p), ("slope", C.c_void_p),
                ("hint", C.c_void_p), ("n", C.c_int32), ("n_hint", C.c_int32),
                ("hint_base", C.c_int32), ("hint_bits", C.c_int32)]


class FieldParams(C.Structure):
    _fields_ = [("dx", C.c_double), ("dt", C.c_double), ("eps", C.c_double),
                ("v0", C.c_double), ("freq", C.c_double), ("mu_i", C.c_double),
                ("D_i", C.c_double), ("D_s", C.c_double), ("n", C.c_int32),
                ("fluid_on", C.c_int32), ("poisson_on", C.c_int32),
                ("skip_moments", C.c_int32), ("ion_scheme", C.c_int32),
                ("max_substeps", C.c_int32)]


class ResamplePolicy(C.Structure):
    _fields_ = [("split_below", C.c_double), ("merge_at", C.c_double),
                ("merge_keep", C.c_double), ("trigger_at", C.c_double),
                ("min_wt", C.c_double), ("split_to", C.c_double),
                ("ref_trigger", C.c_int32), ("max_ppc", C.c_int32),
                ("ref_rescale", C.c_int32), ("pad", C.c_int32),
                ("uniforms", C.c_void_p), ("u_stride", C.c_int64)]


class Mlp(C.Structure):
    _fields_ = [("w", C.c_void_p), ("in_shift", C.c_double * 3), ("in_scale", C.c_double * 3),
                ("out_shift", C.c_double * 3), ("out_scale", C.c_double * 3),
                ("in_lo", C.c_double * 3), ("in_hi", C.c_double * 3)]


class Adam(C.Structure):
    _fields_ = [("lr", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float),
                ("eps", C.c_float), ("bias1", C.c_float), ("bias2", C.c_float)]


BZ_MLP_NPARAM = 8771

# name -> (restype, argtypes); every symbol include/bz.h declares
_P = C.c_void_p
SIGNATURES = {
    "bz_last_error": (C.c_char_p, []),
    "bz_version": (C.c_int, []),
    "bz_ctx_create": (C.c_int, [C.POINTER(_P), C.c_int, C.c_int, C.c_int64, C.c_int32, C.c_int32]),
    "bz_ctx_destroy": (C.c_int, [_P]),
    "bz_set_stream": (C.c_int, [_P, _P]),
    "bz_set_push_shape": (C.c_int, [_P, C.c_int]),
    "bz_set_push_schedule": (C.c_int, [_P, C.c_int]),
    "bz_set_push_grid": (C.c_int, [_P, C.c_int]),
    "bz_set_push_queue": (C.c_int, [_P, C.c_int]),
    "bz_set_push_keys": (C.c_int, [_P, C.c_int32, C.POINTER(C.c_double)]),
    "bz_set_push_keys16": (C.c_int, [_P, C.c_int32]),
    "bz_synchronize": (C.c_int, [_P]),
    "bz_clock_set": (C.c_int, [_P, C.c_int64, C.c_uint64, C.c_double]),
    "bz_clock_get": (C.c_int, [_P, C.POINTER(C.c_int64), C.POINTER(C.c_uint64),
                               C.POINTER(C.c_double), C.POINTER(C.c_uint64)]),
    "bz_counters_reset": (C.c_int, [_P]),
    "bz_clock_advance": (C.c_int, [_P, C.c_double]),
    "bz_timing_enable": (C.c_int, [_P, C.c_int]),
    "bz_timing_get": (C.c_int, [_P, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    "bz_timing_get_n": (C.c_int, [_P, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_int64)]),
    "bz_push_mcc": (C.c_int, [_P, C.POINTER(State), C.POINTER(State), _P, C.POINTER(Params),
                              C.POINTER(XSec), _P, _P, _P]),
    "bz_deposit": (C.c_int, [_P, C.POINTER(State), C.POINTER(Params), _P]),
    "bz_compact": (C.c_int, [_P, C.POINTER(State), C.POINTER(State)]),
    "bz_compact_inplace": (C.c_int, [_P, C.POINTER(State)]),
    "bz_set_collision_counts": (C.c_int, [_P, C.c_int32]),
    "bz_bin_lists": (C.c_int, [_P, _P, C.c_int64, _P, C.c_int64, C.c_int32, C.c_int32, _P, _P,
                               _P, _P]),
    "bz_collision_counts": (C.c_int, [_P, _P, C.c_int32, C.c_int32]),
    "bz_reduction": (C.c_int, [_P, C.c_int32, C.c_int32, C.c_int32, _P, _P, _P, C.c_int64]),
    "bz_reduction_static": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, _P, _P, _P, C.c_int64,
                                      _P]),
    "bz_fill_uniforms": (C.c_int, [_P, _P, C.c_int64, C.c_int64, C.c_uint64, C.c_uint64]),
    "bz_field_step": (C.c_int, [_P, C.POINTER(FieldParams), _P, _P, _P, _P, _P, _P, _P, _P,
                                _P, C.c_int32]),
    "bz_bin_count": (C.c_int, [_P, C.POINTER(State), C.POINTER(Params), C.POINTER(C.c_double),
                               C.c_int32, _P]),
    "bz_bin_sort": (C.c_int, [_P, C.POINTER(State), C.POINTER(Params), C.POINTER(C.c_double),
                              C.c_int32, _P, _P, _P]),
    "bz_resample": (C.c_int, [_P, C.POINTER(State), C.POINTER(State), C.POINTER(Params),
                              _P, _P, _P, _P, C.c_int32, C.POINTER(ResamplePolicy)]),
    "bz_resample_inplace": (C.c_int, [_P, C.POINTER(State), C.POINTER(Params),
                                      C.POINTER(C.c_double), C.c_int32, _P, C.c_int32,
                                      C.POINTER(ResamplePolicy)]),
    "bz_philox4x32_10": (None, [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                C.POINTER(C.c_uint32)]),
    "bz_swarm_step": (C.c_int, [_P, C.POINTER(State), C.POINTER(State), C.POINTER(Params),
                                C.POINTER(XSec), C.c_int32, _P, _P, _P, _P, C.c_int32]),
    "bz_swarm_control": (C.c_int, [_P, C.POINTER(State), C.POINTER(State), C.POINTER(Params),
                                   C.c_int32, _P, _P, _P, _P, C.POINTER(ResamplePolicy)]),
    "bz_swarm_eedf": (C.c_int, [_P, C.POINTER(State), C.c_int32, _P, C.c_int32, C.c_double,
                                C.c_double, C.c_double, _P]),
    "bz_mlp_forward": (C.c_int, [_P, C.POINTER(Mlp), _P, _P, C.c_int64]),
    "bz_mlp_workspace_floats": (C.c_int64, [C.c_int64]),
    "bz_mlp_grad": (C.c_int, [_P, C.POINTER(Mlp), _P, _P, C.c_int64, _P, _P, _P]),
    "bz_mlp_adam": (C.c_int, [_P, C.POINTER(Adam), _P, _P, _P, _P, _P]),
    "bz_mlp_round": (C.c_int, [_P, _P, _P]),
    "bz_closure_source": (C.c_int, [_P, C.POINTER(Mlp), C.c_int32, _P, _P, C.c_double,
                                    C.c_double, _P, _P]),
    "bz_field_substep": (C.c_int, [_P, C.POINTER(FieldParams), _P, _P, _P, _P, _P, _P, _P, _P,
                                   _P, _P]),
    "bz_closure_commit": (C.c_int, [_P, C.c_int32, _P, _P, _P, _P, _P, _P]),
    "bz_graph_begin": (C.c_int, [_P]),
    "bz_graph_end": (C.c_int, [_P, C.POINTER(C.c_int32)]),
    "bz_graph_launch": (C.c_int, [_P, C.c_int32]),
    "bz_graph_destroy": (C.c_int, [_P, C.c_int32]),
}

_lib = None


def load() -> C.CDLL:
    """Load libbz.so (built in-tree).  Raises if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise BzError(f"{LIB_PATH} is missing: build it with `python -m boltzmann_amd.build` "
                      "(there is no CPU fallback)")
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = load().bz_last_error().decode(errors="replace")
        raise BzError(f"{what}: {msg}" if what else msg)


def philox(ctr, key):
    """Host Philox4x32-10 block from the library (used by tests)."""
    lib = load()
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib.bz_philox4x32_10(c, k, o)
    return tuple(o)
