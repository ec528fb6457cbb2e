"""The C-ABI library loads and exports every entry point include/bz.h declares
(no device calls: this runs without a GPU)."""
import ctypes
import re
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]


def declared_functions():
    text = (ROOT / "include" / "bz.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bz_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("bz_ctx_create", "bz_push_mcc", "bz_deposit", "bz_field_step", "bz_bin_sort",
                 "bz_resample", "bz_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from boltzmann_amd import _lib
    lib = _lib.load()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(declared_functions()) <= set(_lib.SIGNATURES)


def test_host_philox_matches_oracle():
    from boltzmann_amd import _lib
    from oracle import oracle as O
    rng = np.random.default_rng(0)
    for _ in range(50):
        c = tuple(int(v) for v in rng.integers(0, 2 ** 32, 4))
        k = tuple(int(v) for v in rng.integers(0, 2 ** 32, 2))
        assert _lib.philox(c, k) == O.philox(c, k)


def test_error_path_without_device():
    """A null context is rejected with a message, not a crash."""
    from boltzmann_amd import _lib
    lib = _lib.load()
    rc = lib.bz_push_mcc(None, None, None, None, None, None, None, None, None)
    assert rc != 0
    assert b"null context" in lib.bz_last_error()


def test_struct_sizes_match_header_layout():
    from boltzmann_amd import _lib
    # 19 doubles + u64 + 6 int32 + int64 + 2 int32
    assert ctypes.sizeof(_lib.Params) == 19 * 8 + 8 + 6 * 4 + 8 + 2 * 4
    assert ctypes.sizeof(_lib.XSec) == 4 * 8 + 4 * 4
    assert ctypes.sizeof(_lib.FieldParams) == 8 * 8 + 6 * 4
    assert ctypes.sizeof(_lib.Mlp) == 8 + 18 * 8
    assert ctypes.sizeof(_lib.Adam) == 6 * 4
    assert ctypes.sizeof(_lib.ResamplePolicy) == 6 * 8 + 4 * 4
